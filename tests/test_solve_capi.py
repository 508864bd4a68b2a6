"""hsg_solve, the C entry of the whole solve (include/hsg_solve.h), replacing the reference's
_hsolve.solve(a, b, config_json, verbose, storage) (python/src/module.cpp:326-345, 458-461).

Without a GPU: the library exports what the header declares; hsg_parse_config completes a config
with the defaults and rejects invalid ones; input errors come back as the header's codes with the
message (checked through ctypes, in this process's interpreter); a C
program linked against the library starts its own interpreter, finds the package and reports the
missing GPU as an internal error. With a GPU: the same solve through ctypes gives the eigenpairs
and report of hsolve_amd.api.solve; the C program's eigenvalues of the 1D Laplacian pencil equal
the closed form; a complex Hermitian pencil (a Bloch cell) comes back with complex vectors and the
eigenpairs of hsolve_amd.api.solve.
"""
import ctypes as C
import json
import re
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "hsolve_amd" / "libhsg_solve.so"


class CSR(C.Structure):
    _fields_ = [("n", C.c_int64), ("indptr", C.c_void_p), ("indices", C.c_void_p), ("values", C.c_void_p),
                ("is_complex", C.c_int32), ("reserved", C.c_int32)]


def _lib():
    if not LIB.exists():
        pytest.skip("libhsg_solve.so not built")
    lib = C.CDLL(str(LIB))
    lib.hsg_solve.argtypes = [C.POINTER(CSR), C.POINTER(CSR), C.c_char_p, C.c_int, C.c_char_p,
                              C.POINTER(C.c_void_p)]
    lib.hsg_solve.restype = C.c_int
    for name in ("hsg_result_order", "hsg_result_count"):
        getattr(lib, name).argtypes, getattr(lib, name).restype = [C.c_void_p], C.c_int64
    lib.hsg_result_is_complex.argtypes, lib.hsg_result_is_complex.restype = [C.c_void_p], C.c_int32
    for name in ("hsg_result_eigenvalues", "hsg_result_eigenvectors"):
        getattr(lib, name).argtypes, getattr(lib, name).restype = [C.c_void_p], C.POINTER(C.c_double)
    for name in ("hsg_result_report_json", "hsg_result_config_json"):
        getattr(lib, name).argtypes, getattr(lib, name).restype = [C.c_void_p], C.c_char_p
    lib.hsg_result_free.argtypes = [C.c_void_p]
    lib.hsg_parse_config.argtypes, lib.hsg_parse_config.restype = [C.c_char_p], C.c_char_p
    lib.hsg_solve_last_error_message.restype = C.c_char_p
    return lib


def _csr(M, keep):
    """A CSR struct over M's arrays (kept alive in `keep`)."""
    M = sp.csr_matrix(M)
    ip = np.ascontiguousarray(M.indptr, dtype=np.int64)
    ix = np.ascontiguousarray(M.indices, dtype=np.int32)
    cplx = np.iscomplexobj(M.data)
    v = np.ascontiguousarray(M.data, dtype=np.complex128 if cplx else np.float64)
    keep += [ip, ix, v]
    return CSR(M.shape[0], ip.ctypes.data, ix.ctypes.data, v.ctypes.data, int(cplx), 0)


def _call(lib, A, B, config, storage=b"auto"):
    keep = []
    a = _csr(A, keep)
    b = _csr(B, keep) if B is not None else None
    out = C.c_void_p()
    cfg = json.dumps(config).encode() if isinstance(config, dict) else config
    rc = lib.hsg_solve(C.byref(a), C.byref(b) if b is not None else None, cfg, 0, storage, C.byref(out))
    return rc, out


def _take(lib, out):
    n, m = lib.hsg_result_order(out), lib.hsg_result_count(out)
    cplx = lib.hsg_result_is_complex(out)
    w = np.ctypeslib.as_array(lib.hsg_result_eigenvalues(out), shape=(m,)).copy() if m else np.zeros(0)
    size = n * m * (2 if cplx else 1)
    v = np.ctypeslib.as_array(lib.hsg_result_eigenvectors(out), shape=(size,)).copy() if size else np.zeros(0)
    v = (v.view(np.complex128) if cplx else v).reshape((n, m), order="F")
    report = json.loads(lib.hsg_result_report_json(out))
    config = json.loads(lib.hsg_result_config_json(out))
    lib.hsg_result_free(out)
    return w, v, report, config


def _laplacian(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()


def _declared():
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "hsg_solve.h").read_text(), flags=re.S)
    return set(re.findall(r"\b(hsg_\w+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = _lib()
    names = _declared()
    assert {"hsg_solve", "hsg_result_free", "hsg_solve_last_error_message"} <= names
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/hsg_solve.h but not exported"


@pytest.mark.parametrize("case", ["storage", "json", "json_list", "indptr", "indices", "order"])
def test_input_errors_come_back_as_codes(case):
    lib = _lib()
    A = _laplacian(10)
    B, cfg, storage = None, {"target": {"mode": "nearest", "sigma": 1.0, "k": 2}}, b"auto"
    if case == "storage":
        storage = b"sideways"
    elif case == "json":
        cfg = b"{not json"
    elif case == "json_list":
        cfg = b"[1, 2]"
    elif case == "order":
        B = sp.identity(11, format="csr")
    keep = []
    a = _csr(A, keep)
    if case == "indptr":
        bad = keep[0].copy()
        bad[3] = 0
        keep.append(bad)
        a.indptr = bad.ctypes.data
    if case == "indices":
        bad = keep[1].copy()
        bad[4] = 10
        keep.append(bad)
        a.indices = bad.ctypes.data
    b = _csr(B, keep) if B is not None else None
    out = C.c_void_p(1234)
    cfgb = json.dumps(cfg).encode() if isinstance(cfg, dict) else cfg
    rc = lib.hsg_solve(C.byref(a), C.byref(b) if b is not None else None, cfgb, 0, storage, C.byref(out))
    assert rc == 1, lib.hsg_solve_last_error_message()  # HSG_SOLVE_INVALID_INPUT
    assert not out.value
    assert b"InvalidInputError" in lib.hsg_solve_last_error_message()


def test_parse_config_completes_and_validates():
    from hsolve_amd import config as cfgmod
    lib = _lib()
    doc = json.loads(lib.hsg_parse_config(b'{"target": {"mode": "nearest", "sigma": 3.0, "k": 5}}'))
    want = cfgmod.default_config()
    want["target"].update({"mode": "nearest", "sigma": 3.0, "k": 5})
    assert doc == json.loads(json.dumps(want))
    assert json.loads(lib.hsg_parse_config(None)) == json.loads(json.dumps(cfgmod.default_config()))
    assert lib.hsg_parse_config(b'{"bogus_key": 1}') is None
    assert b"unknown key" in lib.hsg_solve_last_error_message()
    assert lib.hsg_parse_config(b'{"tol": -1}') is None


def _build_demo(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None or not LIB.exists():
        pytest.skip("gcc or libhsg_solve.so missing")
    import sysconfig
    exe = tmp_path / "solve_demo"
    libdir = sysconfig.get_config_var("LIBDIR")
    pyver = f"python{sys.version_info.major}.{sys.version_info.minor}"
    subprocess.run([gcc, "-O2", "-std=c11", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "capi" / "solve_demo.c"),
                    "-o", str(exe), f"-L{LIB.parent}", "-lhsg_solve", f"-L{libdir}", f"-l{pyver}", "-lm",
                    f"-Wl,-rpath,{LIB.parent}"], check=True, capture_output=True)
    return exe


def test_c_host_reports_the_missing_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the GPU test covers the C host")
    exe = _build_demo(tmp_path)
    res = subprocess.run([str(exe), "200", "1.0", "4"], capture_output=True, text=True, timeout=600, cwd=tmp_path)
    assert res.returncode == 2, res.stdout + res.stderr
    assert "rc 9" in res.stdout and "needs a ROCm GPU" in res.stdout


@pytest.mark.gpu
def test_ctypes_solve_matches_the_python_api(gpu):
    from hsolve_amd import api
    from hsolve_amd.gen import yee
    lib = _lib()
    A, _, g = yee.cube_cavity(6)
    cfg = {"target": {"mode": "nearest", "sigma": 6.5, "k": 8}}
    rc, out = _call(lib, A, None, cfg)
    assert rc == 0, lib.hsg_solve_last_error_message()
    w, v, report, config = _take(lib, out)
    ref = api.solve(A, config=cfg)
    print("capi", w, "api", ref.eigenvalues, flush=True)
    assert report["success"] and ref.success
    assert np.allclose(w, ref.eigenvalues, rtol=1e-12, atol=0)
    assert v.shape == ref.eigenvectors.shape
    assert config["target"]["sigma"] == 6.5 and config["target"]["k"] == 8
    R = A @ v - v * w
    assert np.abs(R).max() <= 1e-8 * np.abs(A).max() * np.abs(v).max()


@pytest.mark.gpu
def test_c_host_laplacian_closed_form(gpu, tmp_path):
    exe = _build_demo(tmp_path)
    n, sigma, k = 2000, 1.0, 4
    res = subprocess.run([str(exe), str(n), str(sigma), str(k)], capture_output=True, text=True, timeout=600,
                         cwd=tmp_path)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    lam = np.array([float(x.split()[1]) for x in res.stdout.splitlines() if x.startswith("lambda ")])
    exact = 1.0 - np.cos(np.arange(1, n + 1) * np.pi / (n + 1))  # A x = lambda (2 I) x
    want = np.sort(exact[np.argsort(np.abs(exact - sigma))[:k]])
    assert lam.size == k
    assert np.allclose(np.sort(lam), want, rtol=1e-10, atol=1e-12)
    # |A x - lambda B x|_inf / |x|_inf against (|A|_inf + |lambda| |B|_inf) = 6: the solve meets
    # tol = 1e-10 in 2-norms, which the infinity-norm ratio may exceed by up to sqrt(n)
    resid = float([x for x in res.stdout.splitlines() if x.startswith("residual ")][0].split()[1])
    assert resid < 1e-8 * 6.0


@pytest.mark.gpu
def test_complex_hermitian_input(gpu):
    # a Bloch-periodic Yee cell with generic phases (genuinely complex A) and its B: the call
    # returns complex vectors and the eigenpairs of hsolve_amd.api.solve on the same pencil
    from hsolve_amd import api
    from hsolve_amd.gen import yee
    lib = _lib()
    A, B, _ = yee.bloch_cell(6, (0.7, 1.3, 2.1))
    assert np.iscomplexobj(A.data) and np.any(A.data.imag != 0)
    cfg = {"target": {"mode": "nearest", "sigma": 25.0, "k": 6}}
    rc, out = _call(lib, A, B, cfg)
    assert rc == 0, lib.hsg_solve_last_error_message()
    w, v, report, _ = _take(lib, out)
    ref = api.solve(A, B, config=cfg)
    print("capi", w, "api", ref.eigenvalues, flush=True)
    assert report["success"] and ref.success
    assert np.iscomplexobj(v) and v.shape == ref.eigenvectors.shape == (A.shape[0], 6)
    assert np.allclose(w, ref.eigenvalues, rtol=1e-12, atol=0)
    R = A @ v - (B @ v) * w
    assert np.abs(R).max() <= 1e-8 * (np.abs(A).max() + abs(w).max() * np.abs(B).max()) * np.abs(v).max()
