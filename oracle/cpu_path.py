"""CPU restatement of the shift-invert path, timed as the CPU baseline (oracle: test and bench
infrastructure only; the product never imports it).

The parts, each the CPU counterpart of the reference's path (MUMPS + OpenBLAS + its block Krylov
driver, SURVEY's CPU-baseline plan):
  * ordering and supernodal symbolic analysis: the engine's host analysis (nested dissection,
    elimination tree, relaxed supernodes; hsg_analyze runs on the host only, no GPU involved),
    timed as part of the CPU path;
  * numeric multifrontal LDL^T with Bunch-Kaufman pivots inside fronts and the inertia read off
    D, and the multi-right-hand-side solve: oracle/cpu_ldlt.cpp, C++ with OpenMP over the levels
    of the assembly tree (and over column blocks of the top fronts' updates), dense kernels from
    the OpenBLAS that scipy ships (dgemm / dtrsm through scipy.linalg.cython_blas, one BLAS thread
    per OpenMP thread);
  * the block Krylov-Schur iteration in the B inner product: the engine's own host driver
    (hsolve_amd.krylov, the restatement of block_lanczos.cpp / lanczos_krylov_schur.cpp) run on
    CPU torch tensors (CpuOps below: the same Gram / update / dot operations on the host);
  * a spectrum slice as the reference's WindowSolver runs one (definite.cpp:1362-1520): exact
    counts nu(lo), nu(hi) from the CPU LDL^T, a shift at the centre, one block Krylov run for
    the counted pairs, Rayleigh-Ritz, and the certificate (pairs found in [lo, hi] == count).

Build: build() compiles cpu_ldlt.cpp with g++ -O3 -fopenmp into oracle/libcpu_ldlt.so.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import time
from pathlib import Path

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import torch

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_ldlt.cpp"
LIB = HERE / "libcpu_ldlt.so"


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-march=x86-64-v3", "-fPIC", "-shared", str(SRC), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"building the CPU LDL^T failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


_lib = None


def _capsule_ptr(cap) -> int:
    get_name = C.pythonapi.PyCapsule_GetName
    get_name.restype = C.c_char_p
    get_name.argtypes = [C.py_object]
    get_ptr = C.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = C.c_void_p
    get_ptr.argtypes = [C.py_object, C.c_char_p]
    return get_ptr(cap, get_name(cap))


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        # idle OpenMP workers of this library sleep instead of spinning: the Krylov layer's BLAS
        # threads run between the library's parallel regions
        os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
        L = C.CDLL(str(LIB))
        P, I64 = C.c_void_p, C.c_int64
        L.mf_set_blas.argtypes = [P, P]
        L.mf_analyze.restype = P
        L.mf_analyze.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P]
        L.mf_factorize.argtypes = [P, P, P]
        L.mf_solve.argtypes = [P, P, C.c_int, I64]
        L.mf_free.argtypes = [P]
        L.mf_last_error.restype = C.c_char_p
        L.mf_factor_entries.restype = C.c_double
        L.mf_factor_entries.argtypes = [P]
        L.mf_gram.argtypes = [I64, P, C.c_int, P, C.c_int, P, P]
        L.mf_threads.restype = C.c_int
        L.mf_update.argtypes = [I64, P, C.c_int, P, P, C.c_int, C.c_double, C.c_double]
        import scipy.linalg.cython_blas as cb
        L.mf_set_blas(_capsule_ptr(cb.__pyx_capi__["dgemm"]), _capsule_ptr(cb.__pyx_capi__["dtrsm"]))
        _lib = L
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _one_blas_thread():
    from threadpoolctl import threadpool_limits
    return threadpool_limits(1, "blas")


class CpuLDLT:
    """Multifrontal LDL^T of K = A - s B (real symmetric, B diagonal or identity) on the CPU over the
    engine's host analysis of the pencil's union pattern."""

    def __init__(self, A, B=None):
        from hsolve_amd.factor import GpuFactorization, Pencil  # host analysis only
        t0 = time.perf_counter()
        self.pencil = Pencil(A, B)
        ana = GpuFactorization(self.pencil)
        fr = ana.fronts()
        order = ana.order()
        self.analysis_s = time.perf_counter() - t0
        self.info = ana.info
        ana.close()
        p = self.pencil
        self.n = p.n
        self._keep = [np.ascontiguousarray(x) for x in (fr["first"], fr["ncol"], fr["m"], fr["parent"], fr["rowoff"],
                                                        fr["rows"], order, p.u_rowptr, p.u_col.astype(np.int32))]
        k = self._keep
        self.h = lib().mf_analyze(self.n, int(fr["first"].size), *[_p(x) for x in k])
        if not self.h:
            raise RuntimeError(lib().mf_last_error().decode())
        self.analysis_s = time.perf_counter() - t0
        self.a_vals = np.where(p.a_index >= 0, p.A.data[np.maximum(p.a_index, 0)], 0.0)
        bd = p._Bfull.data
        self.b_vals = np.where(p.b_index >= 0, bd[np.maximum(p.b_index, 0)], 0.0)
        self.inertia = None

    def factorize(self, sigma: float):
        vals = np.ascontiguousarray(self.a_vals - sigma * self.b_vals)
        out = np.zeros(3, dtype=np.int64)
        with _one_blas_thread():  # the library's OpenMP threads each call a single-threaded BLAS
            rc = lib().mf_factorize(self.h, _p(vals), _p(out))
        if rc != 0:
            raise RuntimeError(lib().mf_last_error().decode())
        self.inertia = tuple(int(v) for v in out)
        return self.inertia

    def count(self, s: float) -> int:
        return self.factorize(s)[0]

    def solve(self, X: np.ndarray) -> np.ndarray:
        """X: (n, p) Fortran-order float64, solved in place."""
        X = np.asfortranarray(X, dtype=np.float64)
        with _one_blas_thread():
            rc = lib().mf_solve(self.h, _p(X), X.shape[1], X.shape[0])
        if rc != 0:
            raise RuntimeError(lib().mf_last_error().decode())
        return X

    def close(self):
        if getattr(self, "h", None):
            lib().mf_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CpuOps:
    """The block operations hsolve_amd.krylov needs, on CPU torch tensors of shape (p, n) (rows are
    vectors): B inner products (diagonal B), Gram matrices, block updates."""

    def __init__(self, A, B=None):
        self.n = A.shape[0]
        self.dtype = torch.float64
        self.np_dtype = np.float64
        self.dev = torch.device("cpu")
        self.is_complex = False
        self.A = sp.csr_matrix(A)
        self.bd = None if B is None else torch.from_numpy(np.ascontiguousarray(B.diagonal(), dtype=np.float64))

    def empty(self, p):
        return torch.empty((p, self.n), dtype=self.dtype)

    def zeros(self, p):
        return torch.zeros((p, self.n), dtype=self.dtype)

    def block(self, a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))

    def coef(self, m):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(m).T, dtype=np.float64))

    def apply_b(self, x, y=None):
        r = x if self.bd is None else x * self.bd[None, :]
        if y is None:
            return r.clone() if self.bd is None else r
        y.copy_(r)
        return y

    # Gram and update run in the C++ library (OpenMP row chunks, one BLAS thread each): the BLAS
    # stays single-threaded for the whole CPU path, since a multithreaded BLAS pool spinning
    # between calls slowed the library's OpenMP solves forty-fold on this host
    def gram(self, v, w, weight=None, out=None):
        v, w = v.contiguous(), w.contiguous()
        k, p = v.shape[0], w.shape[0]
        c = torch.empty((p, k), dtype=torch.float64) if out is None else out
        if k and p:
            cc = np.empty((p, k))
            lib().mf_gram(self.n, _p(v.numpy()), k, _p(w.numpy()), p,
                          _p(weight.numpy()) if weight is not None else None, _p(cc))
            c.copy_(torch.from_numpy(cc))
        return c

    def b_gram(self, v, w, bw=None, out=None):
        return self.gram(v, w, self.bd, out)

    def update(self, v, c, w, alpha=1.0, beta=1.0):
        # W = beta W + alpha V C, C given as (p, k)
        k, p = v.shape[0], w.shape[0]
        if p:
            cc = np.ascontiguousarray(c.numpy(), dtype=np.float64)
            assert w.is_contiguous() and v.is_contiguous()
            lib().mf_update(self.n, _p(v.numpy()), k, _p(cc), _p(w.numpy()), p, float(alpha), float(beta))
        return w

    def dots(self, x, y, weight=None, out=None):
        xn, yn = x.numpy(), y.numpy()
        r = np.einsum("ij,ij->i", xn, yn if weight is None else yn * weight.numpy()[None, :])
        r = torch.from_numpy(np.ascontiguousarray(r))
        if out is not None:
            out.copy_(r)
            return out
        return r

    def b_dots(self, x, y, by=None, out=None):
        return self.dots(x, y, self.bd, out)


class CpuShiftInvert:
    """Op = K^-1 B on (p, n) CPU tensors with the CPU factor (no refinement: FP64 factors)."""

    def __init__(self, fac: CpuLDLT, ops: CpuOps):
        self.fac = fac
        self.ops = ops
        self.applications = 0

    def apply(self, x, y=None):
        bx = self.ops.apply_b(x)
        X = np.asfortranarray(bx.numpy().T)
        self.fac.solve(X)
        self.applications += x.shape[0]
        out = torch.from_numpy(np.ascontiguousarray(X.T))
        if y is not None:
            y.copy_(out)
            return y
        return out


def solve_slice(A, B, lower: float, upper: float, block: int = 8, tol: float = 1e-11, seed: int = 2026,
                fac: CpuLDLT | None = None) -> dict:
    """All eigenvalues in [lower, upper] of A x = lambda B x on the CPU: counts at both ends, one
    shift at the centre, one block Krylov-Schur run for the counted pairs, Rayleigh-Ritz in the
    original pencil, certificate = pairs found in the window equal the count. Returns the pairs and
    a timing split."""
    from hsolve_amd.krylov import LARGEST_MAGNITUDE, KrylovOptions, lanczos_krylov_schur
    with _one_blas_thread():
        return _solve_slice(A, B, lower, upper, block, tol, seed, fac, LARGEST_MAGNITUDE, KrylovOptions,
                            lanczos_krylov_schur)


def _solve_slice(A, B, lower, upper, block, tol, seed, fac, LARGEST_MAGNITUDE, KrylovOptions, lanczos_krylov_schur):
    t0 = time.perf_counter()
    if fac is None:
        fac = CpuLDLT(A, B)
    t_ana = time.perf_counter() - t0
    t1 = time.perf_counter()
    nu_lo = fac.count(lower)
    nu_hi = fac.count(upper)
    want = nu_hi - nu_lo
    t_count = time.perf_counter() - t1
    sigma = 0.5 * (lower + upper)
    t1 = time.perf_counter()
    fac.factorize(sigma)
    t_fac = time.perf_counter() - t1
    ops = CpuOps(A, B)
    op = CpuShiftInvert(fac, ops)
    t1 = time.perf_counter()
    res = lanczos_krylov_schur(ops, op.apply, LARGEST_MAGNITUDE,
                               KrylovOptions(nev=want + 2, tol=tol, seed=seed, block_size=block))
    t_iter = time.perf_counter() - t1
    t1 = time.perf_counter()
    V = res.vectors.numpy().T  # n x nev
    AV = A @ V
    BV = V if B is None else B @ V
    H = V.T @ AV
    M = V.T @ BV
    lam, S = sla.eigh(0.5 * (H + H.T), 0.5 * (M + M.T)) if V.shape[1] else (np.zeros(0), np.zeros((0, 0)))
    X = V @ S
    keep = (lam >= lower) & (lam <= upper)
    lam, X = lam[keep], X[:, keep]
    t_rr = time.perf_counter() - t1
    return {"eigenvalues": lam, "eigenvectors": X, "count_inertia": want, "count_found": int(lam.size),
            "passed": int(lam.size) == want, "sigma": sigma, "op_applies": op.applications,
            "restarts": res.restarts, "time_analysis_s": t_ana, "time_counts_s": t_count,
            "time_factor_s": t_fac, "time_iterate_s": t_iter, "time_rr_s": t_rr,
            "time_total_s": time.perf_counter() - t0, "threads": int(lib().mf_threads())}


def host_description() -> dict:
    """CPU model and thread counts of this host (for the baseline's record)."""
    model = ""
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 0
    return {"cpu_model": model, "nproc": os.cpu_count() or 0, "affinity_cpus": affinity,
            "omp_threads": int(lib().mf_threads())}
