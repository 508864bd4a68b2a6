"""The C ABI: every function include/hsg.h declares is exported by libhsg.so and bound in the
ctypes table with the declared number of arguments (no compute calls: runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "hsg.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(hsg_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_header_declares_the_entry_points():
    decl = _declared()
    for name in ("hsg_analyze", "hsg_factorize", "hsg_factor_solve", "hsg_release_factors", "hsg_factor_destroy",
                 "hsg_last_error_message", "hsg_version", "hsg_timing_read"):
        assert name in decl


def test_library_exports_every_declared_symbol():
    from hsolve_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libhsg.so not built")
    so = C.CDLL(str(_lib.LIB_PATH))
    decl = _declared()
    for name, nargs in decl.items():
        assert hasattr(so, name), f"{name} declared in include/hsg.h but not exported"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
        assert len(_lib.SIGNATURES[name][1]) == nargs, f"{name}: {nargs} declared arguments"
    assert set(_lib.SIGNATURES) == set(decl), "ctypes table and header differ"


def test_version_and_error_channel():
    from hsolve_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libhsg.so not built")
    lib = _lib.lib()
    assert lib.hsg_version().decode().count(".") == 2
    assert lib.hsg_timing_read(99, None, None, None, None) == 1  # invalid phase -> HSG_INVALID_INPUT
    assert b"phase" in lib.hsg_last_error_message()


def test_binding_keeps_the_gil_only_across_launch_only_calls():
    # long entry points (analysis, factorization, solve, clone, frees) release the GIL; the
    # launch-only ones keep it (a thread that released it for microseconds could wait a whole
    # interpreter switch interval to get it back while other slice threads run Python)
    import ctypes as C
    from hsolve_amd import _lib
    assert _lib.RELEASES_GIL <= set(_lib.SIGNATURES)
    assert {"hsg_analyze", "hsg_factorize", "hsg_factor_solve"} <= _lib.RELEASES_GIL
    if not _lib.LIB_PATH.exists():
        pytest.skip("libhsg.so not built")
    b = _lib.Bound(_lib.LIB_PATH)
    for name in _lib.SIGNATURES:
        keeps = bool(getattr(b, name)._flags_ & C._FUNCFLAG_PYTHONAPI)
        assert keeps == (name not in _lib.RELEASES_GIL), name
