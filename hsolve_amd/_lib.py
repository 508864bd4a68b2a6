"""ctypes binding of libhsg.so (include/hsg.h). Loading fails loudly when the library was not
built: there is no CPU fallback for any of its entry points."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from .errors import error_from_code

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libhsg.so"


class AnalysisOptions(C.Structure):
    _fields_ = [("leaf_size", C.c_int32), ("panel_width", C.c_int32), ("small_front_max", C.c_int32),
                ("reserved", C.c_int32 * 5)]


class AnalysisInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("supernodes", C.c_int64), ("levels", C.c_int64),
                ("max_front", C.c_int64), ("max_pivots", C.c_int64), ("factor_entries", C.c_double),
                ("flops", C.c_double), ("panel_bytes_double", C.c_double), ("cb_bytes_double", C.c_double),
                ("analyze_s", C.c_double), ("ordering_s", C.c_double), ("small_fronts", C.c_int64),
                ("large_fronts", C.c_int64), ("front_rows", C.c_int64), ("update_flops", C.c_double),
                ("cb_bytes_unshared", C.c_double), ("workspace_bytes_double", C.c_double)]


class Inertia(C.Structure):
    _fields_ = [("neg", C.c_int64), ("zero", C.c_int64), ("pos", C.c_int64), ("valid", C.c_int32),
                ("two_by_two", C.c_int32)]


_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "hsg_version": (C.c_char_p, []),
    "hsg_last_error_message": (C.c_char_p, []),
    "hsg_analyze": (C.c_int, [_I64, _P, _P, _P, _P, _P, _P, C.POINTER(_P)]),
    "hsg_analysis_get_info": (C.c_int, [_P, C.POINTER(AnalysisInfo)]),
    "hsg_analysis_get_order": (C.c_int, [_P, _P]),
    "hsg_analysis_get_fronts": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "hsg_factorize": (C.c_int, [_P, _D, _D, _P, _P, C.c_int32, _D, _P, C.POINTER(Inertia), C.POINTER(C.c_int32)]),
    "hsg_factorize_batch": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, C.c_int32, _D, _P, C.POINTER(Inertia)]),
    "hsg_factor_solve": (C.c_int, [_P, _P, _I64, _I64, _P]),
    "hsg_lanczos_coefficients": (C.c_int, [C.c_int32, _I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P]),
    "hsg_release_factors": (C.c_int, [_P]),
    "hsg_factor_destroy": (None, [_P]),
    "hsg_factor_clone": (C.c_int, [_P, C.POINTER(_P)]),
    "hsg_release_cached_memory": (C.c_int, [_P]),
    "hsg_csr_spmm": (C.c_int, [_I64, _P, _P, _P, _P, _I64, _P, _I64, _I64, _D, _D, _P]),
    "hsg_diag_spmm": (C.c_int, [_I64, _P, _P, _I64, _P, _I64, _I64, _D, _D, _P]),
    "hsg_shifted_residual": (C.c_int, [_I64, _P, _P, _P, _P, _P, _P, _P, _D, _P, _I64, _P, _I64, _P, _I64, _I64,
                                       _P, _P, _P]),
    "hsg_gram": (C.c_int, [_I64, _P, _I64, _I64, _P, _P, _I64, _I64, _P, _I64, _P]),
    "hsg_gram_noalloc": (C.c_int, [_I64, _P, _I64, _I64, _P, _P, _I64, _I64, _P, _I64, _P]),
    "hsg_block_update": (C.c_int, [_I64, _P, _I64, _I64, _P, _I64, _P, _I64, _I64, _D, _D, _P]),
    "hsg_column_dots": (C.c_int, [_I64, _P, _I64, _P, _I64, _I64, _P, _P, _P]),
    "hsg_axpby": (C.c_int, [_I64, _D, _P, _I64, _D, _P, _I64, _I64, _P]),
    "hsg_pair_sums": (C.c_int, [_I64, _P, _P, _P, _I64, _I64, _P, _P, _P]),
    "hsg_block_orthonormalize": (C.c_int, [C.c_int32, _I64, _P, _I64, _I64, _P, _I64, _I64, _P, _P, _P, _P, _D, _P,
                                           _P, _P, _P, C.POINTER(C.c_int32), _P]),
    "hsg_timing_enable": (C.c_int, [C.c_int]),
    "hsg_timing_read": (C.c_int, [C.c_int, _P, _P, _P, _P]),
}
# complex variants: the same argument lists (values and vectors as interleaved doubles)
for _name in ("csr_spmm", "diag_spmm", "shifted_residual", "gram", "block_update", "column_dots"):
    SIGNATURES["hsg_z" + _name] = SIGNATURES["hsg_" + _name]

# Entry points that run long on the host (analysis, the numeric factorization and its inertia, a
# solve's level-by-level launches, factor copies, a block orthonormalization with its two waits on
# the device) release the GIL while they run. Every other entry
# point only enqueues one or two kernels on a stream and returns within microseconds: those keep the
# GIL. Concurrent spectrum slices drive the device from several Python threads, and a thread that
# gave up the GIL for a short call could wait a full interpreter switch interval to get it back.
RELEASES_GIL = frozenset({"hsg_analyze", "hsg_factorize", "hsg_factorize_batch", "hsg_factor_solve", "hsg_factor_clone", "hsg_release_factors",
                          "hsg_factor_destroy", "hsg_release_cached_memory", "hsg_block_orthonormalize"})


class Bound:
    """libhsg entry points with their ctypes signatures, one attribute per exported name."""

    def __init__(self, path: Path):
        releasing, holding = C.CDLL(str(path)), C.PyDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(releasing if name in RELEASES_GIL else holding, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)


_lib = None


def lib() -> Bound:
    """The loaded library; raises ImportError if it is missing (build with python -m hsolve_amd.build)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("HSG_LIBRARY", str(LIB_PATH)))
        if not path.exists():
            raise ImportError(f"libhsg.so not found at {path}: build it with `python -m hsolve_amd.build` "
                              "(the HIP engine has no CPU fallback)")
        _lib = Bound(path)
    return _lib


def check(code: int) -> None:
    if code != 0:
        msg = lib().hsg_last_error_message().decode(errors="replace")
        raise error_from_code(code, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def version() -> str:
    return lib().hsg_version().decode()
