"""The Python half of libhsg_solve.so's hsg_solve and hsg_parse_config (include/hsg_solve.h): builds
the matrices from the caller's CSR arrays, runs solve() and returns what the C side copies into
its result; validates and completes a config document.

The reference's counterpart is py_solve + to_python (python/src/module.cpp:313-345): parse the
config, take the matrix arguments, run the solve with the solve lock held, and return
(w, v, complex_w, complex_v, report_json, config_json, complex_dtype). The C side passes host
addresses and sizes; nothing here keeps a reference to the caller's memory after the call.
"""
from __future__ import annotations

import ctypes
import json
import warnings

import numpy as np
import scipy.sparse as sp

from .config import config_to_json, report_to_json
from .errors import ERROR_CLASSES, HsolveError, InvalidInputError

INTERNAL = len(ERROR_CLASSES)  # the code of InternalError (1 + its position)


def _array(address: int, ctype, count: int) -> np.ndarray:
    if count == 0:
        return np.zeros(0, dtype=np.dtype(ctype))
    if not address:
        raise InvalidInputError("hsg_solve: a CSR array pointer is NULL")
    return np.ctypeslib.as_array((ctype * count).from_address(address)).copy()


def csr_from_addresses(desc, name: str) -> sp.csr_matrix:
    """An n x n csr_matrix copied from (n, indptr, indices, values, is_complex) host addresses."""
    n, p_ptr, p_idx, p_val, is_complex = desc
    n = int(n)
    if n < 0:
        raise InvalidInputError(f"hsg_solve: {name}.n = {n} is negative")
    indptr = _array(p_ptr, ctypes.c_int64, n + 1)
    if indptr[0] != 0 or np.any(np.diff(indptr) < 0):
        raise InvalidInputError(f"hsg_solve: {name}.indptr must start at 0 and not decrease")
    nnz = int(indptr[n])
    indices = _array(p_idx, ctypes.c_int32, nnz)
    if nnz and (indices.min() < 0 or indices.max() >= n):
        raise InvalidInputError(f"hsg_solve: {name}.indices out of range [0, {n})")
    if is_complex:
        values = _array(p_val, ctypes.c_double, 2 * nnz).view(np.complex128)
    else:
        values = _array(p_val, ctypes.c_double, nnz)
    return sp.csr_matrix((values, indices, indptr), shape=(n, n))


def error_code(exc: BaseException) -> int:
    """hsg.h code of an exception: 1 + the position of its HsolveError class, else internal."""
    if isinstance(exc, HsolveError):
        for i, cls in enumerate(ERROR_CLASSES):
            if isinstance(exc, cls):
                return i + 1
    return INTERNAL


def _config_doc(config_json: str):
    if not config_json:
        return None
    try:
        config = json.loads(config_json)
    except json.JSONDecodeError as e:
        raise InvalidInputError(f"config_json: {e}") from None
    if not isinstance(config, dict):
        raise InvalidInputError("config_json: the top level must be an object")
    return config


def parse_config_json(config_json: str):
    """(code, text): the validated config completed with the defaults as JSON text (code 0), or an
    error code and message (module.cpp:466-467, parse_config then config_to_json)."""
    try:
        from . import config as cfgmod
        return (0, config_to_json(cfgmod.parse_config(_config_doc(config_json) or {}), None))
    except BaseException as e:
        return (error_code(e), f"{type(e).__name__}: {e}")


def solve_addresses(a_desc, b_desc, config_json: str, verbose: int, storage: str):
    """(code, message, w, v_columns, report_json, config_json, is_complex): on success code 0, w
    the eigenvalues (float64, m), v_columns the eigenvectors as an m x n C-ordered array (the
    n x m column-major matrix; complex128 for complex input); on failure code > 0 and message."""
    try:
        from . import api
        A = csr_from_addresses(a_desc, "a")
        B = None if b_desc is None else csr_from_addresses(b_desc, "b")
        if B is not None and B.shape != A.shape:
            raise InvalidInputError(f"B has order {B.shape[0]}, A has {A.shape[0]}")
        config = _config_doc(config_json)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")  # the report's "success" carries what the warning says
            res = api.solve(A, B, config=config, verbose=int(verbose), storage=storage or "auto")
        complex_out = bool(np.iscomplexobj(res.eigenvectors))
        vt = np.ascontiguousarray(res.eigenvectors.T)
        w = np.ascontiguousarray(res.eigenvalues, dtype=np.float64)
        return (0, "", w, vt, report_to_json(res.report, None), config_to_json(res.config, None), complex_out)
    except BaseException as e:  # every failure goes back to the C caller as a code and a message
        return (error_code(e), f"{type(e).__name__}: {e}", None, None, "", "", False)
