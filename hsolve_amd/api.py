"""hsolve.solve() drop-in (python/hsolve/__init__.py:135-171, :241-419).

The signature, the keyword rules, the result type and the error classes are the reference's
contract; the code below is this package's own. Matrices are scipy.sparse, dense numpy arrays or
paths to .npz / .npy / .mtx files. The solve runs in this package's GPU engine
(hsolve_amd.solver); results come back as numpy arrays (eigenvectors n x m in Fortran order) with
the report as a dict of report.json's keys.

Keyword rules (what the caller's keywords do to the configuration, in this order):
  * config: None, a dict (deep-copied; numpy scalars become plain numbers) or a path to a JSON
    file whose top level is an object. Unreadable file: HsolveIOError; bad JSON: InvalidInputError.
  * mode given explicitly: must be one of MODES and sets target.mode. Otherwise lower or upper
    given selects "interval" (an error if sigma is given too), sigma alone selects "nearest", and
    without any of them config.json's target.mode stands.
  * sigma / lower / upper (when not None) and k (when passed) set the target fields; tol and
    threads (when passed) set the top-level fields. Real-valued keywords take ints, floats, numpy
    real scalars and 0-d arrays, never bool or str (TypeError); integer keywords take whatever
    operator.index accepts, never bool (TypeError).
A keyword counts as passed when it appears in the call, whatever its value; the defaults shown in
the signature are the ones config.json itself falls back to.
"""
from __future__ import annotations

import functools
import json
import numbers
import operator
import os
import threading
import warnings
from collections.abc import Mapping
from dataclasses import dataclass
from typing import Any

import numpy as np
import scipy.sparse as sp

from . import config as cfgmod
from .errors import HsolveIOError, HsolveWarning, InvalidInputError


@dataclass(repr=False, eq=False)
class SolveResult:
    """Result of solve() (python/hsolve/__init__.py:135-171)."""

    eigenvalues: np.ndarray
    eigenvectors: np.ndarray
    complex_eigenvalues: np.ndarray
    complex_eigenvectors: np.ndarray
    report: dict
    config: dict

    @property
    def success(self) -> bool:
        return bool(self.report.get("success", False))

    def __repr__(self) -> str:
        n = self.eigenvectors.shape[0]
        return (f"SolveResult(n={n}, eigenvalues={self.eigenvalues.size}, "
                f"complex_eigenvalues={self.complex_eigenvalues.size}, problem={self.report.get('problem')!r}, "
                f"success={self.success})")


MODES = ("nearest", "interval", "smallest", "largest")
STORAGES = ("auto", "full", "upper", "lower")
_SOLVE_LOCK = threading.Lock()  # one solve at a time per process (module.cpp:14-16)


def as_real(value: Any, what: str) -> float:
    """value as a float when it is a real number (bool and str excluded); TypeError otherwise."""
    if isinstance(value, (bool, np.bool_)):
        ok = False
    elif isinstance(value, np.ndarray):
        ok = value.ndim == 0 and value.dtype.kind in "iuf"
    else:
        ok = isinstance(value, numbers.Real)
    if not ok:
        raise TypeError(f"{what} must be a real number, got {type(value).__name__}")
    return float(value)


def as_int(value: Any, what: str) -> int:
    """value as an int when it supports operator.index (bool excluded); TypeError otherwise."""
    if not isinstance(value, (bool, np.bool_)):
        try:
            return operator.index(value)
        except TypeError:
            pass
    raise TypeError(f"{what} must be an integer, got {type(value).__name__}")


def _plain(o: Any) -> Any:
    """JSON encoder hook: numpy scalars (and 0-d arrays) become their Python values."""
    if hasattr(o, "item") and np.ndim(o) == 0:
        return o.item()
    return str(o)


def load_config(config: Any) -> dict:
    """The caller's config argument as a fresh dict of plain JSON values."""
    if config is None:
        return {}
    if isinstance(config, Mapping):
        return json.loads(json.dumps(dict(config), default=_plain))
    if not isinstance(config, (str, bytes, os.PathLike)):
        raise TypeError(f"config must be a dict, a path to config.json or None, got {type(config).__name__}")
    path = os.fsdecode(config)
    try:
        with open(path, encoding="utf-8") as fh:
            text = fh.read()
    except OSError as e:
        raise HsolveIOError(f"cannot read config file {path}: {e}") from e
    try:
        doc = json.loads(text)
    except json.JSONDecodeError as e:
        raise InvalidInputError(f"config file {path}: invalid JSON: {e}") from None
    if not isinstance(doc, dict):
        raise InvalidInputError(f"config file {path}: top level must be a JSON object")
    return doc


def _target_mode(given: frozenset, mode: Any, sigma, lower, upper) -> str | None:
    """target.mode implied by the keywords (None: keep config.json's)."""
    if "mode" in given:
        if not isinstance(mode, str):
            raise TypeError(f"mode must be a string, got {type(mode).__name__}")
        if mode not in MODES:
            raise InvalidInputError(f"mode = {mode!r} is not one of {MODES}")
        return mode
    if lower is not None or upper is not None:
        if sigma is not None:
            raise InvalidInputError("give sigma (nearest) or lower/upper (interval), not both, or pass mode=")
        return "interval"
    return "nearest" if sigma is not None else None


def merge_config(config, given: frozenset = frozenset(), *, sigma=None, k=6, mode="nearest", lower=None,
                 upper=None, tol=1e-10, threads=0) -> dict:
    """The configuration a solve runs with: config (see load_config) overridden by the keywords in
    `given` and by sigma / lower / upper where not None (python/hsolve/__init__.py:241-278), then
    validated and completed by config.parse_config."""
    cfg = load_config(config)
    target = cfg.get("target", {})
    if not isinstance(target, dict):
        raise InvalidInputError("config: 'target' must be an object")
    target = dict(target)
    m = _target_mode(given, mode, sigma, lower, upper)
    if m is not None:
        target["mode"] = m
    for key, value in (("sigma", sigma), ("lower", lower), ("upper", upper)):
        if value is not None:
            target[key] = as_real(value, key)
    if "k" in given:
        target["k"] = as_int(k, "k")
    if target:
        cfg["target"] = target
    if "tol" in given:
        cfg["tol"] = as_real(tol, "tol")
    if "threads" in given:
        cfg["threads"] = as_int(threads, "threads")
    return cfgmod.parse_config(cfg)


def _mirror(M: sp.csr_matrix, storage: str) -> sp.csr_matrix:
    if storage == "full":
        return M
    upper_nnz = sp.triu(M, 1).nnz
    lower_nnz = sp.tril(M, -1).nnz
    if storage == "auto":
        if upper_nnz and lower_nnz:
            return M
        if not upper_nnz and not lower_nnz:
            return M
        storage = "upper" if upper_nnz else "lower"
    T = sp.triu(M) if storage == "upper" else sp.tril(M)
    strict = sp.triu(T, 1) if storage == "upper" else sp.tril(T, -1)
    return (T + strict.conj().T).tocsr()


def _matrix(m, name: str, storage: str, notes: list | None = None, hermitian_tol: float = 1e-10):
    if m is None:
        return None
    if isinstance(m, (str, bytes, os.PathLike)):
        from .io import load_matrix
        loaded = load_matrix(os.fsdecode(m), storage=storage, hermitian_tol=hermitian_tol)
        if notes is not None:
            notes.extend(f"{name}: {n}" for n in loaded.notes)
        return loaded.matrix
    if sp.issparse(m):
        M = sp.csr_matrix(m)
    else:
        if hasattr(m, "matvec") and not isinstance(m, np.ndarray):
            raise TypeError(f"{name}: a matrix-free operator cannot be factorized; pass the matrix entries "
                            "(scipy.sparse, a dense array or a file path)")
        arr = np.asarray(m)
        if arr.dtype == object:
            raise InvalidInputError(f"{name}: cannot interpret {type(m).__name__} as a numeric matrix")
        if arr.ndim != 2:
            raise InvalidInputError(f"{name}: expected a 2-D array, got {arr.ndim}-D")
        M = sp.csr_matrix(arr)
    if M.shape[0] != M.shape[1]:
        raise InvalidInputError(f"{name}: matrix must be square, got shape {M.shape}")
    if M.dtype.kind not in "fciub":
        raise InvalidInputError(f"{name}: unsupported dtype {M.dtype}")
    if M.dtype.kind in "iub":
        M = M.astype(np.float64)
    M.sum_duplicates()
    M.eliminate_zeros()
    return _mirror(M, storage)


def _records_keywords(fn):
    """Calls fn with _given = the set of keyword names the caller wrote (values aside)."""
    @functools.wraps(fn)
    def call(*args, **kwargs):
        return fn(*args, _given=frozenset(kwargs), **kwargs)
    return call


@_records_keywords
def solve(A: Any, B: Any = None, *, sigma: float | None = None, k: int = 6, mode: str = "nearest",
          lower: float | None = None, upper: float | None = None, tol: float = 1e-10, config=None,
          threads: int = 0, verbose: int = 0, storage: str = "auto", _given: frozenset = frozenset()) -> SolveResult:
    """Solve A x = lambda B x (B=None: A x = lambda x) for Hermitian A and HPD B on the GPU.

    Arguments as hsolve.solve (python/hsolve/__init__.py:347-419); the keyword rules are in the
    module docstring. Raises an HsolveError subclass on failure; a run that finishes without
    meeting tol returns with success False and issues an HsolveWarning.
    """
    if storage not in STORAGES:
        raise InvalidInputError(f"storage = {storage!r} is not one of {STORAGES}")
    cfg = merge_config(config, _given, sigma=sigma, k=k, mode=mode, lower=lower, upper=upper, tol=tol,
                       threads=threads)
    as_int(verbose, "verbose")
    notes: list = []
    Am = _matrix(A, "A", storage, notes, cfg["hermitian_tol"])
    Bm = _matrix(B, "B", storage, notes, cfg["hermitian_tol"])
    complex_dtype = np.iscomplexobj(Am.data) or (Bm is not None and np.iscomplexobj(Bm.data))
    from .solver import solve as engine_solve
    with _SOLVE_LOCK:
        w, v, report, cfg_out = engine_solve(Am, Bm, cfg)
    report.pop("_events", None)
    report["warnings"].extend("input: " + n for n in notes)
    if complex_dtype:
        v = v.astype(np.complex128, order="F")
    res = SolveResult(np.asarray(w, dtype=np.float64), v, np.zeros(0, dtype=np.complex128),
                      np.zeros((v.shape[0], 0), dtype=np.complex128, order="F"), report, cfg_out)
    if not res.success:
        warnings.warn("hsolve: the solve finished without meeting the tolerance (see result.report['pairs'] and "
                      "result.report['warnings'])", HsolveWarning, stacklevel=3)
    return res


def default_config() -> dict:
    return cfgmod.default_config()


def version() -> str:
    return cfgmod.VERSION
