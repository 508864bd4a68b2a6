"""hsolve.solve() drop-in (python/hsolve/__init__.py:135-171, :241-419).

Same signature, keyword handling and config merging as the reference; matrices are scipy.sparse,
dense numpy arrays or paths to .npz / .npy / .mtx files. The solve runs in this package's GPU
engine (hsolve_amd.solver) instead of the reference's C++ extension; results come back as numpy
arrays (eigenvectors n x m in Fortran order) with the report as a dict of report.json's keys.
"""
from __future__ import annotations

import json
import numbers
import operator
import os
import threading
import warnings
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


class _Default:
    __slots__ = ("value",)

    def __init__(self, value):
        self.value = value

    def __repr__(self):
        return repr(self.value)


_MODES = ("nearest", "interval", "smallest", "largest")
_STORAGES = ("auto", "full", "upper", "lower")
_SOLVE_LOCK = threading.Lock()  # one solve at a time per process (module.cpp:14-16)


def _real(x, name):
    if isinstance(x, np.ndarray) and x.ndim == 0 and x.dtype.kind in "iuf":
        return float(x)
    if isinstance(x, (bool, np.bool_)) or not isinstance(x, numbers.Real):
        raise TypeError(f"{name} must be a real number, got {type(x).__name__}")
    return float(x)


def _integer(x, name):
    if isinstance(x, (bool, np.bool_)):
        raise TypeError(f"{name} must be an integer, got bool")
    try:
        return operator.index(x)
    except TypeError as e:
        raise TypeError(f"{name} must be an integer, got {type(x).__name__}") from e


def _read_config(config):
    if config is None:
        return {}
    if isinstance(config, dict):
        return json.loads(json.dumps(config, default=lambda o: o.item() if hasattr(o, "item") else str(o)))
    if isinstance(config, (str, bytes, os.PathLike)):
        path = os.fsdecode(config)
        try:
            with open(path, encoding="utf-8") as f:
                text = f.read()
        except OSError as e:
            raise HsolveIOError(f"cannot read config file {path}: {e}") from e
        try:
            doc = json.loads(text)
        except json.JSONDecodeError as e:
            raise InvalidInputError(f"config file {path}: invalid JSON: {e}") from None
        if not isinstance(doc, dict):
            raise InvalidInputError(f"config file {path}: top level must be a JSON object")
        return doc
    raise TypeError(f"config must be a dict, a path to config.json or None, got {type(config).__name__}")


def merge_config(config, sigma, k, mode, lower, upper, tol, threads) -> dict:
    """_merge_config (python/hsolve/__init__.py:241-278)."""
    cfg = _read_config(config)
    target = cfg.get("target", {})
    if not isinstance(target, dict):
        raise InvalidInputError("config: 'target' must be an object")
    target = dict(target)
    explicit_mode = not isinstance(mode, _Default)
    interval_bounds = lower is not None or upper is not None
    if explicit_mode:
        if not isinstance(mode, str):
            raise TypeError(f"mode must be a string, got {type(mode).__name__}")
        if mode not in _MODES:
            raise InvalidInputError(f"mode = {mode!r} is not one of {_MODES}")
        target["mode"] = mode
    elif interval_bounds:
        if sigma is not None:
            raise InvalidInputError("give sigma (nearest) or lower/upper (interval), not both, or pass mode=")
        target["mode"] = "interval"
    elif sigma is not None:
        target["mode"] = "nearest"
    if sigma is not None:
        target["sigma"] = _real(sigma, "sigma")
    if lower is not None:
        target["lower"] = _real(lower, "lower")
    if upper is not None:
        target["upper"] = _real(upper, "upper")
    if not isinstance(k, _Default):
        target["k"] = _integer(k, "k")
    if target:
        cfg["target"] = target
    if not isinstance(tol, _Default):
        cfg["tol"] = _real(tol, "tol")
    if not isinstance(threads, _Default):
        cfg["threads"] = _integer(threads, "threads")
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
            raise TypeError(f"{name}: hsolve factorizes the matrix and needs its entries, not a LinearOperator")
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


def solve(A: Any, B: Any = None, *, sigma: float | None = None, k: int = _Default(6),
          mode: str = _Default("nearest"), lower: float | None = None, upper: float | None = None,
          tol: float = _Default(1e-10), config=None, threads: int = _Default(0), verbose: int = 0,
          storage: str = "auto") -> SolveResult:
    """Solve A x = lambda B x (B=None: A x = lambda x) for Hermitian A and HPD B on the GPU.

    Arguments as hsolve.solve (python/hsolve/__init__.py:347-419). Raises an HsolveError subclass
    on failure; a run that finishes without meeting tol returns with success False and issues an
    HsolveWarning.
    """
    if storage not in _STORAGES:
        raise InvalidInputError(f"storage = {storage!r} is not one of {_STORAGES}")
    cfg = merge_config(config, sigma, k, mode, lower, upper, tol, threads)
    _integer(verbose, "verbose")
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
                      "result.report['warnings'])", HsolveWarning, stacklevel=2)
    return res


def default_config() -> dict:
    return cfgmod.default_config()


def version() -> str:
    return cfgmod.VERSION
