"""On-disk layout of cases, references and results, as the reference's tools use it
(tools/case.py:1-22; formats in its docs/FORMATS.md), so that a case solved by this engine and the
result directory it writes are read and judged exactly as the reference's runs were.

    <root>/cases/<name>/A.npz          CSR, full Hermitian storage (scipy.sparse.save_npz)
    <root>/cases/<name>/B.npz          optional
    <root>/cases/<name>/meta.json      provenance (family, problem, scenarios; never reference values)
    <root>/cases/<name>/config.json    the solver configuration to run the case with
    <root>/refs/<name>/eigenvalues.npy reference eigenvalues, ascending
    <root>/refs/<name>/eigenvectors.npy optional (n x m, Fortran order)
    <root>/refs/<name>/ref.json        kind, note, count, complete, window
    result directory: eigenvalues.npy, eigenvectors.npy (n x m, Fortran order),
                      [complex_eigenvalues.npy, complex_eigenvectors.npy], report.json

A case or reference may also be given as a directory path holding those files directly (the
test fixtures keep a case and its reference in one directory). The root defaults to
$HS_DATA_ROOT, else ./data.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg  # noqa: F401  (sp.linalg.norm)

SCHEMA_VERSION = 1
PROBLEMS = ("HEP", "GHEP", "GHIEP")
REF_KINDS = ("analytic", "discrete-exact", "dense", "solver", "partial-analytic")
SCENARIOS = ("zero-cluster", "multiplicity", "near-singular", "indefinite-B", "complex-hermitian", "memory",
             "clustered", "large")


def default_root() -> Path:
    return Path(os.environ.get("HS_DATA_ROOT", "data"))


def _csr(m) -> sp.csr_matrix:
    m = sp.csr_matrix(m)
    m.sum_duplicates()
    m.sort_indices()
    return m


def _write_json(path: Path, doc: dict) -> None:
    def plain(x):
        if isinstance(x, (np.integer, np.floating)):
            return x.item()
        if isinstance(x, np.ndarray):
            return x.tolist()
        if isinstance(x, Path):
            return str(x)
        raise TypeError(f"not JSON serializable: {type(x).__name__}")
    path.write_text(json.dumps(doc, indent=2, ensure_ascii=False, default=plain) + "\n")


def _defect(m: sp.csr_matrix) -> float:
    nrm = sp.linalg.norm(m)
    return 0.0 if nrm == 0 else float(sp.linalg.norm(m - m.conj().T) / nrm)


@dataclass
class Case:
    name: str
    path: Path
    A: sp.csr_matrix
    B: sp.csr_matrix | None
    meta: dict
    config: dict

    @property
    def n(self) -> int:
        return int(self.A.shape[0])


def _case_dir(name_or_path, root) -> Path:
    p = Path(name_or_path)
    if (p / "A.npz").exists():
        return p
    return Path(root if root is not None else default_root()) / "cases" / str(name_or_path)


def load_case(name_or_path, root=None) -> Case:
    d = _case_dir(name_or_path, root)
    if not (d / "A.npz").exists():
        raise FileNotFoundError(f"no case at {d} (A.npz missing)")
    A = sp.load_npz(d / "A.npz").tocsr()
    B = sp.load_npz(d / "B.npz").tocsr() if (d / "B.npz").exists() else None
    meta = json.loads((d / "meta.json").read_text()) if (d / "meta.json").exists() else {}
    config = json.loads((d / "config.json").read_text()) if (d / "config.json").exists() else {}
    return Case(d.name, d, A, B, meta, config)


def save_case(name: str, A, B=None, *, meta: dict, config: dict | None = None, root=None,
              hermitian_tol: float | None = 1e-12) -> Path:
    """Writes <root>/cases/<name>/; meta must name the family and the problem kind and may not
    carry reference values (those go to save_ref)."""
    if "family" not in meta or "problem" not in meta:
        raise ValueError("meta needs 'family' and 'problem'")
    if meta["problem"] not in PROBLEMS:
        raise ValueError(f"meta['problem'] must be one of {PROBLEMS}")
    unknown = [s for s in meta.get("scenarios", []) if s not in SCENARIOS]
    if unknown:
        raise ValueError(f"unknown scenario tags {unknown}; allowed {SCENARIOS}")
    leaked = [k for k in ("reference", "ref_eigenvalues", "eigenvalues") if k in meta]
    if leaked:
        raise ValueError(f"meta must not carry reference data ({leaked}); use save_ref()")
    d = Path(root if root is not None else default_root()) / "cases" / name
    d.mkdir(parents=True, exist_ok=True)
    A = _csr(A)
    if A.shape[0] != A.shape[1]:
        raise ValueError("A must be square")
    if hermitian_tol is not None and _defect(A) > hermitian_tol:
        raise ValueError(f"A is not Hermitian (defect {_defect(A):.3e})")
    sp.save_npz(d / "A.npz", A, compressed=True)
    doc = dict(meta, schema_version=SCHEMA_VERSION, name=name, n=int(A.shape[0]), nnz_A=int(A.nnz),
               dtype="complex128" if np.iscomplexobj(A.data) else "float64")
    if B is not None:
        B = _csr(B)
        if B.shape != A.shape:
            raise ValueError("B shape differs from A")
        if hermitian_tol is not None and _defect(B) > hermitian_tol:
            raise ValueError(f"B is not Hermitian (defect {_defect(B):.3e})")
        sp.save_npz(d / "B.npz", B, compressed=True)
        doc["nnz_B"] = int(B.nnz)
        if np.iscomplexobj(B.data):
            doc["dtype"] = "complex128"
    elif (d / "B.npz").exists():
        (d / "B.npz").unlink()
    _write_json(d / "meta.json", doc)
    _write_json(d / "config.json", config if config is not None else
                {"target": {"mode": "nearest", "sigma": 0.0, "k": 6}, "tol": 1e-10})
    return d


@dataclass
class Reference:
    eigenvalues: np.ndarray
    eigenvectors: np.ndarray | None
    info: dict = field(default_factory=dict)


def save_ref(name: str, eigenvalues, *, kind: str, note: str = "", complete: bool = False, window=None,
             eigenvectors=None, root=None) -> Path:
    """Writes <root>/refs/<name>/; complete = every eigenvalue of the window (of the whole
    spectrum without a window) is listed, which lets the evaluator detect misses."""
    if kind not in REF_KINDS:
        raise ValueError(f"kind must be one of {REF_KINDS}")
    d = Path(root if root is not None else default_root()) / "refs" / name
    d.mkdir(parents=True, exist_ok=True)
    w = np.asarray(eigenvalues)
    order = np.argsort(w.real if np.iscomplexobj(w) else w, kind="stable")
    np.save(d / "eigenvalues.npy", w[order])
    if eigenvectors is not None:
        np.save(d / "eigenvectors.npy", np.asfortranarray(np.asarray(eigenvectors)[:, order]))
    elif (d / "eigenvectors.npy").exists():
        (d / "eigenvectors.npy").unlink()
    _write_json(d / "ref.json", {"schema_version": SCHEMA_VERSION, "kind": kind, "note": note, "count": int(w.size),
                                 "complete": bool(complete), "window": list(window) if window is not None else None})
    return d


def load_ref(name_or_path, root=None) -> Reference | None:
    p = Path(name_or_path)
    d = p if (p / "eigenvalues.npy").exists() else Path(root if root is not None else default_root()) / "refs" / str(
        name_or_path)
    if not (d / "eigenvalues.npy").exists():
        return None
    vecs = np.load(d / "eigenvectors.npy") if (d / "eigenvectors.npy").exists() else None
    info = json.loads((d / "ref.json").read_text()) if (d / "ref.json").exists() else {}
    return Reference(np.load(d / "eigenvalues.npy"), vecs, info)


@dataclass
class Result:
    eigenvalues: np.ndarray
    eigenvectors: np.ndarray
    complex_eigenvalues: np.ndarray | None
    complex_eigenvectors: np.ndarray | None
    report: dict


def save_result(out_dir, eigenvalues, eigenvectors, report: dict, complex_eigenvalues=None,
                complex_eigenvectors=None) -> Path:
    """Writes a result directory; complex pair files only when non-empty (stale ones removed)."""
    d = Path(out_dir)
    d.mkdir(parents=True, exist_ok=True)
    np.save(d / "eigenvalues.npy", np.asarray(eigenvalues, dtype=float))
    np.save(d / "eigenvectors.npy", np.asfortranarray(eigenvectors))
    for fname, arr in (("complex_eigenvalues.npy", complex_eigenvalues),
                       ("complex_eigenvectors.npy", complex_eigenvectors)):
        if arr is not None and np.size(arr):
            np.save(d / fname, arr)
        elif (d / fname).exists():
            (d / fname).unlink()
    doc = dict(report)
    doc.setdefault("schema_version", SCHEMA_VERSION)
    _write_json(d / "report.json", doc)
    return d


def load_result(out_dir) -> Result:
    d = Path(out_dir)
    ce, cv = d / "complex_eigenvalues.npy", d / "complex_eigenvectors.npy"
    return Result(np.load(d / "eigenvalues.npy"), np.load(d / "eigenvectors.npy"),
                  np.load(ce) if ce.exists() else None, np.load(cv) if cv.exists() else None,
                  json.loads((d / "report.json").read_text()) if (d / "report.json").exists() else {})
