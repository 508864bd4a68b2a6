"""Independent evaluation of a solve (oracle): the reference's correctness bar.

Restates the rules of tools/eval/evaluate.py:
  * norms ||A||_2, ||B||_2: dense eigenvalues up to DENSE_MAX, else ARPACK (eigsh, LM, tol 1e-6)
    (:151-164);
  * per-pair metrics recomputed from the ORIGINAL matrices: abs residual ||Ax - lambda Bx||,
    relative residual, backward error eta = r / ((||A|| + |lambda| ||B||) ||x||) (:181-202);
  * expected set of a target within the tie tolerance tie_rtol = TIE_FACTOR * tol: values tied
    with the k-th (or an interval end) are interchangeable (:338-365);
  * one-to-one matching of computed to reference eigenvalues within MATCH_RTOL (Hungarian,
    scipy linear_sum_assignment, :292-313);
  * completeness: no missing required value, the boundary ties filled, no spurious value against
    a complete reference (:380-428); B-orthonormality ||X^H B X - I||.
Pinned by tests/test_oracle_evaluate.py (exact dense pairs pass; a dropped, duplicated or
perturbed pair fails the rule it breaks).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg as la
import scipy.sparse as sp
import scipy.sparse.linalg as sla
from scipy.optimize import linear_sum_assignment

NORM_TOL = 1e-6
NORM_NCV = 40
MATCH_RTOL = 1e-6
TIE_FACTOR = 10.0
DUP_TOL = 1e-4
DENSE_MAX = 3000


def norm2(M, seed: int = 0) -> float:
    n = M.shape[0]
    if n <= DENSE_MAX:
        return float(np.max(np.abs(la.eigvalsh(M.toarray() if sp.issparse(M) else M))))
    v0 = np.random.default_rng(seed).uniform(-1, 1, n)
    w = sla.eigsh(M, k=1, which="LM", tol=NORM_TOL, ncv=min(NORM_NCV, n), v0=v0, return_eigenvectors=False)
    return float(np.max(np.abs(w)))


def pair_metrics(A, B, lam, X, n2a: float, n2b: float, tol: float):
    lam = np.asarray(lam, dtype=float)
    X = np.asarray(X)
    BX = X if B is None else B @ X
    R = A @ X - BX * lam[None, :]
    r = np.linalg.norm(R, axis=0)
    bx = np.linalg.norm(BX, axis=0)
    xn = np.linalg.norm(X, axis=0)
    la_ = np.abs(lam)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = r / np.where(la_ > 0, la_ * bx, bx)
        eta = r / ((n2a + la_ * n2b) * xn)
    return [{"lambda": float(lam[i]), "abs_residual": float(r[i]), "rel_residual": float(rel[i]),
             "backward_error": float(eta[i]), "accepted": bool(eta[i] <= tol)} for i in range(lam.size)]


@dataclass
class ExpectedSet:
    required: np.ndarray
    optional: np.ndarray
    need_optional: int


def expected_set(ref: np.ndarray, target: dict, scale: float, rtol: float) -> ExpectedSet:
    ref = np.asarray(ref, dtype=float)
    if target["mode"] == "interval":
        lo, hi = float(target["lower"]), float(target["upper"])
        blo, bhi = rtol * (abs(lo) + scale), rtol * (abs(hi) + scale)
        inside = (ref >= lo - blo) & (ref <= hi + bhi)
        boundary = inside & ((ref < lo + blo) | (ref > hi - bhi))
        return ExpectedSet(np.flatnonzero(inside & ~boundary), np.flatnonzero(boundary), 0)
    k = int(target["k"])
    if ref.size <= k:
        return ExpectedSet(np.arange(ref.size), np.empty(0, dtype=np.int64), 0)
    sigma = float(target.get("sigma", 0.0))
    key = {"nearest": np.abs(ref - sigma), "smallest": ref, "largest": -ref}[target["mode"]]
    order = np.argsort(key, kind="stable")
    kth, mu_k = key[order[k - 1]], ref[order[k - 1]]
    band = rtol * (abs(mu_k) + scale)
    required = np.flatnonzero(key < kth - band)
    optional = np.flatnonzero(np.abs(key - kth) <= band)
    if optional.size == k - required.size:
        return ExpectedSet(np.sort(np.concatenate([required, optional])), np.empty(0, dtype=np.int64), 0)
    return ExpectedSet(required, optional, k - required.size)


def match(lam: np.ndarray, ref: np.ndarray, scale: float, rtol: float = MATCH_RTOL) -> np.ndarray:
    lam = np.asarray(lam, dtype=float)
    ref = np.asarray(ref, dtype=float)
    out = np.full(lam.size, -1, dtype=np.int64)
    if lam.size == 0 or ref.size == 0:
        return out
    tol = rtol * (np.abs(ref) + scale) + np.finfo(float).tiny
    # restrict to a neighbourhood of the computed values (references can be long)
    lo, hi = lam.min(), lam.max()
    span = np.flatnonzero((ref >= lo - 2 * tol.max()) & (ref <= hi + 2 * tol.max()))
    if span.size == 0:
        return out
    r = ref[span]
    dist = np.abs(lam[:, None] - r[None, :]) / tol[span][None, :]
    cost = np.where(dist <= 1.0, dist, lam.size + 1.0)
    ri, ci = linear_sum_assignment(cost)
    ok = cost[ri, ci] <= 1.0
    out[ri[ok]] = span[ci[ok]]
    return out


def evaluate(A, B, lam, X, target: dict, ref: np.ndarray | None, tol: float = 1e-10, ref_complete: bool = True,
             n2a: float | None = None, n2b: float | None = None) -> dict:
    """Verdict dict: max_eta, all_accepted, missing, deficit, spurious, orthogonality, passed."""
    lam = np.asarray(lam, dtype=float)
    A = sp.csr_matrix(A)
    Bm = None if B is None else sp.csr_matrix(B)
    if n2a is None:
        n2a = norm2(A)
    if n2b is None:
        n2b = 1.0 if Bm is None else norm2(Bm)
    scale = n2a / n2b
    pm = pair_metrics(A, Bm, lam, X, n2a, n2b, tol)
    max_eta = max((p["backward_error"] for p in pm), default=0.0)
    BX = X if Bm is None else Bm @ X
    G = np.asarray(X).conj().T @ np.asarray(BX)
    orth = float(np.max(np.abs(G - np.eye(G.shape[0])))) if G.size else 0.0
    out = {"max_eta": max_eta, "all_accepted": all(p["accepted"] for p in pm), "orthogonality": orth,
           "missing": [], "deficit": 0, "spurious": [], "distinct": int(lam.size)}
    if ref is not None:
        rt = TIE_FACTOR * tol
        exp = expected_set(ref, target, scale, rt)
        m = match(lam, ref, scale)
        matched = set(int(j) for j in m[m >= 0])
        out["missing"] = [float(ref[j]) for j in exp.required if j not in matched]
        opt_found = sum(1 for j in exp.optional if j in matched)
        out["deficit"] = max(0, exp.need_optional - opt_found)
        if ref_complete:
            out["spurious"] = [float(lam[i]) for i in range(lam.size) if m[i] < 0]
        out["expected"] = int(exp.required.size + exp.need_optional) if target["mode"] != "interval" else \
            int(exp.required.size)
    out["passed"] = (out["all_accepted"] and not out["missing"] and out["deficit"] == 0 and not out["spurious"]
                     and out["orthogonality"] <= 1e-6)
    return out


def evaluate_case(case, result_dir, root=None, tol: float | None = None) -> dict:
    """Judges a result directory against a case (and its reference when one exists) in the
    reference's data layout (tools/eval/evaluate.py:805-840): the case's matrices and target, the
    reference eigenvalues (ref.json 'complete' decides whether unmatched values are spurious), the
    result's eigenvalues and eigenvectors. Returns the verdict document written by main()."""
    from hsolve_amd import case as cio  # the layout reader only; no solver code
    c = cio.load_case(case, root)
    ref = cio.load_ref(c.path if (c.path / "eigenvalues.npy").exists() else c.name, root)
    res = cio.load_result(result_dir)
    target = dict(c.config.get("target", {}))
    rep_target = res.report.get("config", {}).get("target") if isinstance(res.report.get("config"), dict) else None
    if not target and rep_target:
        target = rep_target
    tol = tol if tol is not None else float(c.config.get("tol", 1e-10))
    lam = np.asarray(res.eigenvalues, dtype=float)
    X = np.asarray(res.eigenvectors)
    rv = None if ref is None else np.real(np.asarray(ref.eigenvalues))
    out = evaluate(c.A, c.B, lam, X, target, rv, tol=tol,
                   ref_complete=bool(ref.info.get("complete", False)) if ref is not None else False)
    reasons = []
    if not out["all_accepted"]:
        reasons.append(f"backward error above tol (max {out['max_eta']:.3g} > {tol:g})")
    if out["missing"]:
        reasons.append(f"{len(out['missing'])} required eigenvalue(s) missing")
    if out["deficit"]:
        reasons.append(f"{out['deficit']} tied eigenvalue(s) at the target boundary missing")
    if out["spurious"]:
        reasons.append(f"{len(out['spurious'])} eigenvalue(s) not in the complete reference")
    if out["orthogonality"] > 1e-6:
        reasons.append(f"vectors not B-orthonormal ({out['orthogonality']:.3g})")
    return {"case": c.name, "result": str(result_dir), "n": c.n, "target": target, "tol": tol,
            "reference": None if ref is None else {"kind": ref.info.get("kind"),
                                                   "complete": ref.info.get("complete"),
                                                   "count": int(np.size(ref.eigenvalues))},
            "summary": {"n_pairs": int(lam.size), "max_backward_error": out["max_eta"],
                        "orthogonality": out["orthogonality"]},
            "missing": out["missing"], "spurious": out["spurious"], "deficit": out["deficit"],
            "verdict": "pass" if out["passed"] else "fail", "reasons": reasons}


def main(argv=None) -> int:
    """python -m oracle.evaluate CASE RESULT [--root R] [--out FILE] [--tol T]: writes the verdict
    JSON (default RESULT/evaluation.json) and prints a one-line summary; exit 0 when it passes."""
    import argparse
    import json
    from pathlib import Path
    ap = argparse.ArgumentParser(description="Evaluate a solver result directory against a case.")
    ap.add_argument("case", help="case directory, or a case name under ROOT/cases")
    ap.add_argument("result", type=Path, help="result directory")
    ap.add_argument("--root", help="data root (default: $HS_DATA_ROOT or ./data)")
    ap.add_argument("--out", type=Path, help="output JSON (default: RESULT/evaluation.json)")
    ap.add_argument("--tol", type=float, help="acceptance threshold on eta (default: the case's config tol)")
    args = ap.parse_args(argv)
    doc = evaluate_case(args.case, args.result, args.root, args.tol)
    out = args.out or args.result / "evaluation.json"
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(doc, indent=2) + "\n")
    s = doc["summary"]
    print(f"{doc['case']}: {doc['verdict']}  pairs={s['n_pairs']} max_eta={s['max_backward_error']:.3g} -> {out}")
    for r in doc["reasons"]:
        print(f"  - {r}")
    return 0 if doc["verdict"] == "pass" else 1


if __name__ == "__main__":
    raise SystemExit(main())
