"""Subspace-angle bounds of GPU solves (report.subspace, pairs[j].subspace_sin_bound), against
dense generalized eigenvectors: every group's bound is at least the true largest principal sine
(B inner product) between the returned span and the exact eigenspace, and small for converged
pairs; multiple eigenvalues of a cavity are bounded as groups."""
import numpy as np
import pytest
import scipy.linalg as la

from hsolve_amd.gen import yee

pytestmark = pytest.mark.gpu


def _true_sin(Xg, E, Bd):
    Bh = np.sqrt(Bd)[:, None]
    qa, _ = np.linalg.qr(Bh * Xg)
    qb, _ = np.linalg.qr(Bh * E)
    return float(la.svdvals(qa - qb @ (qb.conj().T @ qa)).max())


@pytest.mark.parametrize("case", ["ghep", "cavity"])
def test_subspace_bounds_dominate_true_angles(gpu, case):
    import hsolve_amd
    if case == "ghep":
        A, B, g = yee.jittered_box(cells=(8, 5, 6), lengths=(2 * np.pi, np.pi, 1.5 * np.pi), seed=4)
        sigma, k = 12.0, 10
        Bd = B.diagonal()
    else:
        A, B, g = yee.cube_cavity(6)
        sigma, k = 9.0, 16
        Bd = np.ones(A.shape[0])
    w, V = la.eigh(A.toarray(), None if B is None else B.toarray())
    res = hsolve_amd.solve(A, B, sigma=sigma, k=k)
    sub = res.report["subspace"]
    print({k_: sub[k_] for k_ in ("performed", "max_sin_bound", "rigorous", "note")},
          [(gr["kind"], len(gr["pairs"]), gr["sin_bound"]) for gr in sub["groups"]], flush=True)
    assert sub["performed"] and sub["max_sin_bound"] is not None and sub["max_sin_bound"] < 1e-6
    assert all(p["subspace_sin_bound"] is not None for p in res.report["pairs"])
    X = res.eigenvectors
    lo, hi = res.report["certificate"]["lower"], res.report["certificate"]["upper"]
    for gr in sub["groups"]:
        idx = gr["pairs"]
        if gr["kind"] == "cluster":
            ex = np.flatnonzero((w >= gr["lower"]) & (w <= gr["upper"]))
        else:
            ex = np.flatnonzero((w >= lo) & (w <= hi))
        assert ex.size == gr["dimension"]
        true = _true_sin(X[:, idx], V[:, ex], Bd)
        assert gr["sin_bound"] >= true, (gr, true)
    if case == "cavity":
        assert any(len(gr["pairs"]) >= 2 for gr in sub["groups"])  # multiple eigenvalues grouped
