"""Subspace-angle bounds (hsolve_amd/subspace.py) against the mathematics, on the host: a fake
device layer runs the same code on CPU tensors. For small dense pencils with known eigenvectors,
perturbed vectors get bounds that are never below the true largest principal angle (B inner
product) between each group's span and its exact eigenspace, and are within a modest factor of
it; clusters of equal eigenvalues are grouped; a window that holds more eigenvalues than pairs
gives a single "window" group."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.linalg as la
import torch

from hsolve_amd import subspace


class FakeOps:
    """The DeviceOps calls subspace.py makes, on CPU tensors (rows of x are vectors)."""

    def __init__(self, A, B):
        self.A = torch.from_numpy(A)
        self.Bd = None if B is None else torch.from_numpy(np.diag(B).copy())
        self.n = A.shape[0]
        self.pencil = SimpleNamespace(b_is_identity=B is None, b_is_diagonal=B is not None)
        self.d = {"b_diag": self.Bd}

    def apply_a(self, x):
        return x @ self.A.T

    def apply_b(self, x):
        return x if self.Bd is None else x * self.Bd[None, :]

    def dots(self, x, y, weight=None, out=None):
        w = 1.0 if weight is None else weight[None, :]
        return (x * y * w).sum(dim=1)

    def b_dots(self, x, y, by=None, out=None):
        return self.dots(x, self.apply_b(y))

    def b_gram(self, v, w, bw=None, out=None):
        return (self.apply_b(w) @ v.T)  # (p, k): C = V^T B W column-major


def _pencil(n, seed, mult=()):
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(1.0, 10.0, n))
    for i, k in mult:  # force k-fold eigenvalues starting at i
        lam[i:i + k] = lam[i]
    lam = np.sort(lam)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    b = rng.uniform(0.5, 2.0, n)
    Bh = np.diag(np.sqrt(b))
    # A = B^{1/2} Q diag(lam) Q^T B^{1/2}: generalized eigenvectors X = B^{-1/2} Q
    A = Bh @ Q @ np.diag(lam) @ Q.T @ Bh
    A = 0.5 * (A + A.T)
    X = np.diag(1.0 / np.sqrt(b)) @ Q
    return A, np.diag(b), lam, X


def _true_sin(Xg, E, B):
    """largest sine of the principal angles between span Xg and span E in the B inner product"""
    Bh = np.sqrt(np.diag(B))[:, None]
    qa, _ = np.linalg.qr(Bh * Xg)
    qb, _ = np.linalg.qr(Bh * E)
    s = la.svdvals(qa - qb @ (qb.T @ qa))
    return float(s.max()) if s.size else 0.0


@pytest.mark.parametrize("noise", [1e-9, 1e-6])
def test_bound_dominates_true_angle_for_separated_pairs(noise):
    A, B, lam, X = _pencil(60, 1)
    rng = np.random.default_rng(2)
    sel = np.arange(20, 30)
    Xp = X[:, sel] + noise * rng.standard_normal((60, sel.size))
    # B-orthonormalize and Rayleigh-Ritz, as the solver returns pairs
    G = Xp.T @ B @ Xp
    L = np.linalg.cholesky(G)
    Xp = Xp @ np.linalg.inv(L).T
    th, S = np.linalg.eigh(Xp.T @ A @ Xp)
    Xp = Xp @ S
    lo, hi = 0.5 * (lam[19] + lam[20]), 0.5 * (lam[29] + lam[30])
    ops = FakeOps(A, B)
    rep, per = subspace.bounds(ops, torch.from_numpy(np.ascontiguousarray(Xp.T)), th, [(lo, hi, 10)], 10.0)
    assert rep["performed"] and all(v is not None for v in per)
    for g in rep["groups"]:
        idx = g["pairs"]
        E = X[:, sel[idx]] if g["kind"] == "cluster" else X[:, sel]
        true = _true_sin(Xp[:, idx], E, B)
        assert g["sin_bound"] >= true * (1 - 1e-6), (g, true)
        assert g["sin_bound"] <= 1e4 * max(true, 1e-14) + 1e-8
    assert rep["max_sin_bound"] == max(per)


def test_multiple_eigenvalue_forms_one_group_and_window_group_when_cut():
    A, B, lam, X = _pencil(50, 3, mult=((10, 3),))
    ev, V = la.eigh(A, B)
    i0 = int(np.argmin(np.abs(ev - lam[10])))
    triple = np.flatnonzero(np.abs(ev - ev[i0]) < 1e-9)
    assert triple.size == 3
    sel = np.arange(triple[0] - 2, triple[-1] + 3)
    lo, hi = 0.5 * (ev[sel[0] - 1] + ev[sel[0]]), 0.5 * (ev[sel[-1]] + ev[sel[-1] + 1])
    ops = FakeOps(A, B)
    x = torch.from_numpy(np.ascontiguousarray(V[:, sel].T))
    rep, per = subspace.bounds(ops, x, ev[sel], [(lo, hi, sel.size)], 10.0)
    sizes = sorted(len(g["pairs"]) for g in rep["groups"])
    assert 3 in sizes and all(g["kind"] == "cluster" for g in rep["groups"])
    assert rep["max_sin_bound"] < 1e-8
    # one copy of the triple not returned: the window holds one eigenvalue more than the pairs
    keep = [j for j in range(sel.size) if sel[j] != triple[-1]]
    rep2, per2 = subspace.bounds(ops, x[keep].contiguous(), ev[sel][keep], [(lo, hi, sel.size)], 10.0)
    assert [g["kind"] for g in rep2["groups"]] == ["window"]
    assert rep2["groups"][0]["dimension"] == sel.size and rep2["groups"][0]["unreturned"] == 1
    true = _true_sin(V[:, sel[keep]], V[:, sel], B)
    assert per2[0] >= true and per2[0] < 1e-8


def test_no_bound_without_a_gap():
    A, B, lam, X = _pencil(30, 5)
    ops = FakeOps(A, B)
    sel = np.arange(5, 10)
    # window end placed on an eigenvalue: zero gap to the outside, no bound
    rep, per = subspace.bounds(ops, torch.from_numpy(np.ascontiguousarray(X[:, sel].T)), lam[sel],
                               [(lam[5], lam[10], 5)], 10.0)
    assert any(v is None for v in per) or rep["max_sin_bound"] is None or rep["max_sin_bound"] == 1.0
