"""Pins of the evaluator restatement (oracle/evaluate.py) on dense problems (CPU).

Exact pairs from LAPACK must pass; each plausible defect (a dropped pair, a duplicated pair, a
perturbed eigenvalue or vector, a missing tie-breaking copy) must fail the rule that catches it,
as the reference's tests/py/test_eval.py checks its evaluator.
"""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

from hsolve_amd.gen import yee
from oracle import evaluate as ev
from oracle import yee_exact


@pytest.fixture(scope="module")
def cavity():
    A, _, g = yee.cube_cavity(5)
    w, V = la.eigh(A.toarray())
    ref = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    return A, w, V, ref


def _nearest(w, V, sigma, k):
    idx = np.argsort(np.abs(w - sigma), kind="stable")[:k]
    return w[idx], V[:, idx]


def test_exact_pairs_pass(cavity):
    A, w, V, ref = cavity
    target = {"mode": "nearest", "sigma": 20.0, "k": 8}
    lam, X = _nearest(w, V, 20.0, 8)
    out = ev.evaluate(A, None, lam, X, target, ref)
    assert out["passed"], out
    assert out["max_eta"] < 1e-13


def test_dropped_pair_is_missing(cavity):
    A, w, V, ref = cavity
    target = {"mode": "nearest", "sigma": 20.0, "k": 8}
    lam, X = _nearest(w, V, 20.0, 9)
    keep = list(range(1, 9))  # skip the nearest: the 9th nearest does not replace it
    out = ev.evaluate(A, None, lam[keep], X[:, keep], target, ref)
    assert not out["passed"] and (out["missing"] or out["deficit"])


def test_duplicate_pair_is_caught(cavity):
    A, w, V, ref = cavity
    target = {"mode": "nearest", "sigma": 20.0, "k": 8}
    lam, X = _nearest(w, V, 20.0, 8)
    lam2 = np.concatenate([lam[:7], lam[:1]])
    X2 = np.concatenate([X[:, :7], X[:, :1]], axis=1)
    out = ev.evaluate(A, None, lam2, X2, target, ref)
    assert not out["passed"]


def test_perturbed_vector_fails_backward_error(cavity):
    A, w, V, ref = cavity
    target = {"mode": "nearest", "sigma": 20.0, "k": 4}
    lam, X = _nearest(w, V, 20.0, 4)
    X = X.copy()
    X[:, 0] += 1e-6 * np.random.default_rng(0).standard_normal(X.shape[0])
    X[:, 0] /= np.linalg.norm(X[:, 0])
    out = ev.evaluate(A, None, lam, X, target, ref)
    assert not out["all_accepted"]


def test_interval_expected_set_and_ties():
    ref = np.array([1.0, 2.0, 2.0, 3.0, 4.0])
    exp = ev.expected_set(ref, {"mode": "interval", "lower": 1.5, "upper": 3.5}, 1.0, 1e-9)
    assert sorted(exp.required.tolist()) == [1, 2, 3]
    exp = ev.expected_set(ref, {"mode": "nearest", "sigma": 2.1, "k": 2}, 1.0, 1e-9)
    assert sorted(exp.required.tolist()) == [1, 2]
    exp = ev.expected_set(ref, {"mode": "nearest", "sigma": 2.9, "k": 2}, 1.0, 1e-9)
    # 3.0 is required, one of the tied 2.0 copies completes the set
    assert exp.required.tolist() == [3] and exp.need_optional == 1 and sorted(exp.optional.tolist()) == [1, 2]


def test_generalized_backward_error_uses_b():
    A, B, g = yee.jittered_box(cells=(4, 3, 3), lengths=(2.0, 1.0, 1.5), seed=2)
    w, V = la.eigh(A.toarray(), B.toarray())
    target = {"mode": "nearest", "sigma": 30.0, "k": 5}
    idx = np.argsort(np.abs(w - 30.0))[:5]
    out = ev.evaluate(A, B, w[idx], V[:, idx], target, w)
    assert out["passed"], out
