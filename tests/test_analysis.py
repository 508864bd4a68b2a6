"""Host symbolic analysis (CPU): ordering and supernodal fronts.

Replaces MUMPS' analysis (src/factor/mumps_factorization.cpp:188-283). Checks, independent of the
engine: the order is a permutation; every nonzero of the true Cholesky factor of an SPD matrix
with the pencil's pattern (dense LAPACK factorization, the oracle) lies inside a front; the tree is
postordered; factor entries and flops agree with a direct count from the fronts.
"""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

from hsolve_amd.gen import yee


def _analysis(A, B=None, **kw):
    from hsolve_amd.factor import GpuFactorization, Pencil
    return GpuFactorization(Pencil(A, B), **kw)


def _spd_with_pattern(A):
    S = sp.csr_matrix(A, dtype=float).copy()
    S.data = np.random.default_rng(0).uniform(-1, 1, S.nnz)
    S = S + S.T
    d = np.asarray(abs(S).sum(axis=1)).ravel() + 1.0
    return (S + sp.diags(d)).toarray()


@pytest.mark.parametrize("cells,kw", [(4, {}), (5, {"leaf_size": 8}), (6, {"leaf_size": 16})])
def test_fronts_cover_cholesky_pattern(cells, kw):
    A, _, g = yee.cube_cavity(cells)
    F = _analysis(A, **kw)
    order = F.order()
    n = A.shape[0]
    assert np.array_equal(np.sort(order), np.arange(n))
    fr = F.fronts()
    M = _spd_with_pattern(A)[np.ix_(order, order)]
    L = la.cholesky(M, lower=True)
    nz = np.abs(L) > 1e-14 * np.abs(L).max()
    covered = np.zeros((n, n), dtype=bool)
    for s in range(fr["first"].size):
        rows = fr["rows"][fr["rowoff"][s]:fr["rowoff"][s + 1]]
        f, c = fr["first"][s], fr["ncol"][s]
        assert np.array_equal(rows[:c], np.arange(f, f + c))
        assert np.all(np.diff(rows) > 0)
        for j in range(f, f + c):
            covered[rows[rows >= j], j] = True
    assert not np.any(nz & ~covered)
    # postorder: parents after children; the columns partition 0..n-1
    assert np.all((fr["parent"] < 0) | (fr["parent"] > np.arange(fr["parent"].size)))
    assert fr["ncol"].sum() == n and fr["first"][0] == 0


def test_counts_agree_with_fronts():
    A, B, g = yee.jittered_box(cells=(8, 6, 7), lengths=(2.0, 1.0, 1.5), seed=1)
    F = _analysis(A, B)
    fr = F.fronts()
    m = fr["m"].astype(float)
    c = fr["ncol"].astype(float)
    entries = np.sum(m * c - c * (c - 1) / 2)
    cb = m - c
    flops = np.sum(c ** 3 / 3 + c ** 2 * cb + c * cb ** 2)
    assert F.info.factor_entries == pytest.approx(entries)
    assert F.info.flops == pytest.approx(flops)


def test_nested_dissection_beats_natural_order():
    # fill of the own ordering against the natural (Yee block) order on a 10^3 cavity
    A, _, g = yee.cube_cavity(10)
    nd = _analysis(A)
    nat = _analysis(A, order=np.arange(A.shape[0], dtype=np.int32))
    assert nd.info.factor_entries < 0.6 * nat.info.factor_entries


def test_given_order_is_kept():
    A, _, g = yee.cube_cavity(5)
    rng = np.random.default_rng(2)
    perm = rng.permutation(A.shape[0]).astype(np.int32)
    F = _analysis(A, order=perm)
    o = F.order()
    assert np.array_equal(np.sort(o), np.arange(A.shape[0]))


def test_invalid_pattern_rejected():
    from hsolve_amd.errors import InvalidInputError
    A = sp.csr_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))
    A.indices[1] = 7  # out of range
    with pytest.raises(Exception):
        _analysis(A)
