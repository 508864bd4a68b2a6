"""The real embedding of a complex Hermitian pencil (host side of the complex factorization path).

E is rebuilt densely from the pattern and the value maps the factorization receives (f_rowptr,
f_col, f_a_index, f_b_index over the extended value arrays [Re v, Im v, -Im v]) and checked
against what the mathematics fixes:
  * E equals the interleaved real form of K = alpha A + beta B entry by entry
    (E(2i, 2j) = E(2i+1, 2j+1) = Re K, E(2i, 2j+1) = -Im K, E(2i+1, 2j) = Im K);
  * E acting on the real view of a complex vector is K acting on the vector;
  * the spectrum of E is that of K with every eigenvalue twice, so inertia(E) = 2 inertia(K);
  * entries whose real (off the diagonal: imaginary) parts vanish in A and B are left out.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from hsolve_amd.factor import Pencil


def _hermitian(rng, n, density=0.15):
    re = sp.random(n, n, density=density, random_state=rng, format="csr")
    im = sp.random(n, n, density=density, random_state=rng, format="csr")
    # some couplings purely real, some purely imaginary: both kinds of pruned entries occur
    m = re + 1j * im
    m = (m + m.conj().T) * 0.5
    m = sp.csr_matrix(m + sp.identity(n) * 2.0)
    m.sort_indices()
    return m


def _dense_embedding(P: Pencil, alpha: float, beta: float) -> np.ndarray:
    a = P.A.data
    b = P._Bfull.data
    a_ext = np.concatenate((a.real, a.imag, -a.imag))
    b_ext = np.concatenate((b.real, b.imag, -b.imag))
    n2 = P.f_n
    E = np.zeros((n2, n2))
    for i in range(n2):
        for e in range(P.f_rowptr[i], P.f_rowptr[i + 1]):
            v = 0.0
            if P.f_a_index[e] >= 0:
                v += alpha * a_ext[P.f_a_index[e]]
            if P.f_b_index[e] >= 0:
                v += beta * b_ext[P.f_b_index[e]]
            E[i, P.f_col[e]] += v
    return E


def _interleaved(K: np.ndarray) -> np.ndarray:
    n = K.shape[0]
    E = np.zeros((2 * n, 2 * n))
    E[0::2, 0::2] = K.real
    E[1::2, 1::2] = K.real
    E[0::2, 1::2] = -K.imag
    E[1::2, 0::2] = K.imag
    return E


def test_embedding_matches_interleaved_real_form():
    rng = np.random.default_rng(3)
    n = 37
    A = _hermitian(rng, n)
    for B in (None, sp.diags(rng.uniform(0.5, 2.0, n), format="csr"), _hermitian(rng, n, 0.05)):
        P = Pencil(A, B)
        assert P.is_complex and P.f_n == 2 * n
        alpha, beta = 1.0, -0.37
        E = _dense_embedding(P, alpha, beta)
        K = (alpha * P.A + beta * P._Bfull).toarray()
        np.testing.assert_array_equal(E, _interleaved(K))
        assert np.array_equal(E, E.T)
        # E on the real view of z is K z
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        ez = E @ z.view(np.float64)
        np.testing.assert_allclose(ez.view(np.complex128), K @ z, rtol=1e-13, atol=1e-13)
        # every eigenvalue of K twice
        wk = np.linalg.eigvalsh(K)
        we = np.linalg.eigvalsh(E)
        np.testing.assert_allclose(we, np.repeat(wk, 2), rtol=1e-12, atol=1e-12)
        assert np.count_nonzero(we < 0) == 2 * np.count_nonzero(wk < 0)


def test_embedding_prunes_structural_zeros():
    # a tridiagonal K with real couplings and one purely imaginary coupling: no -Im/+Im entries for
    # the real couplings, no Re entries for the imaginary one
    n = 6
    K = sp.diags([np.full(n - 1, 1.5), np.full(n, 4.0), np.full(n - 1, 1.5)], [-1, 0, 1], format="lil",
                 dtype=np.complex128)
    K[2, 3], K[3, 2] = 0.7j, -0.7j
    P = Pencil(sp.csr_matrix(K))
    nnz_real_offdiag = 2 * (n - 1) - 2
    # Re entries: diagonal n + real off-diagonals, twice; Im entries: the one coupling both ways, twice
    assert P.f_col.size == 2 * (n + nnz_real_offdiag) + 2 * 2
    E = _dense_embedding(P, 1.0, 0.0)
    np.testing.assert_array_equal(E, _interleaved(K.toarray()))


def test_real_pencil_pattern_is_the_union_pattern():
    rng = np.random.default_rng(4)
    A = _hermitian(rng, 20).real
    P = Pencil(sp.csr_matrix(A))
    assert not P.is_complex and P.f_n == 20
    assert P.f_rowptr is P.u_rowptr and P.f_col is P.u_col
