"""Sylvester inertia counts independent of the GPU engine (oracle).

nu(K) = number of negative eigenvalues of K = A - s B, the count the reference reads from MUMPS'
INFOG(12) (src/factor/mumps_factorization.cpp:358-381, contract include/hsolve/factor.hpp:295-300)
and uses in count_below (src/driver/steps.cpp:549-568).

* dense_inertia: eigenvalues of the dense K (LAPACK); exact for small n.
* sparse_inertia: SuperLU (scipy) in symmetric mode with diagonal pivots only, so P K P^T = L U
  with U = D L^T; the signs of diag(U) give the inertia by Sylvester's law. An independent sparse
  elimination (SuperLU's own ordering and supernodes), usable at 1e4..1e5 unknowns.
Pinned by tests/test_oracle_yee.py::test_sparse_inertia_matches_dense_and_closed_form: both
agree with each other and with the closed-form Yee spectra (oracle/yee_exact.py) at gap points.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as la
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def dense_inertia(K) -> tuple[int, int, int]:
    Kd = K.toarray() if sp.issparse(K) else np.asarray(K)
    w = la.eigvalsh(Kd)
    scale = max(1.0, float(np.max(np.abs(w)))) if w.size else 1.0
    tol = 1e-12 * scale
    return int(np.sum(w < -tol)), int(np.sum(np.abs(w) <= tol)), int(np.sum(w > tol))


def sparse_inertia(K) -> tuple[int, int, int]:
    Kc = sp.csc_matrix(K)
    lu = sla.splu(Kc, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                  options={"SymmetricMode": True})
    if not np.array_equal(lu.perm_r, lu.perm_c):
        raise RuntimeError("SuperLU left the diagonal: inertia not available from this elimination")
    d = lu.U.diagonal()
    return int(np.sum(d < 0)), int(np.sum(d == 0)), int(np.sum(d > 0))


def shifted(A, B, s: float):
    n = A.shape[0]
    Bm = sp.identity(n, format="csr") if B is None else sp.csr_matrix(B)
    return (sp.csr_matrix(A) - s * Bm).tocsr()
