"""CPU restatement of the shift-invert path (oracle; the bench's CPU baseline).

The reference's own binary cannot run here (its LDL^T is MUMPS, Fortran, no gfortran in this
image), so its CPU path is restated with independent components, as the reference's own scipy
baseline does (tools/eval/baseline_scipy.py:109-170):
  * sparse LU of K = A - sigma B with SuperLU (scipy splu, single thread);
  * shift-invert Lanczos in the B inner product (ARPACK eigsh, mode 3, which='LM'), the
    mathematical counterpart of the reference's Krylov-Schur run (src/driver/definite.cpp:213-309);
  * the exact Sylvester certificate of solve_nearest (definite.cpp:630-798): window ends in the
    gaps beyond the selected pairs, nu(b) - nu(a) from sparse LDL^T inertia (oracle/inertia.py)
    compared with the pairs found in the window.
Timed on the host as the baseline; its answers are also checked by the evaluator in the tests.
"""
from __future__ import annotations

import math
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from . import inertia as oinertia


def solve_nearest(A, B, sigma: float, k: int, tol: float = 1e-11, seed: int = 2026) -> dict:
    t0 = time.perf_counter()
    n = A.shape[0]
    Bm = sp.identity(n, format="csc") if B is None else sp.csc_matrix(B)
    K = (sp.csc_matrix(A) - sigma * Bm).tocsc()
    lu = sla.splu(K)
    t_fac = time.perf_counter() - t0
    op = sla.LinearOperator((n, n), matvec=lu.solve, dtype=np.float64)
    Ar = sp.csr_matrix(A)
    Br = None if B is None else sp.csr_matrix(B)
    t_iter = t_cert = 0.0
    nev = k
    for _ in range(3):
        t1 = time.perf_counter()
        v0 = np.random.default_rng(seed).uniform(-1, 1, n)
        w, V = sla.eigsh(Ar, k=nev, M=Br, sigma=sigma, which="LM", OPinv=op, tol=tol, v0=v0,
                         ncv=min(n - 1, max(2 * nev + 1, nev + 16)))
        order = np.argsort(np.abs(w - sigma), kind="stable")
        w, V = w[order], V[:, order]
        t_iter += time.perf_counter() - t1
        # certificate window beyond the k-th nearest, pushed into the gap to the next computed value
        r = float(np.max(np.abs(w[:k] - sigma)))
        beyond = np.abs(w - sigma)[np.abs(w - sigma) > r * (1 + 1e-9)]
        gap = 0.5 * (float(beyond.min()) - r) if beyond.size else max(1e-3 * r, 1e-8 * max(1.0, abs(sigma)))
        a_end, b_end = sigma - r - gap, sigma + r + gap
        t1 = time.perf_counter()
        nu_a = oinertia.sparse_inertia(oinertia.shifted(A, B, a_end))[0]
        nu_b = oinertia.sparse_inertia(oinertia.shifted(A, B, b_end))[0]
        t_cert += time.perf_counter() - t1
        found = int(np.sum((w >= a_end) & (w <= b_end)))
        if nu_b - nu_a == found:
            break
        nev = min(n - 2, nev + (nu_b - nu_a - found) + 2)  # rerun for the missing pairs
    w, V = w[:k], V[:, :k]
    return {"eigenvalues": w, "eigenvectors": V, "count_inertia": nu_b - nu_a, "count_found": found,
            "passed": nu_b - nu_a == found, "time_factor_s": t_fac, "time_iterate_s": t_iter,
            "time_certificate_s": t_cert, "time_total_s": time.perf_counter() - t0, "cores": 1}


def solve_slice(A, B, lower: float, upper: float, tol: float = 1e-11, seed: int = 2026) -> dict:
    """All eigenvalues in [lower, upper] with one shift at the centre: counts at both ends give the
    number wanted, one shift-invert Lanczos run converges them, the counts certify the window
    (one slice of the reference's WindowSolver, definite.cpp:1362-1520)."""
    t0 = time.perf_counter()
    nu_lo = oinertia.sparse_inertia(oinertia.shifted(A, B, lower))[0]
    nu_hi = oinertia.sparse_inertia(oinertia.shifted(A, B, upper))[0]
    want = nu_hi - nu_lo
    t_count = time.perf_counter() - t0
    if want == 0:
        return {"eigenvalues": np.zeros(0), "count_inertia": 0, "count_found": 0, "passed": True,
                "time_total_s": time.perf_counter() - t0, "cores": 1}
    sigma = 0.5 * (lower + upper)
    res = solve_nearest(A, B, sigma, want, tol=tol, seed=seed)
    w = res["eigenvalues"]
    found = int(np.sum((w >= lower) & (w <= upper)))
    res.update({"count_inertia": want, "count_found": found, "passed": found == want,
                "time_certificate_s": res["time_certificate_s"] + t_count, "time_total_s": time.perf_counter() - t0})
    return res


def factor_solve_sample(A, B, sigma: float, perm: np.ndarray, p: int = 8, seed: int = 2026) -> dict:
    """The CPU path's two hot operations on one shifted pencil, timed (the bench's bounded CPU
    sample): the elimination of K = A - sigma B in the fill-reducing order `perm` (new -> old; the
    order the reference hands MUMPS as PERM_IN, include/hsolve/factor.hpp:598-633) with SuperLU
    in symmetric mode and diagonal pivots (the LDL^T-equivalent elimination, one thread), then one
    solve with p right-hand sides (the reference's multi-RHS solve,
    src/factor/mumps_factorization.cpp:458-527). Returns the two times and the LU size."""
    n = A.shape[0]
    Bm = sp.identity(n, format="csc") if B is None else sp.csc_matrix(B)
    perm = np.asarray(perm, dtype=np.int64)
    K = (sp.csc_matrix(A) - sigma * Bm).tocsc()[perm][:, perm].tocsc()
    t0 = time.perf_counter()
    lu = sla.splu(K, permc_spec="NATURAL", diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))
    t_factor = time.perf_counter() - t0
    rhs = np.random.default_rng(seed).standard_normal((n, p))
    t0 = time.perf_counter()
    x = lu.solve(rhs)
    t_solve = time.perf_counter() - t0
    resid = float(np.linalg.norm(K @ x - rhs) / np.linalg.norm(rhs))
    return {"factor_s": t_factor, "solve_s": t_solve, "p": p, "lu_entries": int(lu.L.nnz + lu.U.nnz),
            "relative_residual": resid, "cores": 1}
