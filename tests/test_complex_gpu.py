"""Complex Hermitian pencils on the GPU: Bloch-periodic Yee cells (complex phases on the wrapped
edges), factorized through the real embedding, Krylov in complex arithmetic.

Checks against closed forms and dense solvers:
  * inertia of A - s B (homogeneous cell) = the count of the exact discrete Bloch spectrum below s
    (tools/gen/yee.py:231-260 restated in oracle/yee_exact.py) at gap points;
  * a complex solve K z = b has a componentwise-small residual;
  * interval and nearest targets of a homogeneous cell match the exact spectrum, of a photonic
    cell (eps = 12 inside a sphere) the dense generalized eigenvalues; every pair passes the
    evaluator (backward error <= 1e-10 from the original matrices, B-orthonormal vectors).
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

from hsolve_amd.gen import yee
from oracle import evaluate as ev
from oracle import yee_exact

pytestmark = pytest.mark.gpu

K_GENERIC = (0.7, 1.3, 2.1)   # phases k_d a (radians) with no symmetry: A genuinely complex


def _homogeneous(cells):
    g = yee.YeeGrid((cells, cells, cells), (1.0, 1.0, 1.0), "bloch", K_GENERIC)
    A, B = yee.assemble(g)
    return A, B, g


def test_bloch_inertia_matches_exact_spectrum(gpu):
    from hsolve_amd.factor import GpuFactorization, Pencil
    A, B, g = _homogeneous(6)
    assert np.iscomplexobj(A.data) and np.any(A.data.imag != 0)
    ref = yee_exact.bloch_box_spectrum(g.cells, g.lengths, g.phases)
    F = GpuFactorization(Pencil(A, B))
    pts = yee_exact.gap_points(ref, 1.0, 0.6 * ref[-1])
    for s in pts[:: max(1, len(pts) // 12)]:
        assert F.factorize(1.0, -float(s), precision="double"), s
        inn = F.inertia()
        print("s", s, "neg", inn.neg, "exact", yee_exact.count_below(ref, s), flush=True)
        assert inn.neg == yee_exact.count_below(ref, s)
        assert inn.neg + inn.zero + inn.pos == A.shape[0]


@pytest.mark.parametrize("precision", ["double", "single"])
def test_complex_solve_residual(gpu, precision):
    import torch
    from hsolve_amd.factor import GpuFactorization, Pencil
    A, B, g = yee.bloch_cell(6, K_GENERIC)
    sigma = 13.7
    F = GpuFactorization(Pencil(A, B))
    assert F.factorize(1.0, -sigma, precision=precision)
    rng = np.random.default_rng(1)
    n = A.shape[0]
    b = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    x = torch.from_numpy(b.copy()).to(gpu)
    F.solve(x)
    z = x.cpu().numpy()
    K = (A - sigma * B).tocsr()
    r = (K @ z.T).T - b
    den = (abs(K) @ np.abs(z).T).T + np.abs(b)
    omega = np.max(np.abs(r) / den)
    print("omega", omega, flush=True)
    assert omega <= (1e-12 if precision == "double" else 1e-4)


def test_bloch_interval_homogeneous_vs_exact(gpu):
    import hsolve_amd
    A, B, g = _homogeneous(6)
    ref = yee_exact.bloch_box_spectrum(g.cells, g.lengths, g.phases)
    lo, hi = 20.0, 90.0
    res = hsolve_amd.solve(A, B, lower=lo, upper=hi, config={"slice_size": 24})
    out = ev.evaluate(A, B, res.eigenvalues, res.eigenvectors, {"mode": "interval", "lower": lo, "upper": hi}, ref)
    print("eval", {k: out[k] for k in ("max_eta", "missing", "spurious", "orthogonality")},
          "pairs", res.eigenvalues.size, "slices", len(res.report["slices"]), flush=True)
    assert out["passed"], out
    assert res.success and res.report["certificate"]["passed"] and res.report["complex_scalars"]
    assert res.eigenvectors.dtype == np.complex128
    assert res.report["certificate"]["count_inertia"] == (yee_exact.count_below(ref, hi) -
                                                         yee_exact.count_below(ref, lo))


def test_bloch_photonic_nearest_vs_dense(gpu):
    import hsolve_amd
    A, B, g = yee.bloch_cell(5, K_GENERIC)
    w = la.eigvalsh(A.toarray(), B.toarray())
    sigma, k = 25.0, 12
    res = hsolve_amd.solve(A, B, sigma=sigma, k=k)
    out = ev.evaluate(A, B, res.eigenvalues, res.eigenvectors, {"mode": "nearest", "sigma": sigma, "k": k}, w)
    print("eval", {k_: out[k_] for k_ in ("max_eta", "missing", "spurious", "orthogonality")}, flush=True)
    assert out["passed"], out
    assert res.success and res.report["certificate"]["passed"]


def test_bloch_concurrent_slices_repeatable(gpu):
    # photonic cell, window on three streams twice: bitwise the same pairs
    from hsolve_amd import solver
    A, B, g = yee.bloch_cell(6, K_GENERIC)
    cfg = {"target": {"mode": "interval", "lower": 10.0, "upper": 60.0}, "tol": 1e-10, "slice_size": 16}
    l1, v1, r1, _ = solver.solve(A, B, cfg, streams=3)
    l2, v2, r2, _ = solver.solve(A, B, cfg, streams=3)
    print("pairs", l1.size, "slices", len(r1["slices"]), flush=True)
    assert r1["certificate"]["passed"] and len(r1["slices"]) >= 2
    assert np.array_equal(l1, l2) and np.array_equal(v1, v2)
    w = la.eigvalsh(A.toarray(), B.toarray())
    out = ev.evaluate(A, B, l1, v1, cfg["target"], w)
    assert out["passed"], out
