"""Single-precision working factors with FP64 iterative refinement, on the GPU.

Mirrors the reference's RefinementOfSinglePrecisionFactorReachesDoubleAccuracy
(tests/cpp/test_ops.cpp:164-225): a single-precision factor alone gives K^-1 B x only to about
eps_single * cond(K); FP64 refinement against the exact A and B (ShiftInvertOperator,
src/ops/shift_invert_operator.cpp:83-177) restores double accuracy, one block solve per
correction; a loose tolerance stops earlier. The reference value is the operator applied with an
FP64 factor. Then whole solves with factor_precision = "single": the single-precision attempt
must be the result (no fallback to double), every pair certified by the evaluator.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg as la

from hsolve_amd.gen import yee
from oracle import evaluate as ev
from oracle import yee_exact

pytestmark = pytest.mark.gpu

K_GENERIC = (0.7, 1.3, 2.1)


def _max_rel_diff(y, z):
    """max over columns of ||y_c - z_c||_inf / ||z_c||_inf (rows of the (p, n) blocks)."""
    return float(np.max(np.max(np.abs(y - z), axis=1) / np.max(np.abs(z), axis=1)))


def _cases():
    A, B, _ = yee.jittered_box(cells=(12, 8, 10), lengths=(2 * np.pi, np.pi, 1.5 * np.pi), seed=3)
    yield "ghep", A, B, 17.3
    A, _, _ = yee.cube_cavity(10)
    yield "standard", A, None, 9.4


@pytest.mark.parametrize("case", ["ghep", "standard"])
def test_refinement_of_single_factor_reaches_double_accuracy(gpu, case):
    import torch
    from hsolve_amd.factor import GpuFactorization, Pencil
    from hsolve_amd.ops import DeviceOps, ShiftInvertOperator, SolveRefinement
    name, A, B, sigma = next(c for c in _cases() if c[0] == case)
    pencil = Pencil(A, B)
    ops = DeviceOps(pencil)
    F64 = GpuFactorization(pencil)
    assert F64.factorize(1.0, -sigma, precision="double")
    F32 = F64.clone()
    assert F32.factorize(1.0, -sigma, precision="single")
    assert F32.factored_precision == "single"
    n = A.shape[0]
    x = torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, (5, n))).to(gpu)
    z = ShiftInvertOperator(F64, ops, sigma).apply(x.clone()).cpu().numpy()

    plain = ShiftInvertOperator(F32, ops, sigma).apply(x.clone()).cpu().numpy()
    plain_err = _max_rel_diff(plain, z)
    print(case, "single alone", plain_err, flush=True)
    assert 1e-9 < plain_err < 1e-2  # single precision alone is not double-accurate

    refine = SolveRefinement(max_steps=10, rel_tol=1e-14)
    op = ShiftInvertOperator(F32, ops, sigma, refine)
    before = 0
    for p in (1, 3, 5):
        y = op.apply(x[:p].clone()).cpu().numpy()
        err = _max_rel_diff(y, z[:p])
        steps = op.refinement_steps - before
        before = op.refinement_steps
        print(case, "p", p, "refined", err, "rel residual", op.last_rel_residual, "steps", steps, flush=True)
        assert err <= 1e-11
        assert op.last_rel_residual <= 1e-13 or op.last_backward_error <= 8 * np.finfo(float).eps
        assert 1 <= steps <= refine.max_steps
    assert op.applications == 1 + 3 + 5

    loose = ShiftInvertOperator(F32, ops, sigma, SolveRefinement(max_steps=10, rel_tol=1e-8))
    loose.apply(x.clone())
    tight = ShiftInvertOperator(F32, ops, sigma, refine)
    tight.apply(x.clone())
    assert loose.last_rel_residual <= 1e-8
    assert loose.refinement_steps < tight.refinement_steps
    F32.close()
    F64.close()


def _check_single(res, A, B, target, ref):
    prec = res.report["precision"]
    print("precision", {k: prec[k] for k in ("requested", "used", "fallback")}, flush=True)
    assert prec["used"] == "single" and not prec["fallback"]
    assert res.report["factor_options"]["precision"] == "single"
    out = ev.evaluate(A, B, res.eigenvalues, res.eigenvectors, target, ref)
    print("eval", {k: out[k] for k in ("max_eta", "missing", "spurious", "orthogonality")}, flush=True)
    assert out["passed"], out
    assert res.success and res.report["certificate"]["passed"]
    assert out["max_eta"] <= 1e-10


def test_single_precision_nearest_ghep_no_fallback(gpu):
    import hsolve_amd
    A, B, g = yee.jittered_box(cells=(8, 5, 6), lengths=(2 * np.pi, np.pi, 1.5 * np.pi), seed=4)
    w = la.eigvalsh(A.toarray(), B.toarray())
    res = hsolve_amd.solve(A, B, sigma=12.0, k=10, config={"factor_precision": "single"})
    _check_single(res, A, B, {"mode": "nearest", "sigma": 12.0, "k": 10}, w)


def test_single_precision_interval_ghep_no_fallback(gpu):
    import hsolve_amd
    A, B, g = yee.jittered_box(cells=(10, 6, 8), lengths=(2 * np.pi, np.pi, 1.5 * np.pi), seed=5)
    w = la.eigvalsh(A.toarray(), B.toarray())
    res = hsolve_amd.solve(A, B, lower=10.0, upper=40.0, config={"factor_precision": "single", "slice_size": 24})
    _check_single(res, A, B, {"mode": "interval", "lower": 10.0, "upper": 40.0}, w)
    assert len(res.report["slices"]) >= 2


def test_single_precision_bloch_cell_no_fallback(gpu):
    # complex Hermitian pencil: the working factor is the single-precision LDL^T of the embedding
    # (complex64 accuracy), refined in complex128 arithmetic
    import hsolve_amd
    g = yee.YeeGrid((6, 6, 6), (1.0, 1.0, 1.0), "bloch", K_GENERIC)
    A, B = yee.assemble(g)
    ref = yee_exact.bloch_box_spectrum(g.cells, g.lengths, g.phases)
    res = hsolve_amd.solve(A, B, lower=20.0, upper=90.0, config={"factor_precision": "single", "slice_size": 24})
    _check_single(res, A, B, {"mode": "interval", "lower": 20.0, "upper": 90.0}, ref)
