"""The million-unknown PEC box (80 x 70 x 64 Yee cells, n = 1,045,014) on one GPU: the factorization
fits because contribution blocks share storage across levels (their lifetimes run from the
front's level to its parent's), 14 GB for the blocks instead of 275 GB with one block per front.

* inertia of the unjittered FIT pencil equal to the closed-form count (oracle/yee_exact.py) at three
  mid-gap shifts, low, middle and near the shift of the nearest solve (the reference's counts at
  gap points, tests/cpp/test_factor.cpp:198-265, at the size of its million-unknown case,
  README.en.md:116);
* the nearest 20 eigenvalues at a mid-spectrum shift, judged by the evaluator against the closed
  form (eta <= 1e-10 from the original matrices, no missed or spurious eigenvalue), certificate
  counts equal to the closed form.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from hsolve_amd.gen import yee
from oracle import evaluate as ev
from oracle import yee_exact

pytestmark = pytest.mark.gpu
CELLS = (80, 70, 64)
LENGTHS = (80 / 64 * math.pi, 70 / 64 * math.pi, math.pi)
# a quarter of the spectrum's top 12 / h^2, moved to the middle of the gap above it: 0.25 * 12 / h^2
# is itself an eigenvalue of this box (a shift on an eigenvalue is perturbed by 1e-8 relative,
# which leaves K too ill-conditioned for the FP64 solves to resolve the other 19 pairs)
TOP_QUARTER = 0.25 * 12.0 / (math.pi / 64) ** 2


@pytest.fixture(scope="module")
def box80():
    A, B, g = yee.jittered_box(CELLS, LENGTHS, seed=2026, jitter=0.0)
    assert A.shape[0] == 1045014
    ref = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    return A, B, ref, float(yee_exact.gap_points(ref, TOP_QUARTER, TOP_QUARTER + 0.1)[0])


def test_box80_inertia_matches_closed_form(gpu, box80):
    from hsolve_amd.factor import GpuFactorization, Pencil
    A, B, ref, sigma = box80
    F = GpuFactorization(Pencil(A, B))
    info = F.info
    print(f"panels {info.panel_bytes_double / 1e9:.2f} GB, contribution blocks {info.cb_bytes_double / 1e9:.2f} GB "
          f"(one per front: {info.cb_bytes_unshared / 1e9:.1f} GB), flops {info.flops:.3g}", flush=True)
    assert info.cb_bytes_double < 0.1 * info.cb_bytes_unshared
    low = yee_exact.gap_points(ref, 40.0, 41.0)[0]
    mid = yee_exact.gap_points(ref, 600.0, 601.0)[0]
    near = yee_exact.gap_points(ref, sigma - 0.5, sigma + 0.5)[0]
    for s in (low, mid, near):
        assert F.factorize(1.0, -float(s), precision="double")
        inn = F.inertia()
        print(f"s={s:.10g} neg={inn.neg} closed form={yee_exact.count_below(ref, s)}", flush=True)
        assert inn.neg == yee_exact.count_below(ref, float(s)) and inn.zero == 0
    F.close()


def test_box80_nearest_k20_matches_closed_form(gpu, box80):
    import hsolve_amd
    A, B, ref, sigma = box80
    target = {"mode": "nearest", "sigma": sigma, "k": 20}
    res = hsolve_amd.solve(A, B, sigma=sigma, k=20)
    out = ev.evaluate(A, B, res.eigenvalues, res.eigenvectors, target, ref)
    cert = res.report["certificate"]
    print("eval", {k: out[k] for k in ("max_eta", "missing", "deficit", "spurious", "orthogonality")},
          "cert", {k: cert[k] for k in ("count_inertia", "count_found", "passed")},
          "time", round(res.report["time_total_s"], 2), flush=True)
    assert out["passed"], out
    assert res.success and cert["passed"]
    assert cert["count_inertia"] == yee_exact.count_below(ref, cert["upper"]) - yee_exact.count_below(ref, cert["lower"])
