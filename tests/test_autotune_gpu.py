"""The autotuner on the GPU: every candidate plan of a pattern is measured, the cheapest is cached
and used by the next solve of that pattern; whatever the plan, the inertia is the same and the
solve is certified against the closed form."""
import numpy as np
import pytest

from hsolve_amd.gen import yee
from oracle import evaluate as ev
from oracle import yee_exact

pytestmark = pytest.mark.gpu


def test_tune_measures_candidates_and_the_solve_uses_the_plan(gpu, tmp_path, monkeypatch):
    import hsolve_amd
    from hsolve_amd import autotune
    from hsolve_amd.factor import GpuFactorization, Pencil
    monkeypatch.setenv("HSG_TUNE_CACHE", str(tmp_path / "tuning.json"))
    monkeypatch.setattr(autotune, "_memory", {})
    A, _, g = yee.cube_cavity(10)
    w = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    r = autotune.tune(A, None, sigma=9.0, reps=2)
    print("chosen", r["options"], "cost", r["cost"], "default", r["default_cost"], flush=True)
    assert len(r["table"]) >= 8 and r["cost"] <= min(c["cost"] for c in r["table"])
    # every plan eliminates the same matrix: identical inertia at a gap point
    s = float(yee_exact.gap_points(w, 8.0, 10.0)[0])
    for c in r["table"][:4]:
        F = GpuFactorization(Pencil(A, None), **c["options"])
        assert F.factorize(1.0, -s)
        assert F.inertia().neg == yee_exact.count_below(w, s)
        F.close()
    res = hsolve_amd.solve(A, None, sigma=9.0, k=12)
    assert any("autotuner" in n for n in res.report["parallel"]["notes"])
    out = ev.evaluate(A, None, res.eigenvalues, res.eigenvectors, {"mode": "nearest", "sigma": 9.0, "k": 12}, w)
    assert out["passed"] and res.report["certificate"]["passed"]
