"""The serving loop on the GPU (hsolve_amd.serve): a seeded Poisson stream of mixed requests (PEC
boxes with B = I or jittered mass at random sigma and k, Bloch cells with complex phases over a
window) runs with several requests on the device at once; every result equals the same request
solved alone (bitwise: each request owns its factorization and Krylov scratch, all reductions
have fixed orders), passes the evaluator (eta <= 1e-10 recomputed from A and B, B-orthonormal
vectors) and its certificate. Under a tight memory budget admission keeps the predicted memory
of the running requests within it, and a request larger than the budget fails alone with
OutOfMemoryBudgetError."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import evaluate as ev

pytestmark = pytest.mark.gpu


def _stream(count, seed):
    from hsolve_amd import workloads
    return workloads.serving_stream(count, seed=seed, rate=25.0, n_min=2e3, n_max=2.5e4, bloch_fraction=0.25,
                                    bloch_cells=6)


def test_results_in_the_mix_equal_results_alone(gpu):
    from hsolve_amd import serve, solver
    reqs = _stream(10, 7)
    out, summary = serve.run_stream(reqs, slots=3)
    print("summary", {k: summary[k] for k in ("completed", "latency_p50_ms", "latency_p99_ms", "peak_running",
                                                "analyses", "analysis_hits")}, flush=True)
    assert summary["completed"] == len(reqs) and summary["peak_running"] >= 2
    for r, res in zip(reqs, out):
        assert isinstance(res, serve.ServeResult), res
        lam, X, rep, _ = solver.solve(r["A"], r["B"], r["config"], streams=1)
        e = ev.evaluate(r["A"], r["B"], res.eigenvalues, res.eigenvectors, r["config"]["target"], None)
        print(r["kind"], r["n"], "pairs", res.eigenvalues.size, "max eta", e["max_eta"], "latency",
              round(res.latency_s, 3), flush=True)
        assert np.array_equal(res.eigenvalues, lam) and np.array_equal(res.eigenvectors, X)
        assert e["passed"] and res.report["certificate"]["passed"] and res.report["success"]


def test_admission_respects_a_tight_budget(gpu):
    from hsolve_amd import serve
    reqs = _stream(6, 11)
    # the largest single prediction, then a budget that holds it but not all requests together
    from hsolve_amd.factor import GpuFactorization
    from hsolve_amd.solver import prepare
    from hsolve_amd import config as cfgmod
    needs = []
    for r in reqs:
        p = prepare(r["A"], r["B"], r["config"])
        f = GpuFactorization(p.pencil)
        needs.append(serve.predict_bytes(f, cfgmod.parse_config(r["config"]), p.pencil.n, p.pencil.is_complex))
        f.close()
    budget = 1.05 * max(needs)
    out, summary = serve.run_stream(reqs, slots=4, memory_budget_gb=budget / 2 ** 30)
    print("needs", needs, "budget", budget, "peak", summary["peak_predicted_bytes"], flush=True)
    assert summary["completed"] == len(reqs)
    assert summary["peak_predicted_bytes"] <= budget
    for res in out:
        assert res.report["certificate"]["passed"] and res.report["success"]


def test_request_larger_than_the_budget_fails_alone(gpu):
    from hsolve_amd import serve
    from hsolve_amd.errors import OutOfMemoryBudgetError
    from hsolve_amd.gen import yee
    small = yee.cube_cavity(5)[0]
    big = yee.cube_cavity(12)[0]
    # predictions: about 0.8 MiB for the 5^3 cavity, 42 MiB for the 12^3 one
    with serve.Server(slots=2, memory_budget_gb=0.01) as srv:
        f_small = srv.submit(small, None, {"target": {"mode": "nearest", "sigma": 9.0, "k": 6}})
        f_big = srv.submit(big, None, {"target": {"mode": "nearest", "sigma": 9.0, "k": 6}})
        with pytest.raises(OutOfMemoryBudgetError):
            f_big.result()
        assert f_small.result().report["certificate"]["passed"]
