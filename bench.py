#!/usr/bin/env python3
"""Headline benchmark: certified eigenpairs per second of the shift-invert engine on one MI355X.

Workload (box48): the generalized curl-curl pencil of a PEC box [0,2pi] x [0,pi] x [0,1.5pi] on a
48 x 24 x 36 Yee grid (n = 117,036), A = C^T C / h^2, B = diagonal edge mass with seeded jitter
1 + 0.05 U(-1, 1) (seed 2026), target: all eigenvalues in [40, b] with b set by exact inertia
bisection so that the window holds 500. One step is one complete interval solve (spectrum slices,
factorizations, block Krylov-Schur runs, the exact Sylvester certificate of every slice), started
from a pencil already resident in HBM and leaving the eigenvectors there.

A pair counts as certified when its backward error
    eta = ||A x - lambda B x|| / ((||A||_2 + |lambda| ||B||_2) ||x||) <= 1e-10,
recomputed here from the original matrices, and every slice's inertia certificate passed.
value = certified pairs / wall seconds over the timed steps (higher is better).

Also printed, in the same JSON line: p50/p99 step latency; the plain form (FP64 factors, one
request at a time) beside the chosen form; the roofline of the backward sweep (its launches
timed with HIP events on the engine's stream during the plain form's timed steps, HBM traffic
from the committed counter profile) and of the factorization's update GEMM; and the CPU baseline
(the oracle's C++/OpenMP multifrontal LDL^T and the same block Krylov-Schur driver on the host
cores, solving the first spectrum slice of the window end to end).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-plain | --plain-only]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import re
import sys
import time

import numpy as np

# The chosen form runs up to 8 spectrum slices and counts on concurrent streams; HIP maps streams
# onto GPU_MAX_HW_QUEUES hardware queues (HIP's default is 4), and streams that share a queue run
# one after the other. This process owns the device, so it sets 8 queues before HIP initializes
# (HSG_HW_QUEUES overrides the number; the value used is reported in the JSON line).
os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("HSG_HW_QUEUES", "8")

METRIC = "certified eigpairs/s, GPU chosen vs plain @bwd err≤1e-10; p50/p99 latency"
UNIT = "eigpairs/s"
TOL = 1e-10
LOWER, COUNT, SEED = 40.0, 500, 2026
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q / 100.0
    f, c = math.floor(k), math.ceil(k)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def norm2_estimate(A, seed=SEED):
    """||A||_2 of a symmetric matrix by scipy's Lanczos (independent of the engine)."""
    import scipy.sparse.linalg as sla
    v0 = np.random.default_rng(seed).uniform(-1.0, 1.0, A.shape[0])
    w = sla.eigsh(A, k=1, which="LM", v0=v0, tol=1e-6, return_eigenvectors=False)
    return float(abs(w[0]))


def backward_errors(A_dev, b_diag_dev, lam, X, na, nb):
    """eta per pair, X (m, n) on the device; A as a torch CSR tensor, B diagonal."""
    import torch
    if X.shape[0] == 0:
        return np.zeros(0)
    lam_d = torch.as_tensor(np.asarray(lam), dtype=torch.float64, device=X.device)
    AX = (A_dev @ X.T).T
    R = AX - lam_d[:, None] * (X * b_diag_dev[None, :])
    rn = torch.linalg.norm(R, dim=1)
    xn = torch.linalg.norm(X, dim=1)
    eta = rn / ((na + lam_d.abs() * nb) * xn)
    return eta.cpu().numpy()


PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_bwd_task_box48.md")


def measured_traffic(alg_bytes_per_launch):
    """HBM bytes per backward-sweep launch (k_bwd_task and k_bwd_reduce together) from the committed
    counter summary (FETCH_SIZE and
    WRITE_SIZE runs scaled by a copy of known size), used when it was taken on the same kernel
    source (hash of solve.hpp) and the same launches (algorithmic bytes per launch within 1 %);
    otherwise None."""
    try:
        text = open(PROFILE, encoding="utf-8").read()
        vals = dict(re.findall(r"^\s*(traffic_bytes_per_launch|algorithmic_bytes_per_launch|kernel_source_sha256_16)"
                               r" = ([0-9a-f]+)", text, re.M))
        t, a = float(vals["traffic_bytes_per_launch"]), float(vals["algorithmic_bytes_per_launch"])
        src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hsolve_amd", "csrc", "kernels", "solve.hpp")
        sha = hashlib.sha256(open(src, "rb").read()).hexdigest()[:16]
    except (OSError, KeyError, ValueError):
        return None, None
    if sha != vals["kernel_source_sha256_16"]:
        return None, "profile taken on another version of the solve kernels: not used"
    if alg_bytes_per_launch <= 0 or abs(a - alg_bytes_per_launch) > 0.01 * alg_bytes_per_launch:
        return None, "profile taken on other launches: not used"
    return t, "profiles/k_bwd_task_box48.md (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-calibrated)"


def run_form(name, prepared, cfg, steps, warmup, check, streams):
    """Warm-up, then `steps` timed solves; returns latencies, certified counts, last result."""
    import torch
    from hsolve_amd import profiling, solver
    for i in range(warmup):
        log(f"{name}: warm-up step {i + 1}/{warmup}")
        solver.solve(config=cfg, prepared=prepared, vectors="device", streams=streams)
    torch.cuda.synchronize()
    profiling.enable(True)
    lat, cert, reps = [], [], []
    last = None
    for i in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lam, X, rep, _ = solver.solve(config=cfg, prepared=prepared, vectors="device", streams=streams)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        eta = check(lam, X)
        slices_ok = all(s.get("passed", False) for s in rep.get("slices", [])) if rep.get("slices") else bool(
            rep["certificate"].get("passed", False))
        in_window = np.all((np.asarray(lam) >= cfg["target"]["lower"]) & (np.asarray(lam) <= cfg["target"]["upper"]))
        good = int(np.count_nonzero(eta <= TOL)) if (slices_ok and in_window) else 0
        lat.append(dt)
        cert.append(good)
        reps.append(rep)
        log(f"{name}: step {i + 1}/{steps} {dt:.3f} s, {len(lam)} pairs, certified {good}, max eta "
            f"{float(eta.max()) if eta.size else float('nan'):.3g}, slices ok {slices_ok}")
        last = (lam, X, rep)
    phases = {ph: profiling.read(ph) for ph in profiling.PHASES}
    profiling.enable(False)
    return {"lat": lat, "cert": cert, "reps": reps, "last": last, "phases": phases}


def summary(r):
    tot = sum(r["lat"])
    return {"value": sum(r["cert"]) / tot if tot > 0 else 0.0, "ms_per_step": 1e3 * tot / len(r["lat"]),
            "latency_p50_ms": 1e3 * percentile(r["lat"], 50), "latency_p99_ms": 1e3 * percentile(r["lat"], 99),
            "certified_per_step": r["cert"]}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="run the plain form alone (for counter runs of the kernels the roofline is taken on)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES bytes (the reference kernel that "
                         "hardware byte counters are scaled by)")
    ap.add_argument("--workload", choices=("box48", "bloch16"), default="box48",
                    help="box48 (default, the headline interval solve) or bloch16 (64 Bloch k-points of a "
                         "photonic cell, complex Hermitian, solved as one batch)")
    ap.add_argument("--workers", type=int, default=8, help="bloch16: concurrent pencils of the batch")
    ap.add_argument("--extras", default="autotune,single,bloch16,serving,box80",
                    help="comma-separated extra measurements added to the box48 line as keys: autotune (the "
                         "launch-plan autotuner, run before the timed steps; the chosen form uses its plan), single (the "
                         "single-precision working-factor form of box48), bloch16 (the 64-k-point batch, one "
                         "step), serving (a Poisson request stream), box80 (the million-order nearest solve); "
                         "'' for none")
    ap.add_argument("--serve-requests", type=int, default=40, help="serving: requests in the stream")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: each unit of this workload (a spectrum slice, an inertia count) "
              "runs whole on one device, and runs on more than one GPU are not part of this build; "
              f"got --gpus {args.gpus} (WORLD_SIZE {world})", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        ap.error("--steps must be >= 1 and --warmup >= 0")

    import torch
    from hsolve_amd import solver, workloads
    from hsolve_amd.factor import GpuFactorization
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU (torch.cuda.is_available() is False)", file=sys.stderr)
        sys.exit(1)

    if args.copy_calibration > 0:
        nel = args.copy_calibration // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {8 * nel} bytes copied device to device")
        del src, dst
    if args.workload == "bloch16":
        return bench_bloch(args)
    t0 = time.perf_counter()
    w = workloads.box48(seed=SEED, lower=LOWER, count=COUNT)
    A, B = w.A, w.B
    upper = float(w.target["upper"])
    log(f"workload box48 n={A.shape[0]} window [{LOWER}, {upper:.12g}] ({COUNT} eigenvalues by inertia), "
        f"set-up {time.perf_counter() - t0:.1f} s")

    base_cfg = {"target": {"mode": "interval", "lower": LOWER, "upper": upper}, "tol": TOL}
    chosen_cfg = dict(base_cfg, factor_precision="auto")
    plain_cfg = dict(base_cfg, factor_precision="double")
    prepared = solver.prepare(A, B, chosen_cfg)  # untimed: the pencil resident in HBM

    # independent backward errors: ||A||_2 by scipy Lanczos, ||B||_2 = max diagonal (B diagonal)
    na = norm2_estimate(A)
    bd = B.diagonal()
    nb = float(np.max(np.abs(bd)))
    A_dev = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)),
                                    torch.from_numpy(A.indices.astype(np.int64)),
                                    torch.from_numpy(A.data.astype(np.float64)), size=A.shape).cuda()
    b_dev = torch.from_numpy(bd.astype(np.float64)).cuda()

    def check(lam, X):
        return backward_errors(A_dev, b_dev, lam, X, na, nb)

    # chosen: the engine's own choices (working precision, concurrent slice streams, the launch
    # plan its autotuner measured for this pattern); plain: FP64 factors, one slice after the
    # other, the default launch plan
    extras = [e for e in args.extras.split(",") if e]
    tuning = None
    if "autotune" in extras and not args.plain_only:
        tuning = guarded_extra("autotune", lambda: autotune_extra(prepared, 0.5 * (LOWER + upper)))
    if args.plain_only:
        chosen_cfg, args.no_plain = plain_cfg, False
    os.environ["HSG_AUTOTUNE"] = "0" if args.plain_only else "1"
    chosen = run_form("plain" if args.plain_only else "chosen", prepared, solver_cfg(chosen_cfg), args.steps,
                      args.warmup, check, 1 if args.plain_only else None)
    cs = summary(chosen)
    rep = chosen["reps"][-1]
    same_form = rep["precision"]["used"] == "double"
    plain, plain_s = None, None
    if not args.no_plain and not args.plain_only:
        os.environ["HSG_AUTOTUNE"] = "0"
        plain = run_form("plain", prepared, solver_cfg(plain_cfg), args.steps, args.warmup, check, 1)
        os.environ["HSG_AUTOTUNE"] = "1"
        plain_s = summary(plain)

    # roofline: the backward sweep's kernels (the dominant ones of this workload), timed in the
    # plain form's steps: with slices on concurrent streams an event interval on one stream also
    # spans other streams' kernels, so the single-stream run gives the kernels' own time
    ph = plain["phases"] if plain_s is not None else chosen["phases"]
    bw = ph["backward"]
    achieved = bw["work"] / bw["seconds"] / 1e9 if bw["seconds"] > 0 else 0.0
    traffic, traffic_src = measured_traffic(bw["work"] / bw["launches"] if bw["launches"] else 0.0)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "hsg::k_bwd_task (+ hsg::k_bwd_reduce)", "launches": bw["launches"],
                "timed_in": "plain form (one stream)" if plain_s is not None else "chosen form",
                "avg_launch_us": 1e6 * bw["avg_launch_s"],
                "algorithmic_bytes_per_launch": bw["work"] / bw["launches"] if bw["launches"] else 0.0}
    phases = {k: {"seconds": v["seconds"], "launches": v["launches"], "calls": v["calls"],  # the same run
                  "rate": v["rate"] / (1e12 if v["unit"] == "flop" else 1e9),
                  "rate_unit": "TFLOP/s" if v["unit"] == "flop" else "GB/s"} for k, v in ph.items()}

    factor_roof = guarded_extra("factor roofline", lambda: factor_roofline(prepared, 0.5 * (LOWER + upper)))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = guarded_extra("cpu baseline", lambda: cpu_baseline(A, B, rep, na, nb))
    out = {
        "metric": METRIC, "value": cs["value"], "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": cs["ms_per_step"], "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64" if same_form else "f32+f64", "data": "synthetic",
        "config": {"workload": "box48: PEC box GHEP, 48x24x36 Yee cells, n=117036, jittered diagonal B (seed 2026), "
                               f"all {COUNT} eigenvalues in [40, b] by inertia bisection",
                   "n": int(A.shape[0]), "lower": LOWER, "upper": upper, "pairs": COUNT, "tol": TOL,
                   "form": "plain" if args.plain_only else "chosen", "precision_used": rep["precision"]["used"],
                   "slice_streams": rep["parallel"]["notes"][-1] if rep["parallel"]["notes"] else "1 stream",
                   "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"])},
        "latency_p50_ms": cs["latency_p50_ms"], "latency_p99_ms": cs["latency_p99_ms"],
        "certified_per_step": cs["certified_per_step"],
        "plain": plain_s,
        "chosen_over_plain": (cs["value"] / plain_s["value"]) if plain_s and plain_s["value"] > 0 else None,
        "roofline": roofline,
        "factor_roofline": factor_roof,
        "phases": phases,
        "solve_phases_s": {form: {ph: r["reps"][-1][f"time_{ph}_s"]
                                  for ph in ("setup", "factor", "certificate", "iterate", "verify")}
                           for form, r in (("chosen", chosen), ("plain", plain)) if r is not None},
        "cpu_baseline": cpu,
    }
    if tuning is not None:
        out["autotune"] = tuning
    if "single" in extras and not args.plain_only:
        out["single_form"] = guarded_extra("single", lambda: single_form(prepared, base_cfg, check, args))
    if "bloch16" in extras:
        out["bloch16_batch"] = guarded_extra("bloch16", lambda: bloch_batch_extra(args))
    if "serving" in extras:
        out["serving"] = guarded_extra("serving", lambda: serving_extra(args))
    if "box80" in extras:
        out["box80"] = guarded_extra("box80", lambda: box80_extra(args))
    print(json.dumps(out), flush=True)


def cpu_baseline(A, B, rep, na, nb):
    """The CPU path (oracle/cpu_path.py: the engine's host analysis, the C++/OpenMP multifrontal
    LDL^T with OpenBLAS kernels, the same block Krylov-Schur driver on host tensors) on a bounded
    sample of the box48 workload: the first spectrum slice of the chosen GPU run, [40, cut], solved
    end to end on the host cores (analysis, the two exact counts, the factorization at the slice
    centre, the Krylov run, Rayleigh-Ritz); certified pairs (eta <= 1e-10 recomputed here, count
    equal to the pairs found) per second."""
    from oracle import cpu_path
    cpu_path.lib()
    sl = rep["slices"][0]
    lo, hi = float(sl["lower"]), float(sl["upper"])
    log(f"cpu baseline: slice [{lo:.10g}, {hi:.10g}] of box48 on the host ({cpu_path.host_description()})")
    r = cpu_path.solve_slice(A, B, lo, hi)
    lam, X = r["eigenvalues"], r["eigenvectors"]
    eta = (np.linalg.norm(A @ X - (B @ X) * lam[None, :], axis=0) /
           ((na + np.abs(lam) * nb) * np.linalg.norm(X, axis=0))) if lam.size else np.zeros(0)
    good = int(np.count_nonzero(eta <= TOL)) if r["passed"] else 0
    host = cpu_path.host_description()
    return {"value": good / r["time_total_s"], "unit": UNIT, "cores": host["omp_threads"], "kind": "port",
            "sample": (f"the first spectrum slice [{lo:.6g}, {hi:.6g}] of the box48 window ({r['count_inertia']} "
                       f"eigenvalues) solved end to end on the CPU path: analysis {r['time_analysis_s']:.2f} s, "
                       f"2 exact counts {r['time_counts_s']:.2f} s, factorization {r['time_factor_s']:.2f} s, block "
                       f"Krylov-Schur {r['time_iterate_s']:.2f} s ({r['op_applies']} applications), Rayleigh-Ritz "
                       f"{r['time_rr_s']:.2f} s; {good} certified pairs in {r['time_total_s']:.2f} s"),
            "host": host, "max_eta": float(eta.max()) if eta.size else None, "certified": good,
            "seconds": r["time_total_s"], "phases_s": {k: r[k] for k in ("time_analysis_s", "time_counts_s",
                                                                          "time_factor_s", "time_iterate_s",
                                                                          "time_rr_s")}}


def guarded_extra(name, fn):
    """An extra measurement; a failure is reported in its key and does not lose the headline line."""
    t0 = time.perf_counter()
    try:
        r = fn()
    except Exception as e:  # noqa: BLE001 (reported in the JSON line)
        log(f"extra {name} failed: {type(e).__name__}: {e}")
        return {"error": f"{type(e).__name__}: {e}"}
    r["wall_s"] = time.perf_counter() - t0
    log(f"extra {name} done in {r['wall_s']:.1f} s")
    return r


MFMA_PEAK_TFS = {"double": 78.6, "single": 157.3}  # MI355X dense matrix-core peaks, FP64 / FP32


def factor_roofline(prepared, sigma, reps=3):
    """Front-update GEMMs (k_gemm_update on the matrix cores) and whole factorizations of box48 at
    sigma, FP64 and FP32, each launch timed with HIP events on the factorization's stream:
    achieved = algorithmic update flops (lower-triangle outputs, from the symbolic analysis) / the
    launches' summed time, against the matrix-core peak of that precision."""
    import torch
    from hsolve_amd import profiling
    from hsolve_amd.factor import GpuFactorization
    F = GpuFactorization(prepared.pencil)
    out = {}
    for prec in ("double", "single"):
        F.factorize(1.0, -sigma, precision=prec)  # warm-up (allocation, first-use costs)
        torch.cuda.synchronize()
        profiling.enable(True, detail=True)
        for _ in range(reps):
            F.factorize(1.0, -sigma, precision=prec)
        torch.cuda.synchronize()
        up, fa = profiling.read("update"), profiling.read("factor")
        profiling.enable(False)
        peak = MFMA_PEAK_TFS[prec]
        ach = up["rate"] / 1e12
        out[prec] = {"bound": "mfma", "kernel": "hsg::k_gemm_update", "achieved": ach, "peak": peak,
                     "unit": "TFLOP/s", "frac": ach / peak, "update_flops_per_factorization": up["work"] / reps,
                     "update_ms_per_factorization": 1e3 * up["seconds"] / reps,
                     "launches_per_factorization": up["launches"] / reps,
                     "factorization_ms": 1e3 * fa["seconds"] / reps,
                     "factorization_tflops": fa["rate"] / 1e12,
                     "factorization_frac": fa["rate"] / 1e12 / peak,
                     "factorization_flops": fa["work"] / reps}
    F.close()
    return out


def autotune_extra(prepared, sigma):
    """The engine's autotuner on the box48 pattern (untimed, once per pattern, like a warm-up): the
    measured launch plans and the one the chosen form then uses (cache in a temporary file)."""
    import tempfile
    from hsolve_amd import autotune
    os.environ.setdefault("HSG_TUNE_CACHE", os.path.join(tempfile.mkdtemp(prefix="hsg-tune-"), "tuning.json"))
    t0 = time.perf_counter()
    r = autotune.tune(prepared.pencil, sigma=sigma)
    return {"chosen": r["options"], "cost_s": r["cost"], "default_cost_s": r["default_cost"],
            "tuning_s": time.perf_counter() - t0,
            "candidates": [{"options": c["options"], "factor_ms": 1e3 * c["factor_s"],
                            "first_solve8_ms": 1e3 * c["first_solve_s"], "solve8_ms": 1e3 * c["solve_s"],
                            "analysis_s": c["analysis_s"], "flops": c["flops"]} for c in r["table"]]}


def single_form(prepared, base_cfg, check, args):
    """box48 with single-precision working factors (FP32 on the matrix cores) refined to FP64 in
    every operator application; counts stay FP64. One warm-up and up to two timed steps."""
    from hsolve_amd import solver
    cfg = dict(base_cfg, factor_precision="single")
    r = run_form("single", prepared, solver_cfg(cfg), min(args.steps, 2), 1, check, None)
    sm = summary(r)
    rep = r["reps"][-1]
    fs = rep["factor_stats"]
    return {"value": sm["value"], "unit": UNIT, "ms_per_step": sm["ms_per_step"],
            "certified_per_step": sm["certified_per_step"], "precision_used": rep["precision"]["used"],
            "fallback": rep["precision"]["fallback"], "factorizations": fs["factorizations"],
            "solves": fs["solves"], "rhs_solved": fs["rhs_solved"],
            "solve_phases_s": {ph: rep[f"time_{ph}_s"] for ph in ("setup", "factor", "certificate", "iterate")}}


def bloch_batch_extra(args):
    """bloch16 chosen form (solve_batch), one warm-up and one timed step."""
    import argparse as _ap
    a = _ap.Namespace(**vars(args))
    a.steps, a.warmup, a.no_plain, a.no_cpu_baseline = 1, 1, True, True
    res = bench_bloch(a, emit=False)
    keep = ("value", "unit", "ms_per_step", "certified_per_step", "config")
    return {k: res[k] for k in keep if k in res}


def serving_extra(args):
    """A scaled-down serving stream (the million-request stream's mix at smaller sizes): PEC boxes of
    5e3 - 1.5e5 unknowns at random sigma and k in 5..40, 20 % Bloch 16^3 cells over their window,
    Poisson arrivals at 6 requests per second, 4 slots; every result certified (eta <= 1e-10
    recomputed on the host, certificate passed)."""
    import scipy.sparse.linalg as sla
    from hsolve_amd import serve, workloads
    reqs = workloads.serving_stream(args.serve_requests, seed=SEED, rate=6.0, n_max=1.5e5)
    out, sm = serve.run_stream(reqs, slots=4, vectors="host")
    certified = pairs = 0
    for r, res in zip(reqs, out):
        if not isinstance(res, serve.ServeResult):
            continue
        lam, X = res.eigenvalues, res.eigenvectors
        pairs += lam.size
        if not lam.size:
            continue
        A, B = r["A"], r["B"]
        v0 = np.random.default_rng(SEED).uniform(-1, 1, A.shape[0])
        na = float(np.max(np.abs(sla.eigsh(A, k=1, which="LM", tol=1e-6, v0=v0, return_eigenvectors=False))))
        nb = 1.0 if B is None else float(np.max(np.abs(B.diagonal())))
        BX = X if B is None else B @ X
        eta = np.linalg.norm(A @ X - BX * lam[None, :], axis=0) / ((na + np.abs(lam) * nb) * np.linalg.norm(X, axis=0))
        if res.report["certificate"]["passed"]:
            certified += int(np.count_nonzero(eta <= TOL))
    kinds = {}
    for r in reqs:
        kinds[r["kind"]] = kinds.get(r["kind"], 0) + 1
    return {"requests": len(reqs), "kinds": kinds, "n_range": [min(r["n"] for r in reqs), max(r["n"] for r in reqs)],
            "rate_per_s": 6.0, "slots": 4, "completed": sm["completed"], "failed": sm["failed"],
            "latency_p50_ms": sm["latency_p50_ms"], "latency_p99_ms": sm["latency_p99_ms"],
            "throughput_req_s": sm["throughput_req_s"], "pairs": pairs, "certified_pairs": certified,
            "certified_pairs_per_s": certified / sm["makespan_s"], "makespan_s": sm["makespan_s"],
            "analyses": sm["analyses"], "analysis_hits": sm["analysis_hits"], "peak_running": sm["peak_running"],
            "peak_predicted_gb": sm["peak_predicted_bytes"] / 2 ** 30, "budget_gb": sm["budget_gb"]}


def box80_extra(args):
    """The million-order GHEP (80 x 70 x 64 cells, n = 1,045,014, jittered diagonal B, seed 2026):
    the nearest 20 eigenvalues at a mid-spectrum shift with the exact certificate, one solve from a
    device-resident pencil (its analysis inside the timed region, as in box48); eta recomputed on
    the host from A and B."""
    import scipy.sparse.linalg as sla
    import torch
    from hsolve_amd import solver, workloads
    w = workloads.box80()
    A, B = w.A, w.B
    cfg = {"target": dict(w.target), "tol": TOL}
    prepared = solver.prepare(A, B, cfg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam, X, rep, _ = solver.solve(config=solver_cfg(cfg), prepared=prepared, vectors="host")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    v0 = np.random.default_rng(SEED).uniform(-1, 1, A.shape[0])
    na = float(abs(sla.eigsh(A, k=1, which="LM", tol=1e-6, v0=v0, return_eigenvectors=False)[0]))
    nb = float(np.max(np.abs(B.diagonal())))
    eta = np.linalg.norm(A @ X - (B @ X) * lam[None, :], axis=0) / ((na + np.abs(lam) * nb) * np.linalg.norm(X, axis=0))
    cert = rep["certificate"]
    good = int(np.count_nonzero(eta <= TOL)) if cert["passed"] else 0
    fs = rep["factor_stats"]
    return {"n": int(A.shape[0]), "sigma": w.target["sigma"], "k": w.target["k"], "seconds": dt,
            "certified": good, "certified_pairs_per_s": good / dt, "max_eta": float(eta.max()) if eta.size else None,
            "certificate": {k: cert[k] for k in ("lower", "upper", "count_inertia", "count_found", "passed")},
            "phases_s": {ph: rep[f"time_{ph}_s"] for ph in ("setup", "factor", "certificate", "iterate", "verify")},
            "factorizations": fs["factorizations"], "factor_entries": fs["factor_entries"],
            "flops_per_factorization": fs["estimate"]["flops"], "factor_mb": fs["factor_mb_used"],
            "device_in_use_gb_after": (lambda fm: (fm[1] - fm[0]) / 2 ** 30)(torch.cuda.mem_get_info()),
            "note": "one cold solve (first factorization of this pattern in the process)"}


def bench_bloch(args, emit=True):
    """bloch16: the 64 k-points of a 4 x 4 x 4 grid over [0, pi/a]^3 of a periodic 16^3 Yee cell with
    eps = 12 inside a centred sphere of radius 0.3 a (n = 12,288 complex unknowns per k-point), all
    eigenvalues in [0.01, 1.0] (2 pi / a)^2 per k-point. One step solves the whole batch: chosen =
    solve_batch (one analysis per factorized pattern, pencils on --workers concurrent streams),
    plain = one k-point after the other, FP64, one stream, an analysis each. Certified as box48:
    eta <= 1e-10 from the original matrices (computed on the host) with every certificate passed."""
    import scipy.sparse.linalg as sla
    import torch
    from hsolve_amd import batch, profiling, solver, workloads
    t0 = time.perf_counter()
    kpts = workloads.bloch_kpoints(4)
    wls = [workloads.bloch16(k) for k in kpts]
    target = dict(wls[0].target)
    base_cfg = {"target": target, "tol": TOL}
    chosen_cfg = dict(base_cfg, factor_precision="auto")
    plain_cfg = dict(base_cfg, factor_precision="double")
    prepared = [solver.prepare(w.A, w.B, chosen_cfg) for w in wls]
    norms = []
    for w in wls:
        v0 = np.random.default_rng(SEED).uniform(-1, 1, w.A.shape[0])
        na = float(np.max(np.abs(sla.eigsh(w.A, k=1, which="LM", tol=1e-6, v0=v0, return_eigenvectors=False))))
        norms.append((na, float(np.max(np.abs(w.B.diagonal())))))
    log(f"workload bloch16: {len(wls)} k-points, n={wls[0].A.shape[0]} complex, window [{target['lower']:.6g}, "
        f"{target['upper']:.6g}], set-up {time.perf_counter() - t0:.1f} s")

    def certified(results):
        good = total = 0
        for w, (na, nb), (lam, X, rep, _) in zip(wls, norms, results):
            lam = np.asarray(lam)
            total += lam.size
            if not lam.size:
                continue
            Xh = X.cpu().numpy().T
            R = w.A @ Xh - (w.B @ Xh) * lam[None, :]
            eta = np.linalg.norm(R, axis=0) / ((na + np.abs(lam) * nb) * np.linalg.norm(Xh, axis=0))
            ok = rep["certificate"]["passed"] and all(s.get("passed", False) for s in rep.get("slices", []))
            inside = np.all((lam >= target["lower"]) & (lam <= target["upper"]))
            if ok and inside:
                good += int(np.count_nonzero(eta <= TOL))
        return good, total

    def run(name, fn):
        for i in range(args.warmup):
            log(f"{name}: warm-up step {i + 1}/{args.warmup}")
            fn()
        torch.cuda.synchronize()
        profiling.enable(True)
        lat, cert = [], []
        for i in range(args.steps):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            res = fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t1
            good, total = certified(res)
            lat.append(dt)
            cert.append(good)
            log(f"{name}: step {i + 1}/{args.steps} {dt:.3f} s, {total} pairs, certified {good}")
        phases = {ph: profiling.read(ph) for ph in profiling.PHASES}
        profiling.enable(False)
        return {"lat": lat, "cert": cert, "reps": [r[2] for r in res], "phases": phases}

    chosen = run("chosen", lambda: batch.solve_batch(prepared, solver_cfg(chosen_cfg), workers=args.workers,
                                                     vectors="device"))
    cs = summary(chosen)
    plain_s, plain = None, None
    if not args.no_plain:
        plain = run("plain", lambda: [solver.solve(config=solver_cfg(plain_cfg), prepared=p, vectors="device",
                                                   streams=1) for p in prepared])
        plain_s = summary(plain)
    ph = plain["phases"] if plain_s is not None else chosen["phases"]
    bw = ph["backward"]
    achieved = bw["work"] / bw["seconds"] / 1e9 if bw["seconds"] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "traffic_source": "no counter summary taken on this workload",
                "kernel": "hsg::k_bwd_task (+ hsg::k_bwd_reduce)", "launches": bw["launches"],
                "timed_in": "plain form (one stream)" if plain_s is not None else "chosen form",
                "avg_launch_us": 1e6 * bw["avg_launch_s"],
                "algorithmic_bytes_per_launch": bw["work"] / bw["launches"] if bw["launches"] else 0.0}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import eigs_cpu
        from hsolve_amd.factor import GpuFactorization
        w = wls[len(wls) // 2]
        emb = GpuFactorization(prepared[len(wls) // 2].pencil).order()
        seen, perm = set(), []
        for e in emb:                     # the embedding's order, each complex unknown at its first copy
            if e // 2 not in seen:
                seen.add(e // 2)
                perm.append(e // 2)
        sigma_s = 0.5 * (target["lower"] + target["upper"])
        smp = eigs_cpu.factor_solve_sample(w.A, w.B, sigma_s, np.asarray(perm), p=8)
        facs = sum(r["factor_stats"]["factorizations"] for r in chosen["reps"])
        rhs = sum(r["factor_stats"]["rhs_solved"] for r in chosen["reps"])
        est = facs * smp["factor_s"] + rhs * smp["solve_s"] / smp["p"]
        cert_last = chosen["cert"][-1]
        cpu = {"value": cert_last / est if est > 0 else 0.0, "unit": UNIT, "cores": smp["cores"], "kind": "port",
               "sample": (f"one complex SuperLU elimination (symmetric mode, diagonal pivots) of K = A - {sigma_s:.6g} B "
                          f"of the k-point {[round(float(x), 4) for x in kpts[len(wls) // 2]]}: {smp['factor_s']:.3f} s, and one "
                          f"8-right-hand-side solve: {smp['solve_s']:.4f} s, scaled by the batch's {facs} "
                          f"factorizations and {rhs} right-hand sides ({est:.1f} s for {cert_last} certified pairs)")}
    out = {
        "metric": METRIC, "value": cs["value"], "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": cs["ms_per_step"], "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": "bloch16: 64 Bloch k-points (4x4x4 over [0, pi/a]^3) of a periodic 16^3 Yee cell, "
                               "eps = 12 in a centred sphere of radius 0.3a, n=12288 complex, all eigenvalues in "
                               "[0.01, 1.0](2pi/a)^2 per k-point",
                   "n": int(wls[0].A.shape[0]), "kpoints": len(wls), "tol": TOL, "form": "chosen",
                   "workers": args.workers,
                   "batch": chosen["reps"][0]["parallel"]["notes"][-1] if chosen["reps"][0]["parallel"]["notes"] else ""},
        "latency_p50_ms": cs["latency_p50_ms"], "latency_p99_ms": cs["latency_p99_ms"],
        "certified_per_step": cs["certified_per_step"],
        "plain": plain_s,
        "chosen_over_plain": (cs["value"] / plain_s["value"]) if plain_s and plain_s["value"] > 0 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def solver_cfg(cfg):
    return json.loads(json.dumps(cfg))


if __name__ == "__main__":
    main()
