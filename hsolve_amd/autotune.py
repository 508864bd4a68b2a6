"""Autotuner of the engine's per-pattern launch plan: which fronts go to the one-workgroup LDS
kernel and which to the blocked panel path (small_front_max: the split of every level's schedule),
the panel width of the blocked path (its diagonal-block and GEMM tile shape along K), and the
nested-dissection leaf size (how deep the ordering goes, which sets the front sizes).

The reference calibrates its CPU cost model per machine (src/driver/steps.cpp:612-677,
1236-1377); here the candidates are measured on the device itself: for each candidate the pattern
is analysed with those options and one factorization plus one block solve of `p` right-hand
sides at a representative shift are timed (HIP events; the first factorization of each candidate
is a warm-up). The cost is analysis_weight * t_analysis + factor_weight * (t_factor +
t_first_solve) + solve_weight * t_solve (the first solve after a factorization also builds the
inverse panels), weighted as a spectrum slice uses the factor (one factorization, tens of block
solves). The host analysis (ordering and symbolic analysis, the faster of two runs) counts with
weight 0.5 against one slice's device time: a solve runs it once with the device idle, while its
slices' device work overlaps on concurrent streams (box48: 6 slices, about 3.6 kernels in flight,
so 3.6 / 6 = 0.6). Without it, leaf 32 won on device time (17.6 against 18.0 ms per factorization)
while its deeper dissection cost 10-15 ms more host analysis, and box48 interval solves measured
0.549 s against 0.537 s per step with the shallower default plan (means of 6 steps). The best options are
kept per pattern (the hash of the factorized pattern and value maps) in a JSON cache, and
GpuFactorization objects created by the solver for that pattern use them from then on.

    cache file: $HSG_TUNE_CACHE, else ~/.cache/hsolve_amd/tuning.json;  HSG_AUTOTUNE=0 ignores it.

Tuning changes only the plan (ordering depth, kernel split, tile width): factorizations with any
options are exact eliminations of the same matrix, with the same inertia.
"""
from __future__ import annotations

import json
import os
import threading
import time
from itertools import product
from pathlib import Path

import numpy as np
import torch

from .factor import GpuFactorization, Pencil

SMALL_FRONT_MAX = (64, 96, 128)
PANEL_WIDTH = (32, 64)
LEAF_SIZE = (32, 64, 128)
DEFAULTS = {"leaf_size": 64, "panel_width": 64, "small_front_max": 96}

_lock = threading.Lock()
_memory: dict[str, dict] = {}


def cache_path() -> Path:
    p = os.environ.get("HSG_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "hsolve_amd" / "tuning.json"


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(key: str, entry: dict) -> None:
    path = cache_path()
    with _lock:
        doc = _load()
        doc[key] = entry
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            tmp = path.with_suffix(".tmp")
            tmp.write_text(json.dumps(doc, indent=1) + "\n")
            os.replace(tmp, path)
        except OSError:
            pass  # an unwritable cache only loses the result for later processes
        _memory[key] = entry


def pattern_key(pencil: Pencil) -> str:
    """The pattern hash of serve.pattern_key, remembered on the pencil."""
    k = getattr(pencil, "_tune_key", None)
    if k is None:
        from .serve import pattern_key as pk
        k = pk(pencil)
        pencil._tune_key = k
    return k


def options_for(pencil: Pencil) -> dict:
    """The tuned options of this pattern, or {} (the engine's defaults) when none are cached or
    HSG_AUTOTUNE=0."""
    if os.environ.get("HSG_AUTOTUNE", "1") == "0":
        return {}
    with _lock:
        if not _memory and not cache_path().exists():
            return {}  # nothing tuned on this host: no pattern hash needed
    key = pattern_key(pencil)
    with _lock:
        if key not in _memory:
            _memory[key] = _load().get(key)
        entry = _memory[key]
    return dict(entry["options"]) if entry else {}


def _time_candidate(pencil: Pencil, opts: dict, sigma: float, p: int, reps: int) -> dict:
    t0 = time.perf_counter()
    F = GpuFactorization(pencil, **opts)
    t_ana = time.perf_counter() - t0
    t0 = time.perf_counter()
    GpuFactorization(pencil, **opts).close()  # a second analysis: the first may carry warm-up costs
    t_ana = min(t_ana, time.perf_counter() - t0)
    x = torch.randn((p, pencil.n), dtype=torch.complex128 if pencil.is_complex else torch.float64,
                    device=torch.device("cuda", torch.cuda.current_device()),
                    generator=torch.Generator(device="cuda").manual_seed(7))
    F.factorize(1.0, -sigma, precision="double")
    y = x.clone()
    F.solve(y)
    torch.cuda.synchronize()
    tf, t1, ts = [], [], []
    for _ in range(reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        F.factorize(1.0, -sigma, precision="double")
        ev[1].record()
        y.copy_(x)
        F.solve(y)  # the first solve after a factorization also builds the inverse panels
        ev[2].record()
        y.copy_(x)
        F.solve(y)
        ev[3].record()
        torch.cuda.synchronize()
        tf.append(ev[0].elapsed_time(ev[1]) * 1e-3)
        t1.append(ev[1].elapsed_time(ev[2]) * 1e-3)
        ts.append(ev[2].elapsed_time(ev[3]) * 1e-3)
    info = F.info
    F.close()
    return {"options": opts, "factor_s": float(np.median(tf)), "first_solve_s": float(np.median(t1)),
            "solve_s": float(np.median(ts)), "analysis_s": t_ana, "flops": float(info.flops),
            "factor_entries": float(info.factor_entries)}


def tune(A, B=None, sigma: float = 0.0, p: int = 8, reps: int = 3, factor_weight: float = 1.0,
         solve_weight: float = 30.0, candidates=None, analysis_weight: float = 0.5) -> dict:
    """Measures the candidate plans for the pattern of (A, B) at shift sigma, stores the best in the
    cache and returns {"options", "cost", "table"} (table: every candidate's timings)."""
    pencil = A if isinstance(A, Pencil) else Pencil(A, B)
    default_grid = candidates is None
    if candidates is None:
        candidates = [dict(leaf_size=DEFAULTS["leaf_size"], panel_width=pw, small_front_max=sf)
                      for sf, pw in product(SMALL_FRONT_MAX, PANEL_WIDTH)]
    def cost(r):
        return (analysis_weight * r["analysis_s"] + factor_weight * (r["factor_s"] + r["first_solve_s"]) +
                solve_weight * r["solve_s"])

    table = [_time_candidate(pencil, c, sigma, p, reps) for c in candidates]
    for r in table:
        r["cost"] = cost(r)
    best = min(table, key=lambda r: r["cost"])
    if default_grid:
        # then the ordering depth at the best kernel split
        extra = [dict(best["options"], leaf_size=ls) for ls in LEAF_SIZE if ls != best["options"]["leaf_size"]]
        more = [_time_candidate(pencil, c, sigma, p, reps) for c in extra]
        for r in more:
            r["cost"] = cost(r)
        table += more
        best = min(table, key=lambda r: r["cost"])
    default = next((r for r in table if r["options"] == DEFAULTS), None)
    entry = {"options": best["options"], "cost": best["cost"], "sigma": sigma, "p": p,
             "weights": [factor_weight, solve_weight, analysis_weight],
             "default_cost": default["cost"] if default else None, "n": pencil.n,
             "device": torch.cuda.get_device_name(), "tuned_at": time.time()}
    _store(pattern_key(pencil), entry)
    return {"options": best["options"], "cost": best["cost"], "table": table,
            "default_cost": entry["default_cost"]}
