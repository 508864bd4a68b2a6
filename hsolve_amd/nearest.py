"""Nearest-k target with the exact Sylvester certificate and deflated reruns.

Restates solve_nearest (src/driver/definite.cpp:461-847) without the structure-candidate and
randomized-certificate options (both off by default, include/hsolve/solver.hpp:177, :213):
  1. factorize at sigma (near-singular handling), one Krylov run for the k wanted pairs;
  2. select the k nearest, verify, rescue pairs above tol;
  3. up to three certificate rounds: window ends pushed to the middle of the gap to the next
     Ritz hint beyond the selected pairs (:669-677), counts nu(b) - nu(a) on double
     factorizations, passed iff equal to the computed pairs in the window; otherwise keep the
     converged pairs, refactorize at sigma and rerun deflated for missing + 2 pairs with
     lead = missing (:759-797).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .driver_core import (Block, Ctx, DefiniteRun, block_metrics, count_batched, count_below, factorize_near, keep_converged,
                          lanczos_run, nearest_beyond, refactorize_at, select_nearest, verify_and_rescue)
from .errors import InvalidInputError, SingularError
from .krylov import LARGEST_MAGNITUDE


def auto_block_size(cfg: dict, nev: int) -> int:
    """Krylov block width: configured, else 8 for >= 16 wanted pairs, 4 for >= 8, else 1 (the
    reference's P4 rule p = 4 for nev >= 8, src/driver/steps.cpp:1316-1330, widened for the GPU
    where a p-column solve costs about as much as one column)."""
    if cfg["block_size"] > 0:
        return int(cfg["block_size"])
    if nev >= 16:
        return 8
    return 4 if nev >= 8 else 1


def rerun_block_size(cfg: dict, p: int, missing: int) -> int:
    """steps.cpp:1247-1265."""
    q = p if cfg["block_size"] > 0 else min(p, 4)
    return max(max(q, 1), min(4, (missing + 1) // 2 * 2))


def drop_found_pairs(run: DefiniteRun, sel_lambda, sigma: float, drop: int):
    """Fault injection (solver.hpp:179-184, definite.cpp:637-646): remove copies of the largest
    multiple eigenvalue among the selected pairs (one copy stays), then the farthest selected."""
    lam = run.all.lam
    sel = sorted(sel_lambda, key=lambda l: abs(l - sigma))
    removed = []
    groups = {}
    for l in sel:
        key = None
        for g in groups:
            if abs(g - l) <= 1e-8 * max(1.0, abs(l)):
                key = g
                break
        groups.setdefault(key if key is not None else l, []).append(l)
    victims = []
    for g, members in sorted(groups.items(), key=lambda kv: -len(kv[1])):
        if len(members) > 1:
            victims.extend(members[1:])
    victims += [l for l in reversed(sel) if l not in victims]
    victims = victims[:drop]
    keep = []
    used = [False] * len(victims)
    for j in range(lam.size):
        hit = -1
        for q, v in enumerate(victims):
            if not used[q] and lam[j] == v:
                hit = q
                break
        if hit >= 0:
            used[hit] = True
            removed.append(float(lam[j]))
        else:
            keep.append(j)
    run.all = run.all.take(keep)
    return removed


def solve_nearest(ctx: Ctx):
    """Returns (eigenvalues, vectors (m, n) device tensor, metrics)."""
    cfg, rep = ctx.cfg, ctx.rep
    n = ctx.pencil.n
    t = cfg["target"]
    k = int(t["k"])
    if k > n:
        raise InvalidInputError(f"target.k = {k} exceeds the matrix order {n}")
    nev = min(n, k)
    sigma, detection_off = factorize_near(ctx, float(t["sigma"]))
    rep["sigma_used"] = sigma
    p = auto_block_size(cfg, nev)
    rep["block_size"] = p
    run = DefiniteRun()
    lanczos_run(ctx, sigma, LARGEST_MAGNITUDE, nev, p, None, run)
    if not run.converged:
        ctx.warn("Krylov-Schur did not converge all requested pairs within max_restarts")
    zero_tol = cfg["zero_filter_tol"] * ctx.scale if cfg["zero_filter_tol"] >= 0 else -1.0
    state = {"sel": None, "metrics": None}

    def prepare():
        refactorize_at(ctx, sigma, detection_off)

    def select_and_verify():
        idx, dropped = select_nearest(run.all.lam, t["sigma"], k, zero_tol)
        rep["zero_modes_filtered"] = dropped
        sel = run.all.take(idx)
        sel, metrics = verify_and_rescue(ctx, sel, ctx.n2a, ctx.n2b, prepare)
        order, _ = select_nearest(sel.lam, t["sigma"], sel.size, -1.0)
        metrics = [metrics[j] for j in order]
        sel = sel.take(order)
        state["sel"], state["metrics"] = sel, metrics

    with ctx.clock.phase("verify"):
        select_and_verify()

    want_cert = cfg["inertia_check"] != "off" and cfg["certificate"] != "off"
    if want_cert and cfg["certificate_fault_drop"] > 0:
        removed = drop_found_pairs(run, list(state["sel"].lam), t["sigma"], cfg["certificate_fault_drop"])
        ctx.warn(f"certificate_fault_drop: {len(removed)} found pair(s) removed before the completeness check ("
                 + ", ".join(f"{l:.17g}" for l in removed) + ")")
        with ctx.clock.phase("verify"):
            select_and_verify()
    cert = rep["certificate"]

    def found_in(lo, hi):
        c = 0
        for l in run.all.lam:
            if math.isfinite(l) and lo <= l <= hi and not (zero_tol >= 0 and abs(l) <= zero_tol):
                c += 1
        return c

    for rnd in range(3 if want_cert else 0):
        sel = state["sel"]
        if sel.size == 0:
            break
        tiny = 1e-12 * max(ctx.scale, abs(t["sigma"]))
        hints = list(run.ritz_hints) + list(run.all.lam)

        def push_out(edge, direction):
            nb = nearest_beyond(hints, edge, direction, tiny)
            gap_default = max(1e-3 * abs(edge - t["sigma"]), 1e-8 * max(ctx.scale, 1e-300))
            return edge + direction * gap_default if math.isnan(nb) else 0.5 * (edge + nb)

        r = max(abs(l - t["sigma"]) for l in sel.lam)
        a_end = push_out(t["sigma"] - r, -1.0)
        b_end = push_out(t["sigma"] + r, +1.0)
        short_by = min(k, n) - sel.size
        step = 1e-9 * max(ctx.scale, abs(t["sigma"]))
        t0_cert = ctx.clock.t["certificate"]
        # both ends in one batched factorization when the solve may hold two at once; an end where
        # K is singular is then counted on its own (count_below moves it)
        nu_a, nu_b = count_batched(ctx, [a_end, b_end], 2) if ctx.streams > 1 else (None, None)
        if nu_a is None:
            nu_a, a_end, _, _ = count_below(ctx, a_end, -1.0, step, s_is_gap=True)
        if nu_b is None:
            nu_b, b_end, _, _ = count_below(ctx, b_end, +1.0, step, s_is_gap=True)
        cert["performed"] = True
        cert["lower"], cert["upper"] = a_end, b_end
        cert["count_inertia"] = nu_b - nu_a
        found = found_in(a_end, b_end)
        cert["count_found"] = found
        cert["passed"] = cert["count_inertia"] == found
        cert["method"] = "exact"
        cert["rigorous"] = cert["passed"]
        cert["failure_probability"] = 0.0 if cert["passed"] else 1.0
        cert["statement"] = (f"Sylvester inertia: nu({b_end:.17g}) - nu({a_end:.17g}) = {cert['count_inertia']} "
                             f"eigenvalues in the window; {found} computed")
        cert["seconds"] += ctx.clock.t["certificate"] - t0_cert
        ctx.event("certificate", round=rnd, lower=a_end, upper=b_end, count_inertia=cert["count_inertia"],
                  count_found=found, passed=cert["passed"], short_by=short_by)
        if cert["count_inertia"] < found or rnd == 2:
            break
        if cert["passed"]:
            if short_by == 0:
                break
            missing = short_by
            ctx.warn(f"certificate: only {sel.size} of {k} pairs selected; deflated rerun for the nearest ones beyond")
        else:
            missing = max(cert["count_inertia"] - found, short_by)
            ctx.warn(f"certificate: {cert['count_inertia'] - found} eigenvalue(s) missing in the window; deflated rerun")
        keep_converged(run, ctx.n2a, ctx.n2b, cfg["tol"])
        nev_rerun = min(n - run.all.size, missing + 2)
        if nev_rerun < 1:
            break
        refactorize_at(ctx, sigma, detection_off)
        defl = run.all.x.contiguous()
        p_rerun = rerun_block_size(cfg, p, missing)
        lanczos_run(ctx, sigma, LARGEST_MAGNITUDE, nev_rerun, p_rerun, defl if defl.shape[0] else None, run, missing)
        with ctx.clock.phase("verify"):
            select_and_verify()
    sel, metrics = state["sel"], state["metrics"]
    if cert["performed"] and cert["passed"] and sel.size < min(k, n):
        cert["statement"] += (f"; but only {sel.size} of {k} pairs were found: the window covers the found pairs, "
                              "not the k-th nearest eigenvalue")
    if cert["performed"] and not cert["passed"]:
        if cert["count_inertia"] < cert["count_found"]:
            ctx.warn(f"certificate: more computed eigenvalues than inertia allows (duplicates/ghosts) in "
                     f"[{cert['lower']:.17g}, {cert['upper']:.17g}]")
        else:
            ctx.warn(f"certificate failed: inertia {cert['count_inertia']} vs found {cert['count_found']}")
    if not want_cert:
        cert["note"] = "disabled by configuration"
    rep["restarts"] = run.restarts
    rep["op_applies"] = run.applies
    all_ok = sel.size == k
    if not all_ok:
        ctx.warn(f"only {sel.size} of {k} pairs found")
    return sel, metrics, all_ok and (not cert["performed"] or cert["passed"])
