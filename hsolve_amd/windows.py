"""Window targets (interval / smallest / largest) by spectrum slicing with exact counts.

Restates WindowSolver (src/driver/definite.cpp:927-2275) for B Hermitian positive definite,
without the zero band (zero_filter_tol < 0, the default) and without helper processes:
  * every count is an exact double-precision Sylvester count nu(s) = #{lambda < s}, remembered
    (count, :1044-1057), so each slice knows how many eigenvalues it holds;
  * cuts by inertia search: regula falsi on the counts with the Illinois rule, bisection after a
    repeated end count, geometric expansion for open ends, clusters that counts cannot split end the
    search beside them (find_count, :1061-1144; gap_point, :1146-1190);
  * plan: slices of at most c_max eigenvalues marching up from the lower end (:1192-1229);
  * shift: the slice centre, or a one-sided anchor at an end beyond which nothing lies (:1232-1246);
  * run_slice: one Krylov run for the slice's count, selection within an end resolution, rescue,
    ghosts dropped, deflated reruns while pairs are missing (with a multiple eigenvalue's own
    shift), at most missing + 4 (<= 32) rounds, stop after 2 stalls (:1362-1520);
  * rank-revealing de-duplication of pairs near the cuts (:1522-1556), assignment of pairs to
    windows with the end checks (assign_to_windows, :66-134), target order of the output.
The reference hands slices to helper processes (definite.cpp:1586-1837); here independent slices
run at the same time on separate streams of the one device, each on its own factorization object
over the shared analysis, and their results are merged in slice order (so they do not depend on
which slice finished first). With more than one stream the two end counts of an interval run at
the same time, and so do counts at the cuts plan() is expected to choose (linear interpolation of
the end counts); plan() then finds them remembered. Every count is exact, so the certificate is
the same; only the cut positions may differ from the one-stream sequence.
"""
from __future__ import annotations

import dataclasses
import math
import threading
from dataclasses import dataclass

import numpy as np
import torch

from . import config as cfgmod
from . import streams as stream_pool
from .driver_core import (Block, Ctx, DefiniteRun, FactorHeld, PhaseClock, block_metrics, count_below,
                          factorize_near, keep_converged, lanczos_run, rayleigh_ritz_definite, refactorize_at,
                          try_refactorize_at, verify_and_rescue, TRUSTWORTHY)
from .errors import InvalidInputError, NotConvergedError
from .krylov import LARGEST_ALGEBRAIC, LARGEST_MAGNITUDE, SMALLEST_ALGEBRAIC
from .nearest import rerun_block_size

INF = math.inf
CLUSTER_SHRINK = 1.0 / 32.0
EQUAL_WIDTH = 1e-8
EXTRA_RERUN_ROUNDS = 4
MAX_RERUN_ROUNDS = 32
STALL_ROUNDS = 2
MULTIPLET_OFFSET = 0.01
# shifts counted inside an interval together with its ends (with enough worker streams); see
# WindowSolver.blind_points. Five: seven concurrent counts in the first round (box48: 58 ms against
# 29 + 44 ms for the ends and then the interpolated cuts, six slices of 73-93 eigenvalues); six
# made seven smaller slices and a slower round of eight counts (77 ms)
BLIND_POINTS = 5


class _NeedsCount(Exception):
    """A dry run of plan() reached a shift that has not been counted."""


def end_resolution(scale: float, s: float) -> float:
    """definite.cpp:63-65."""
    return 1e-9 * max(scale, abs(s)) if math.isfinite(s) else 0.0


@dataclass
class CountPoint:
    s: float
    nu: int
    detection_off: bool = False
    near: bool = False


@dataclass
class Slice:
    lo: CountPoint
    hi: CountPoint
    want: int = -1
    focus: float = math.nan

    def count(self):
        return self.hi.nu - self.lo.nu

    def required(self):
        return self.want if self.want >= 0 else self.count()


def slice_block_size(cfg: dict, nev: int) -> int:
    """Slices: configured width, else 4 for >= 8 wanted pairs (steps.cpp:1316-1330), 8 from 32 on
    (GPU: a wider block costs little per column)."""
    if cfg["block_size"] > 0:
        return int(cfg["block_size"])
    if nev >= 32:
        return 8
    return 4 if nev >= 8 else 1


class WindowSolver:
    def __init__(self, ctx: Ctx, streams: int = 1):
        self._main = ctx
        self._tls = threading.local()      # a slice worker's own context
        self._lock = threading.Lock()      # the count cache
        self.streams = max(1, int(streams))
        self._slot_list = []
        self.n = ctx.pencil.n
        self.scale = ctx.scale
        self.known: list[CountPoint] = []
        self.slices: list[Slice] = []
        self.reports: list[dict] = []
        self.pool_x = []
        self.pool_lam = []
        self.pool_metrics = []
        self.pool_slice = []
        self.restarts = 0
        self.applies = 0
        # counts made ahead of the solve (a batch's window ends, factorized with null-pivot
        # detection at exactly these shifts): the points count() would have produced there
        for s, nu in ctx.counted:
            self.remember(CountPoint(float(s), int(nu)))
            ctx.event("count", s=float(s), nu=int(nu), near_eigenvalue=False, batched=True)

    @property
    def c(self) -> Ctx:
        """The context of the calling thread: a slice worker's own, else the solve's."""
        return getattr(self._tls, "ctx", None) or self._main

    # ------------------------------------------------------------------ counts
    def minus_inf(self):
        return CountPoint(-INF, 0)

    def plus_inf(self):
        return CountPoint(INF, self.n)

    def remember(self, p: CountPoint):
        for i, q in enumerate(self.known):
            if q.s == p.s:
                self.known[i] = p
                return
        self.known.append(p)
        self.known.sort(key=lambda q: q.s)

    def count(self, s: float, outward: float) -> CountPoint:
        with self._lock:
            for p in self.known:
                if p.s == s:
                    return p
        if getattr(self, "_dry", False):
            raise _NeedsCount(s)
        step = 1e-10 * max(self.scale, abs(s))
        nu, s_used, off, near = count_below(self.c, s, outward, step, s_is_gap=False)
        p = CountPoint(s, nu, off, near)
        with self._lock:
            self.remember(p)
            if s_used != s:
                self.remember(CountPoint(s_used, nu, off, near))
        self.c.event("count", s=s, nu=nu, near_eigenvalue=near)
        return p

    def width_min(self, x, y):
        return 1e-8 * max(self.scale, abs(x), abs(y))

    def expansion_step(self):
        return 1e-3 * max(self.scale, 1e-300)

    def find_count(self, a: CountPoint, b: CountPoint, g_lo: int, g: int, g_hi: int, start: float, side_above: bool,
                   guess: float = math.nan):
        """definite.cpp:1061-1144: (point, cluster, cl_lo, cl_hi)."""
        same_side = last_side = 0
        lw = rw = 1.0
        expansions = 0
        pair_lo = pair_hi = -1
        pair_width = INF
        bisect = False

        def bracket():
            lft, rgt, best = a, b, None
            for p in self.known:
                if not (a.s < p.s < b.s):
                    continue
                if g_lo <= p.nu <= g_hi:
                    if best is None or abs(p.nu - g) < abs(best.nu - g):
                        best = p
                elif p.nu < g_lo:
                    lft = p
                elif rgt.s == b.s:
                    rgt = p
            return best, lft, rgt

        def beside(lft, rgt):
            pt = self.gap_point(rgt, +1.0, b) if side_above else self.gap_point(lft, -1.0, a)
            self.c.event("count_cluster", lower=lft.s, upper=rgt.s, eigenvalues=rgt.nu - lft.nu, point=pt.s)
            return pt, True, lft, rgt

        for _ in range(40):
            best, lft, rgt = bracket()
            if best is not None:
                return best, False, None, None
            lf, rf = math.isfinite(lft.s), math.isfinite(rgt.s)
            if lf and rf:
                if lft.nu != pair_lo or rgt.nu != pair_hi:
                    pair_lo, pair_hi, pair_width = lft.nu, rgt.nu, rgt.s - lft.s
                if rgt.s - lft.s <= max(self.width_min(lft.s, rgt.s), CLUSTER_SHRINK * pair_width):
                    return beside(lft, rgt)
            d = self.expansion_step() * (4.0 ** expansions)
            if math.isfinite(guess) and lft.s < guess < rgt.s:
                s = guess
                guess = math.nan
            elif not lf and not rf:
                s = start
            elif not lf:
                s = rgt.s - d
                expansions += 1
            elif not rf:
                s = lft.s + d
                expansions += 1
            elif bisect:
                s = 0.5 * (lft.s + rgt.s)
            else:
                fl = lw * (lft.nu - g)
                fr = rw * (rgt.nu - g)
                frac = min(max(-fl / (fr - fl), 0.05), 0.95)
                s = lft.s + frac * (rgt.s - lft.s)
            p = self.count(s, +1.0)
            bisect = lf and rf and (p.nu == lft.nu or p.nu == rgt.nu)
            sd = -1 if p.nu < g_lo else (1 if p.nu > g_hi else 0)
            same_side = same_side + 1 if (sd != 0 and sd == last_side) else 1
            last_side = sd
            if same_side >= 2 and sd < 0:
                rw *= 0.5
            if same_side >= 2 and sd > 0:
                lw *= 0.5
            if same_side == 1:
                lw = rw = 1.0
        best, lft, rgt = bracket()
        if best is not None:
            return best, False, None, None
        if not (math.isfinite(lft.s) and math.isfinite(rgt.s)):
            raise NotConvergedError(f"inertia search found no point with {g} eigenvalues below it")
        return beside(lft, rgt)

    def gap_point(self, p: CountPoint, direction: float, limit: CountPoint) -> CountPoint:
        """definite.cpp:1146-1190."""
        far, nxt = p, limit
        seq = self.known if direction > 0 else list(reversed(self.known))
        for q in seq:
            if direction > 0 and not (p.s < q.s < limit.s):
                continue
            if direction < 0 and not (limit.s < q.s < p.s):
                continue
            if q.nu != p.nu:
                nxt = q
                break
            far = q
        if nxt.s == limit.s and math.isfinite(limit.s) and limit.nu == p.nu:
            far = limit

        def middle(x):
            m = CountPoint(0.5 * (p.s + x), p.nu)
            self.remember(m)
            return m

        if far.s != p.s:
            return middle(far.s)
        delta = 0.5 * abs(nxt.s - p.s) if math.isfinite(nxt.s) else self.expansion_step()
        for _ in range(3):
            q = self.count(p.s + direction * delta, direction)
            if q.nu == p.nu and direction * (q.s - p.s) > 0:
                return middle(q.s)
            delta *= 0.25
        return CountPoint(p.s, p.nu, p.detection_off, True)

    def plan(self, w: Slice, c_target: int, c_max: int):
        """definite.cpp:1192-1229."""
        a, b = w.lo, w.hi
        if w.want >= 0:
            return [w]
        cuts = [a]
        cur = a
        while b.nu - cur.nu > c_max:
            remaining = b.nu - cur.nu
            step = (remaining + 1) // 2 if remaining <= 2 * c_target else c_target
            g_lo = cur.nu + max(1, step // 2)
            g_hi = cur.nu + min(c_max, remaining - 1)
            start = cur.s if math.isfinite(cur.s) else b.s
            guess = math.nan
            if len(cuts) >= 2 and math.isfinite(cuts[-2].s) and math.isfinite(cur.s) and cur.nu > cuts[-2].nu:
                prev = cuts[-2]
                guess = cur.s + (cur.s - prev.s) * step / (cur.nu - prev.nu)
            elif math.isfinite(cur.s) and math.isfinite(b.s) and b.s > cur.s:
                guess = cur.s + (b.s - cur.s) * step / (b.nu - cur.nu)
            p, _, _, _ = self.find_count(cur, b, g_lo, cur.nu + step, max(g_lo, g_hi), start, True, guess)
            if p.nu <= cur.nu or not (p.s < b.s):
                break
            cuts.append(p)
            cur = p
        cuts.append(b)
        return [Slice(cuts[i], cuts[i + 1]) for i in range(len(cuts) - 1) if cuts[i + 1].nu > cuts[i].nu]

    def anchor_of(self, s: Slice) -> str:
        if not math.isfinite(s.hi.s):
            return "lower"
        if not math.isfinite(s.lo.s):
            return "upper"
        if s.lo.nu == 0:
            return "centre" if s.hi.near else "upper"
        if s.hi.nu == self.n:
            return "centre" if s.lo.near else "lower"
        return "centre"

    # ------------------------------------------------------------------ slices
    def factorize_anchor(self, p: CountPoint):
        if try_refactorize_at(self.c, p.s, p.detection_off):
            return p.s, p.detection_off
        self.c.warn(f"one-sided shift {p.s:.17g} refactorized as a near-singular shift")
        return factorize_near(self.c, p.s)

    def multiplet_shift(self, sel: Block, s: Slice, tried):
        l = sorted(x for x in sel.lam if s.lo.s <= x <= s.hi.s)
        best, best_size = math.nan, 1
        i = 0
        while i < len(l):
            j = i + 1
            while j < len(l) and l[j] - l[j - 1] <= EQUAL_WIDTH * max(self.scale, abs(l[j])):
                j += 1
            mean = sum(l[i:j]) / (j - i)
            seen = any(abs(t - mean) <= EQUAL_WIDTH * max(self.scale, abs(mean)) for t in tried)
            if j - i > best_size and not seen:
                best_size, best = j - i, mean
            i = j
        return best

    def multiplet_rerun_shift(self, mean, sel: Block, s: Slice):
        gl = gr = INF
        for x in sel.lam:
            d = x - mean
            if abs(d) <= EQUAL_WIDTH * max(self.scale, abs(mean)):
                continue
            if d > 0:
                gr = min(gr, d)
            else:
                gl = min(gl, -d)
        if not math.isfinite(gl) and math.isfinite(s.lo.s):
            gl = mean - s.lo.s
        if not math.isfinite(gr) and math.isfinite(s.hi.s):
            gr = s.hi.s - mean
        gap = gr if (math.isfinite(gr) and (gr >= gl or not math.isfinite(gl))) else -gl
        return mean + MULTIPLET_OFFSET * gap if math.isfinite(gap) else mean

    def run_slice(self, s: Slice, sid: int):
        c = self.c
        cfg = c.cfg
        count = s.count()
        which = LARGEST_MAGNITUDE
        anchor = self.anchor_of(s)
        if anchor == "centre":
            sigma, det_off = factorize_near(c, 0.5 * (s.lo.s + s.hi.s))
        elif anchor == "upper":
            p = self.gap_point(s.hi, +1.0, self.plus_inf()) if s.hi.near else s.hi
            sigma, det_off = self.factorize_anchor(p)
            which = SMALLEST_ALGEBRAIC
        else:
            p = self.gap_point(s.lo, -1.0, self.minus_inf()) if s.lo.near else s.lo
            sigma, det_off = self.factorize_anchor(p)
            which = LARGEST_ALGEBRAIC
        sr = {"lower": s.lo.s, "upper": s.hi.s, "sigma": sigma, "which": which, "count_inertia": count,
              "count_found": -1, "nev": 0, "reruns": 0, "restarts": 0, "op_applies": 0, "passed": False}
        d_lo, d_hi = end_resolution(self.scale, s.lo.s), end_resolution(self.scale, s.hi.s)
        run = DefiniteRun()
        st = {"sel": None, "metrics": None, "inside": 0, "near_in": 0, "near_out": 0}
        shift_state = {"sigma": sigma, "det_off": det_off}

        def select():
            idx = [j for j in range(run.all.size) if math.isfinite(run.all.lam[j])
                   and s.lo.s - d_lo <= run.all.lam[j] <= s.hi.s + d_hi]
            sel = run.all.take(idx)
            sel, metrics = verify_and_rescue(c, sel, c.n2a, c.n2b,
                                             lambda: refactorize_at(c, shift_state["sigma"], shift_state["det_off"]))
            keep = []
            inside = near_in = near_out = 0
            for j in range(sel.size):
                if not (metrics[j].backward_error <= max(TRUSTWORTHY, cfg["tol"])):
                    continue
                l = sel.lam[j]
                if l < s.lo.s - d_lo or l > s.hi.s + d_hi:
                    continue
                keep.append(j)
                if s.lo.s <= l <= s.hi.s:
                    inside += 1
                    if l < s.lo.s + d_lo or l > s.hi.s - d_hi:
                        near_in += 1
                else:
                    near_out += 1
            st["sel"] = sel.take(keep)
            st["metrics"] = [metrics[j] for j in keep]
            st["inside"], st["near_in"], st["near_out"] = inside, near_in, near_out

        def missing():
            return count - st["inside"] - st["near_out"]

        sr["nev"] = min(self.n, s.required())
        p = slice_block_size(cfg, sr["nev"])
        if sid == 0:
            c.rep["block_size"] = p
        lanczos_run(c, sigma, which, sr["nev"], p, None, run)
        with c.clock.phase("verify"):
            select()
        tried = []
        stalls = 0
        max_rounds = min(MAX_RERUN_ROUNDS, missing() + EXTRA_RERUN_ROUNDS)
        rnd = 0
        while missing() > 0 and rnd < max_rounds and stalls < STALL_ROUNDS:
            miss = missing()
            keep_converged(run, c.n2a, c.n2b, cfg["tol"])
            nev_rerun = min(self.n - run.all.size, miss + 2)
            if nev_rerun < 1:
                break
            m_sigma = self.multiplet_shift(st["sel"], s, tried)
            c.event("slice_rerun", slice=sid, missing=miss, multiplet=m_sigma)
            defl = run.all.x.contiguous()
            p_rerun = rerun_block_size(cfg, p, miss)
            if math.isfinite(m_sigma):
                tried.append(m_sigma)
                ms, off = factorize_near(c, self.multiplet_rerun_shift(m_sigma, st["sel"], s), on_eigenvalue=True)
                shift_state["sigma"], shift_state["det_off"] = ms, off
                lanczos_run(c, ms, LARGEST_MAGNITUDE, nev_rerun, p_rerun, defl if defl.shape[0] else None, run)
            else:
                refactorize_at(c, sigma, det_off)
                shift_state["sigma"], shift_state["det_off"] = sigma, det_off
                lanczos_run(c, sigma, which, nev_rerun, p_rerun, defl if defl.shape[0] else None, run)
            sr["reruns"] += 1
            with c.clock.phase("verify"):
                select()
            stalls = 0 if missing() < miss else stalls + 1
            rnd += 1
        inside = st["inside"]
        sr["count_found"] = inside
        sr["passed"] = inside - st["near_in"] <= count <= inside + st["near_out"]
        sr["restarts"] = run.restarts
        sr["op_applies"] = run.applies
        c.event("slice", slice=sid, lower=s.lo.s, upper=s.hi.s, sigma=sigma, which=which, count=count,
                found=inside, passed=sr["passed"], reruns=sr["reruns"], restarts=run.restarts,
                op_applies=run.applies)
        if not sr["passed"]:
            c.warn(f"slice [{s.lo.s:.17g}, {s.hi.s:.17g}]: inertia {s.required()} vs found {inside}")
        self.reports[sid] = sr
        sel = st["sel"]
        self.slice_pool[sid] = ([sel.x[j:j + 1] for j in range(sel.size)], [float(l) for l in sel.lam[:sel.size]],
                                list(st["metrics"]))

    # ------------------------------------------------------------------ slices
    def run_slices(self):
        """Every slice, then their pairs pooled in slice order."""
        ns = len(self.slices)
        self.reports = [None] * ns
        self.slice_pool = [None] * ns
        workers = min(ns, self.streams)
        if workers <= 1:
            for i, sl in enumerate(self.slices):
                self.run_slice(sl, i)
        else:
            self.run_slices_concurrently(workers)
        for sid in range(ns):
            xs, lams, mets = self.slice_pool[sid]
            self.pool_x.extend(xs)
            self.pool_lam.extend(lams)
            self.pool_metrics.extend(mets)
            self.pool_slice.extend([sid] * len(xs))
            self.restarts += self.reports[sid]["restarts"]
            self.applies += self.reports[sid]["op_applies"]

    def run_slices_concurrently(self, workers: int):
        """Slices handed out in order to the worker slots; the block size of slice 0 is the solve's."""
        ctxs = self._concurrently(len(self.slices), lambda sid: self.run_slice(self.slices[sid], sid), "iterate")
        main = self._main
        if ctxs and ctxs[0] is not None:
            main.rep["block_size"] = ctxs[0].rep["block_size"]
        par = main.rep["parallel"]
        par["helper_slices"] = len(self.slices)
        par["notes"].append(f"{len(self.slices)} spectrum slices on {workers} concurrent streams of one device")

    # ------------------------------------------------------------------ workers
    def _slots(self, k: int):
        """k worker slots {fac, stream, held}. Slot 0 is the solve's own factorization object; the
        others are clones over the same analysis, made once per solve and reused by every
        concurrent phase (counts, slices), each with a stream of its own."""
        main = self._main
        while len(self._slot_list) < k:
            (st,) = stream_pool.checkout(1)
            fac = main.fac if not self._slot_list else main.fac.clone()
            self._slot_list.append({"fac": fac, "stream": st, "held": FactorHeld()})
        return self._slot_list[:k]

    def release_slots(self):
        """The end of the solve: the clones' factorizations are freed and the slots' streams go back
        to the process pool (slot 0's factorization belongs to the caller)."""
        torch.cuda.synchronize()
        for s in self._slot_list[1:]:
            s["fac"].close()
        stream_pool.checkin([s["stream"] for s in self._slot_list])
        for s in self._slot_list:
            s["stream"] = None

    def _concurrently(self, njobs: int, job, phase: str):
        """Runs job(i) for i < njobs on min(njobs, streams) threads, one worker slot each, jobs handed
        out in order. Every job runs in a context of its own (self.c inside the job: the slot's
        factorization, a fresh report and event list); the jobs' events, warnings and certificate
        fields are merged into the solve's in job order afterwards, so the outcome does not depend on
        which job finished first. The first failure in job order is re-raised. The waiting is charged
        to `phase` of the solve's clock. Returns the job contexts."""
        main = self._main
        workers = max(1, min(njobs, self.streams))
        slots = self._slots(workers)
        slots[0]["held"] = main.held
        torch.cuda.synchronize()
        ctxs = [None] * njobs
        errors = [None] * njobs
        nxt = [0]
        order_lock = threading.Lock()

        def worker(w: int):
            slot = slots[w]
            with torch.cuda.stream(slot["stream"]):
                while True:
                    with order_lock:
                        i = nxt[0]
                        nxt[0] += 1
                        if any(e is not None for e in errors):
                            i = njobs  # a job failed: hand out nothing more
                    if i >= njobs:
                        break
                    rep = cfgmod.new_report()
                    rep["certificate"]["note"] = ""
                    ctx = dataclasses.replace(main, fac=slot["fac"], held=slot["held"], clock=PhaseClock(), events=[],
                                              rep=rep, certified=[])
                    ctxs[i] = ctx
                    self._tls.ctx = ctx
                    try:
                        job(i)
                    except BaseException as e:  # re-raised in job order below
                        errors[i] = e
                    finally:
                        slot["held"] = ctx.held
                        self._tls.ctx = None
                    if errors[i] is not None:
                        break
                slot["stream"].synchronize()

        with main.clock.phase(phase):
            threads = [threading.Thread(target=worker, args=(w,), name=f"hsolve-worker-{w}") for w in range(workers)]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        torch.cuda.synchronize()
        main.held = slots[0]["held"]
        ms = main.fac.stats
        for slot in slots[1:]:
            st = slot["fac"].stats
            ms.factorizations += st.factorizations
            ms.solves += st.solves
            ms.rhs_solved += st.rhs_solved
            ms.factorize_s += st.factorize_s
            ms.solve_s += st.solve_s
            st.factorizations = st.solves = st.rhs_solved = 0
            st.factorize_s = st.solve_s = 0.0
        for i in range(njobs):
            if errors[i] is not None:
                raise errors[i]
            ctx = ctxs[i]
            if ctx is None:
                continue
            r = ctx.rep
            main.events.extend(ctx.events)
            main.rep["warnings"].extend(r["warnings"])
            main.rep["shift_retries"] = max(main.rep["shift_retries"], r["shift_retries"])
            main.rep["certificate"]["endpoint_moves"] += r["certificate"]["endpoint_moves"]
            if r["certificate"]["note"] and not main.rep["certificate"]["note"]:
                main.rep["certificate"]["note"] = r["certificate"]["note"]
        return ctxs

    def count_all(self, points):
        """Counts at several (shift, outward) points; with more than one stream the ones not yet
        known are counted at the same time on the worker slots (one FP64 factorization each).
        (Batched factorizations of these shifts, GpuFactorization.count_batch, were measured no
        faster on box48 than the concurrent streams: both are bound by the same launch latencies.)"""
        with self._lock:
            known = {p.s for p in self.known}
        todo = list(dict.fromkeys(s for s, _ in points if s not in known))
        if self.streams > 1 and len(todo) > 1:
            side = {s: o for s, o in points}
            self._concurrently(len(todo), lambda i: self.count(todo[i], side[todo[i]]), "certificate")
        return [self.count(s, o) for s, o in points]

    def blind_points(self, lower: float, upper: float) -> list:
        """Shifts counted in the same round as an interval's ends, before anything about the
        window's eigenvalues is known: as many as the worker streams beyond the two ends hold (at
        most BLIND_POINTS), equally spaced inside the window. plan() takes its cuts among them when
        they fall within its size limits; otherwise a second round counts the interpolated cuts."""
        k = min(BLIND_POINTS, self.streams - 2)
        if k <= 0 or not (math.isfinite(lower) and math.isfinite(upper)) or not upper > lower:
            return []
        return [lower + (upper - lower) * j / (k + 1) for j in range(1, k + 1)]

    def plan_from_known(self, w: Slice, c_target: int, c_max: int) -> bool:
        """Whether plan() finds every cut of w among the points counted so far (a dry run that
        counts nothing and leaves the known points as they were)."""
        with self._lock:
            saved = list(self.known)
        self._dry = True
        try:
            self.plan(w, c_target, c_max)
            return True
        except _NeedsCount:
            return False
        finally:
            self._dry = False
            with self._lock:
                self.known = saved

    def speculative_cuts(self, w: Slice, c_target: int, c_max: int):
        """The shifts plan() is likely to count at first: its sequence of eigenvalue-count targets
        over w, each placed by linear interpolation between the counts at w's ends (the guess plan()
        itself starts from). Counting them ahead, at the same time, leaves plan() mostly with cache
        hits; where a guess misses, plan() searches on from the known points as usual."""
        a, b = w.lo, w.hi
        if w.want >= 0 or not (math.isfinite(a.s) and math.isfinite(b.s)) or b.nu <= a.nu:
            return []
        out = []
        cur = a.nu
        while b.nu - cur > c_max:
            remaining = b.nu - cur
            cur += (remaining + 1) // 2 if remaining <= 2 * c_target else c_target
            out.append(a.s + (b.s - a.s) * (cur - a.nu) / (b.nu - a.nu))
        # where the density grows over the window the cuts hold fewer eigenvalues than their
        # targets, more than c_max remain after the last one and plan() splits that stretch once
        # more: its midpoint is counted ahead as well
        if out and b.nu - cur > c_target // 2:
            out.append(0.5 * (out[-1] + b.s))
        return out

    def deduplicate(self) -> int:
        """definite.cpp:1522-1556: groups within 2e-9 relative holding pairs of several slices are
        replaced by a B-orthonormal basis of their span (B-Gram directions below 1e-4 dropped)."""
        m = len(self.pool_lam)
        order = sorted(range(m), key=lambda q: self.pool_lam[q])
        drop = [False] * m
        removed = 0
        i = 0
        while i < m:
            j = i + 1
            multi = False
            while j < m:
                l0, l1 = self.pool_lam[order[j - 1]], self.pool_lam[order[j]]
                if l1 - l0 > 2e-9 * max(self.scale, abs(l0), abs(l1)):
                    break
                multi = multi or any(self.pool_slice[order[q]] != self.pool_slice[order[j]] for q in range(i, j))
                j += 1
            if multi:
                g = j - i
                grp = Block(self.c.ops, torch.cat([self.pool_x[order[i + q]] for q in range(g)], 0),
                            [self.pool_lam[order[i + q]] for q in range(g)])
                grp.recompute_products()
                r = rayleigh_ritz_definite(grp, 1e-4)
                if r < g:
                    mr = block_metrics(grp, self.c.n2a, self.c.n2b)
                    for q in range(g):
                        col = order[i + q]
                        if q < r:
                            self.pool_x[col] = grp.x[q:q + 1].clone()
                            self.pool_lam[col] = float(grp.lam[q])
                            self.pool_metrics[col] = mr[q]
                        else:
                            drop[col] = True
                    removed += g - r
            i = j
        if removed:
            keep = [q for q in range(m) if not drop[q]]
            self.pool_x = [self.pool_x[q] for q in keep]
            self.pool_lam = [self.pool_lam[q] for q in keep]
            self.pool_metrics = [self.pool_metrics[q] for q in keep]
            self.pool_slice = [self.pool_slice[q] for q in keep]
        return removed

    def assign_to_windows(self, windows):
        """definite.cpp:66-134 (no zero band): pairs owned by the window holding them; a window
        short of its count takes pairs numerically on its ends only where a count beyond the end
        shows nothing there."""
        lam = self.pool_lam
        m = len(lam)
        owner = [-1] * m
        found = [0] * len(windows)
        for j in range(m):
            for w, win in enumerate(windows):
                if win.lo.s <= lam[j] <= win.hi.s:
                    if win.want < 0 or found[w] < win.required():
                        owner[j] = w
                        found[w] += 1
                    break
        for w, win in enumerate(windows):
            if win.want >= 0 or found[w] == win.required():
                continue
            d_lo, d_hi = end_resolution(self.scale, win.lo.s), end_resolution(self.scale, win.hi.s)
            short = found[w] < win.required()
            on_end = []
            for j in range(m):
                l = lam[j]
                if short:
                    if owner[j] >= 0:
                        continue
                    if win.lo.s - d_lo <= l < win.lo.s:
                        on_end.append((win.lo.s - l, j, False))
                    elif win.hi.s < l <= win.hi.s + d_hi:
                        on_end.append((l - win.hi.s, j, True))
                elif owner[j] == w:
                    if l < win.lo.s + d_lo:
                        on_end.append((l - win.lo.s, j, False))
                    elif l > win.hi.s - d_hi:
                        on_end.append((win.hi.s - l, j, True))
            on_end.sort(key=lambda t: t[0])
            clear = {False: None, True: None}
            for dist, j, upper in on_end:
                if found[w] == win.required():
                    break
                if short:
                    if clear[upper] is None:
                        end = win.hi if upper else win.lo
                        direction = 1.0 if upper else -1.0
                        beyond = end.s + direction * 2.0 * end_resolution(self.scale, end.s)
                        clear[upper] = self.count(beyond, direction).nu == end.nu
                    if not clear[upper]:
                        continue
                    owner[j] = w
                    found[w] += 1
                else:
                    owner[j] = -1
                    found[w] -= 1
        return [j for j in range(m) if owner[j] >= 0], found

    # ------------------------------------------------------------------ solve
    def solve(self):
        c = self.c
        cfg, rep = c.cfg, c.rep
        t = cfg["target"]
        mode = t["mode"]
        n = self.n
        if mode != "interval" and t["k"] > n:
            raise InvalidInputError(f"target.k = {t['k']} exceeds the matrix order {n}")
        c_target = cfg["slice_size"] if cfg["slice_size"] > 0 else 96
        c_max = cfg["slice_size"] if cfg["slice_size"] > 0 else c_target + c_target // 4
        windows = []
        cert = rep["certificate"]
        k = t["k"]
        if mode == "interval":
            blind = [(x, +1.0) for x in self.blind_points(t["lower"], t["upper"])]
            lo, hi = self.count_all([(t["lower"], -1.0), (t["upper"], +1.0)] + blind)[:2]
            windows.append(Slice(lo, hi))
            cert["lower"], cert["upper"] = lo.s, hi.s
        elif mode == "smallest":
            a, b = self.minus_inf(), self.plus_inf()
            g_lo = k
            g_hi = min(g_lo + 16, max(g_lo, n - 1))
            p, clus, cl_lo, cl_hi = self.find_count(a, b, g_lo, min(g_lo + 8, g_hi), g_hi, 0.0, True)
            windows.append(Slice(a, p))
        else:
            a, b = self.minus_inf(), self.plus_inf()
            g_hi = n - k
            g_lo = max(g_hi - 16, min(g_hi, 1))
            p, clus, cl_lo, cl_hi = self.find_count(a, b, g_lo, max(g_hi - 8, g_lo), g_hi, 0.0, False)
            windows.append(Slice(p, b))
        # open ends of one-sided windows: close at a point beyond the spectrum end
        for w in windows:
            if not math.isfinite(w.lo.s):
                w.lo = self.find_count(self.minus_inf(), w.hi, 0, 0, 0, w.hi.s, False)[0]
            if not math.isfinite(w.hi.s):
                w.hi = self.find_count(w.lo, self.plus_inf(), n, n, n, w.lo.s, True)[0]
        total = sum(w.count() for w in windows)
        required = sum(w.required() for w in windows)
        cert["performed"] = True
        cert["method"] = "exact"
        if mode == "smallest":
            cert["lower"], cert["upper"] = -INF, windows[-1].hi.s
        elif mode == "largest":
            cert["lower"], cert["upper"] = windows[0].lo.s, INF
        cert["count_inertia"] = required
        if mode != "interval" and required < k:
            c.warn(f"the target window holds only {required} eigenvalue(s), fewer than k = {k}")
        if total == 0:
            cert["count_found"] = 0
            cert["passed"] = True
            cert["note"] = cert["note"] or "window is empty"
            empty = Block(c.ops, None, [])
            return empty, [], mode == "interval"
        if self.streams > 1:
            # a second round of counts only for windows whose cuts the known points do not supply
            todo = [w for w in windows if not self.plan_from_known(w, c_target, c_max)]
            self.count_all([(x, +1.0) for w in todo for x in self.speculative_cuts(w, c_target, c_max)])
        for w in windows:
            self.slices.extend(self.plan(w, c_target, c_max))
        c.event("slices", slices=len(self.slices))
        self.run_slices()
        with c.clock.phase("verify"):
            self.deduplicate()
            inn, found = self.assign_to_windows(windows)
        cert["count_found"] = len(inn)
        cert["passed"] = all(found[w] == windows[w].required() for w in range(len(windows)))
        cert["rigorous"] = cert["passed"]
        cert["failure_probability"] = 0.0 if cert["passed"] else 1.0
        cert["statement"] = (f"double-precision Sylvester counts at every window end and slice cut: {required} "
                             f"eigenvalue(s) required, {len(inn)} distinct found pair(s)")
        if not cert["passed"]:
            c.warn(f"certificate failed: inertia {cert['count_inertia']} vs found {cert['count_found']} over "
                   f"{len(self.slices)} slice(s)")
        if len(self.slices) > 1:
            cert["note"] = f"{len(self.slices)} slices" + ("; " + cert["note"] if cert["note"] else "")
        inn.sort(key=lambda q: self.pool_lam[q])
        if mode == "largest":
            inn.reverse()
        if mode != "interval" and len(inn) > k:
            inn = inn[:k]
        rep["slices"] = [r for r in self.reports if r is not None]
        rep["restarts"] = self.restarts
        rep["op_applies"] = self.applies
        if rep["slices"]:
            rep["sigma_used"] = rep["slices"][0]["sigma"]
        if inn:
            x = torch.cat([self.pool_x[q] for q in inn], 0)
        else:
            x = None
        sel = Block(c.ops, x, [self.pool_lam[q] for q in inn])
        metrics = [self.pool_metrics[q] for q in inn]
        want_out = total if mode == "interval" else k
        ok = len(inn) == want_out and cert["passed"]
        if len(inn) != (total if mode == "interval" else min(k, required)):
            c.warn(f"only {len(inn)} of {total if mode == 'interval' else min(k, required)} pairs found")
        return sel, metrics, ok


def solve_window(ctx: Ctx, streams: int = 1):
    w = WindowSolver(ctx, streams)
    try:
        return w.solve()
    finally:
        w.release_slots()
