"""Building blocks of the HEP / GHEP driver: phases, norms, candidate blocks, Rayleigh-Ritz,
shift handling and inertia counts.

Restated from the reference driver:
  * PhaseClock: exclusive phases setup / factor / iterate / verify / certificate
    (src/driver/internal.hpp:48-70; they add up to time_total_s, src/driver/solver.cpp:682-688);
  * estimate_norm2: 30 plain Lanczos steps from a seeded start, max |Ritz| as the lower bound of
    ||M||_2 used in the backward error (src/driver/norms.cpp:80, include/hsolve/solver.hpp:7-12);
  * Block + rayleigh_ritz_definite: rank-revealing Rayleigh-Ritz on the original pencil, dropping
    B-Gram directions below 1e-8 of the largest (src/driver/projection.cpp:89-130,
    src/driver/definite.cpp:143-150), pair_metrics (projection.cpp:53-70);
  * factorize_near / refactorize_at: the singular-shift rules (src/driver/steps.cpp:243-327);
  * count_below: Sylvester counts with the null-pivot retries (src/driver/steps.cpp:549-568);
  * lanczos_run: one shift-invert Krylov run appended to the candidate block
    (src/driver/definite.cpp:213-309).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
import torch

from . import _lib
from .errors import InternalError, SingularError
from .factor import GpuFactorization, Pencil, ptr, stream_ptr
from .krylov import KrylovOptions, lanczos_krylov_schur
from .ops import DeviceOps, InexactFactorFailure, ShiftInvertOperator, SolveRefinement

EPS = float(np.finfo(np.float64).eps)
TRUSTWORTHY = 1e-6          # definite.cpp:145
DEPENDENT_GRAM = 1e-8       # definite.cpp:150
INEXACT_REFINE_STEPS = 10   # internal.hpp:235
INEXACT_REFINE_TOL_FACTOR = 0.1  # internal.hpp:236


# ------------------------------------------------------------------------------------ phases

class PhaseClock:
    PHASES = ("setup", "factor", "iterate", "verify", "certificate")

    def __init__(self):
        self.t = {p: 0.0 for p in self.PHASES}
        self.stack = []
        self.last = time.perf_counter()
        self.start = self.last

    def _switch(self):
        now = time.perf_counter()
        if self.stack:
            self.t[self.stack[-1]] += now - self.last
        self.last = now

    def push(self, phase):
        self._switch()
        self.stack.append(phase)

    def pop(self):
        self._switch()
        self.stack.pop()

    def phase(self, name):
        clock = self

        class _Scope:
            def __enter__(self_inner):
                clock.push(name)

            def __exit__(self_inner, *exc):
                clock.pop()
                return False

        return _Scope()

    def total(self):
        self._switch()
        return sum(self.t.values())


# ------------------------------------------------------------------------------------ context

@dataclass
class FactorHeld:
    shift: float = math.nan
    detection_off: bool = False
    precision: str = ""


@dataclass
class Ctx:
    cfg: dict
    rep: dict
    clock: PhaseClock
    pencil: Pencil
    ops: DeviceOps
    fac: GpuFactorization
    n2a: float = 1.0
    n2b: float = 1.0
    scale: float = 1.0
    held: FactorHeld = field(default_factory=FactorHeld)
    precision: str = "double"      # precision of the working (Krylov) factors
    streams: int = 1               # concurrent streams for spectrum slices
    events: list = field(default_factory=list)
    certified: list = field(default_factory=list)
    counted: list = field(default_factory=list)  # (shift, nu): exact counts made ahead (batched)

    def warn(self, msg: str):
        self.rep["warnings"].append(msg)

    def event(self, name: str, **kw):
        kw["event"] = name
        self.events.append(kw)


# ------------------------------------------------------------------------------------ norms

def estimate_norm2(ops: DeviceOps, apply, steps: int = 30, seed: int = 2026) -> float:
    """Largest |Ritz value| of `steps` plain Lanczos steps (norms.cpp:60-95). GPU products and dots,
    full reorthogonalization of the few vectors, tridiagonal eigenvalues on the host.

    For A and B (apply = ops.apply_a / ops.apply_b) the whole recurrence is one library call
    (hsg_lanczos_coefficients): it runs on the device without reading anything back, and the
    coefficients are read once at the end, where the breakdown test truncates the tridiagonal at
    the first beta_j <= 1e-13 max(|alpha_j|, beta_{j-1}) exactly as a loop stopping there would
    (the steps before a breakdown do not depend on the ones after it). Other operators run the
    same recurrence step by step from Python."""
    n = ops.n
    steps = min(steps, n)
    rng = np.random.default_rng([seed, 0x6E6F726D])
    v = ops.block(rng.uniform(-1.0, 1.0, (1, n)))
    v.mul_(1.0 / math.sqrt(float(ops.dots(v, v).cpu()[0])))
    basis = torch.empty((steps + 1, n), dtype=ops.dtype, device=ops.dev)
    basis[0:1].copy_(v)
    matrix = ops.lanczos_matrix(apply)
    if matrix is not None:
        coef = torch.zeros(2 * steps, dtype=torch.float64, device=ops.dev)
        work = torch.empty(2 * (steps + 1), dtype=torch.float64, device=ops.dev)
        rowptr, col, val, diag = matrix
        _lib.call("hsg_lanczos_coefficients", int(ops.is_complex), n, ptr(rowptr), ptr(col), ptr(val), ptr(diag),
                  ptr(basis), steps, ptr(coef), ptr(work), stream_ptr())
        ab = coef.cpu().numpy()
        a_all, b_all = ab[:steps], ab[steps:]
    else:
        a_all, b_all = np.zeros(steps), np.zeros(steps)
        for j in range(steps):
            w = apply(basis[j:j + 1])
            a_all[j] = float(ops.dots(basis[j:j + 1], w).cpu()[0])
            for _ in range(2):
                c = ops.gram(basis[:j + 1], w)
                ops.update(basis[:j + 1], c, w, alpha=-1.0, beta=1.0)
            if j + 1 == steps:
                break
            b_all[j] = math.sqrt(max(float(ops.dots(w, w).cpu()[0]), 0.0))
            if b_all[j] <= 1e-13 * max(abs(a_all[j]), b_all[j - 1] if j else 0.0, 1e-300):
                break
            basis[j + 1:j + 2].copy_(w * (1.0 / b_all[j]))
    alpha, beta = [], []
    for j in range(steps):
        aj = float(a_all[j])
        alpha.append(aj)
        if j + 1 == steps:
            break
        bj = float(b_all[j])
        scale = max(abs(aj), beta[-1] if beta else 0.0)
        if bj <= 1e-13 * max(scale, 1e-300):
            break
        beta.append(bj)
    T = np.diag(alpha)
    for i, b in enumerate(beta[:len(alpha) - 1]):
        T[i + 1, i] = T[i, i + 1] = b
    w = np.linalg.eigvalsh(T)
    return float(np.max(np.abs(w))) if w.size else 0.0


# ------------------------------------------------------------------------------------ blocks

class Block:
    """Candidate pairs X, AX, BX (device, (m, n)) and lambda (host)."""

    def __init__(self, ops: DeviceOps, x: torch.Tensor | None = None, lam=None):
        self.ops = ops
        n = ops.n
        self.x = x if x is not None else torch.empty((0, n), dtype=ops.dtype, device=ops.dev)
        self.lam = np.asarray(lam if lam is not None else [], dtype=np.float64)
        self.ax = None
        self.bx = None

    @property
    def size(self):
        return int(self.x.shape[0])

    def recompute_products(self):
        self.ax = self.ops.apply_a(self.x) if self.size else self.x.clone()
        self.bx = self.ops.apply_b(self.x) if self.size else self.x.clone()

    def take(self, idx) -> "Block":
        idx = np.asarray(idx, dtype=np.int64)
        t = torch.from_numpy(idx).to(self.ops.dev)
        b = Block(self.ops, self.x.index_select(0, t).contiguous(), self.lam[idx])
        b.ax = self.ax.index_select(0, t).contiguous()
        b.bx = self.bx.index_select(0, t).contiguous()
        return b

    def append(self, other: "Block"):
        if self.size == 0:
            self.x, self.ax, self.bx, self.lam = other.x, other.ax, other.bx, other.lam.copy()
            return
        self.x = torch.cat([self.x, other.x], 0)
        self.ax = torch.cat([self.ax, other.ax], 0)
        self.bx = torch.cat([self.bx, other.bx], 0)
        self.lam = np.concatenate([self.lam, other.lam])


def _replace_by_product(ops: DeviceOps, m: torch.Tensor, coef: np.ndarray) -> torch.Tensor:
    out = ops.empty(coef.shape[1])
    ops.update(m, ops.coef(coef), out, 1.0, 0.0)
    return out


def rayleigh_ritz_definite(blk: Block, drop_tol: float = DEPENDENT_GRAM) -> int:
    """projection.cpp:89-130: X^H B X eigen-decomposition, directions below drop_tol * mu_max
    dropped, then X^H A X on the B-orthonormal rest; X, AX, BX rotated, lambda ascending."""
    m = blk.size
    if m == 0:
        blk.lam = np.zeros(0)
        return 0
    ops = blk.ops
    g = ops.gram(blk.x, blk.ax).cpu().numpy().T
    h = ops.gram(blk.x, blk.bx).cpu().numpy().T
    g = 0.5 * (g + g.conj().T)
    h = 0.5 * (h + h.conj().T)
    mu, V = np.linalg.eigh(h)
    keep = [q for q in range(m) if mu[q] > drop_tol * mu[-1] and mu[q] > 0.0]
    r = len(keep)
    Y = V[:, keep] / np.sqrt(mu[keep])
    gr = Y.conj().T @ g @ Y
    gr = 0.5 * (gr + gr.conj().T)
    theta, Z = np.linalg.eigh(gr)
    coef = Y @ Z
    blk.x = _replace_by_product(ops, blk.x, coef)
    blk.ax = _replace_by_product(ops, blk.ax, coef)
    blk.bx = _replace_by_product(ops, blk.bx, coef)
    blk.lam = theta.copy()
    return r


@dataclass
class PairMetrics:
    abs_residual: float
    rel_residual: float
    backward_error: float


def block_metrics(blk: Block, n2a: float, n2b: float) -> list:
    if blk.size == 0:
        return []
    lam = np.where(np.isfinite(blk.lam), blk.lam, 0.0)
    s = blk.ops.pair_sums(blk.ax, blk.bx, blk.x, lam)
    out = []
    for j in range(blk.size):
        r2, b2, x2 = s[j]
        ar = math.sqrt(max(r2, 0.0))
        l = abs(blk.lam[j])
        if not math.isfinite(blk.lam[j]):
            out.append(PairMetrics(math.inf, math.inf, math.inf))
            continue
        rel = ar / max((l if l > 0 else 1.0) * math.sqrt(b2), 1e-300)
        be = ar / max((n2a + l * n2b) * math.sqrt(x2), 1e-300)
        out.append(PairMetrics(ar, rel, be))
    return out


# ------------------------------------------------------------------------------------ shifts

def factor_precision_code(ctx: Ctx) -> str:
    return ctx.precision


def inexact(ctx: Ctx) -> bool:
    return ctx.precision != "double"


def factorize_near(ctx: Ctx, sigma0: float, on_eigenvalue: bool = False):
    """steps.cpp:243-305; returns (sigma_used, detection_off)."""
    fac = ctx.fac
    with ctx.clock.phase("factor"):
        h = ctx.held
        if h.shift == sigma0 and not h.detection_off and h.precision == ctx.precision:
            return sigma0, False
        ctx.held = FactorHeld()
        fac.set_null_pivot_detection(True)
        if fac.factorize(1.0, -sigma0, precision=ctx.precision):
            ctx.held = FactorHeld(sigma0, False, ctx.precision)
            return sigma0, False
        nulls = fac.inertia().zero
        if inexact(ctx):
            fac.set_null_pivot_detection(False)
            ok = fac.factorize(1.0, -sigma0, precision=ctx.precision)
            fac.set_null_pivot_detection(True)
            ctx.event("inexact_singular_shift", sigma=sigma0, null_pivots=nulls, without_detection=ok)
            if not ok:
                raise InexactFactorFailure(
                    f"the single-precision factorization of A - sigma B is singular at sigma = {sigma0:.6g} even "
                    f"without null-pivot detection ({nulls} null pivots)")
            ctx.held = FactorHeld(sigma0, True, ctx.precision)
            return sigma0, True
        cluster = (not on_eigenvalue) and nulls >= 2
        may_perturb = ctx.cfg["max_shift_retries"] > 0
        perturbed = sigma0 + 1e-8 * max(ctx.scale, abs(sigma0))
        ctx.event("singular_shift", sigma=sigma0, null_pivots=nulls, cluster=cluster)

        def at_perturbed():
            if not fac.factorize(1.0, -perturbed, precision=ctx.precision):
                return False
            if not on_eigenvalue:
                ctx.rep["shift_retries"] = max(ctx.rep["shift_retries"], 1)
                ctx.warn(f"shift perturbed to {perturbed:.17g} (A - sigma B singular at the requested shift)")
            ctx.held = FactorHeld(perturbed, False, ctx.precision)
            return True

        if not cluster and may_perturb and at_perturbed():
            return perturbed, False
        fac.set_null_pivot_detection(False)
        if fac.factorize(1.0, -sigma0, precision=ctx.precision):
            fac.set_null_pivot_detection(True)
            if not on_eigenvalue:
                ctx.rep["shift_retries"] = max(ctx.rep["shift_retries"], 2 if may_perturb else 1)
                ctx.warn(f"shift {sigma0:.17g} lies inside an eigenvalue cluster denser than the null-pivot "
                         "threshold; factorized there without null-pivot detection")
            ctx.held = FactorHeld(sigma0, True, ctx.precision)
            return sigma0, True
        fac.set_null_pivot_detection(True)
        if cluster and may_perturb and at_perturbed():
            return perturbed, False
        raise SingularError(f"A - sigma B is exactly singular at and near sigma = {sigma0}")


def try_refactorize_at(ctx: Ctx, sigma: float, detection_off: bool) -> bool:
    h = ctx.held
    if h.shift == sigma and h.detection_off == detection_off and h.precision == ctx.precision:
        return True
    with ctx.clock.phase("factor"):
        ctx.held = FactorHeld()
        ctx.fac.set_null_pivot_detection(not detection_off)
        ok = ctx.fac.factorize(1.0, -sigma, precision=ctx.precision)
        ctx.fac.set_null_pivot_detection(True)
        if ok:
            ctx.held = FactorHeld(sigma, detection_off, ctx.precision)
        return ok


def refactorize_at(ctx: Ctx, sigma: float, detection_off: bool):
    if try_refactorize_at(ctx, sigma, detection_off):
        return
    if inexact(ctx):
        raise InexactFactorFailure(f"the inexact refactorization at the working shift {sigma:.6g} is flagged singular")
    raise InternalError(f"refactorization at the working shift {sigma} failed")


def count_once(ctx: Ctx, s: float, detection: bool):
    """One exact (double, full-rank) Sylvester count: nu(A - s B) or None when singular."""
    fac = ctx.fac
    fac.set_null_pivot_detection(detection)
    ok = fac.factorize(1.0, -s, precision="double")
    fac.set_null_pivot_detection(True)
    inn = fac.inertia()
    # the working factors are replaced by this count's factorization
    ctx.held = FactorHeld(s, not detection, "double") if ok else FactorHeld()
    return inn.neg if ok else None


def count_batched(ctx: Ctx, shifts, limit: int):
    """Exact counts nu(A - s B) at several shifts from FP64 factorizations with null-pivot
    detection made together (GpuFactorization.count_batch), at most `limit` per batch (the memory
    of `limit` concurrent factorizations): per shift nu, or None where K is singular there (the
    caller then counts that shift with count_below, which moves it). A count equals count_once(s,
    True). The factors ctx.fac held are replaced."""
    fac = ctx.fac
    one = float(fac.info.panel_bytes_double + fac.info.cb_bytes_double + fac.info.workspace_bytes_double)
    with ctx.clock.phase("certificate"):
        fac.set_null_pivot_detection(True)
        ctx.held = FactorHeld()
        got = fac.count_batch([(ctx.pencil, 1.0, -s) for s in shifts], memory_bytes=max(1, int(limit)) * one)
    return [g.neg if g.valid else None for g in got]


def count_below(ctx: Ctx, s: float, outward: float, step: float, s_is_gap: bool = False):
    """steps.cpp:549-568: (nu, s_final, detection_off, near_eigenvalue)."""
    near = False
    with ctx.clock.phase("certificate"):
        for tries in range(8):
            for detection in (True, False):
                if not detection and tries == 0 and not s_is_gap:
                    s += outward * 0.1 * step
                nu = count_once(ctx, s, detection)
                if nu is not None:
                    if not detection and not ctx.rep["certificate"]["note"]:
                        ctx.rep["certificate"]["note"] = ("endpoint inertia counted without null-pivot detection "
                                                          "(dense cluster)")
                    return nu, s, not detection, near
                near = True
            s += outward * step * (4.0 ** tries)
            ctx.rep["certificate"]["endpoint_moves"] += 1
    raise SingularError(f"cannot find a nonsingular certificate endpoint near {s}")


# ------------------------------------------------------------------------------------ runs

@dataclass
class DefiniteRun:
    all: Block = None
    ritz_hints: list = field(default_factory=list)
    restarts: int = 0
    applies: int = 0
    converged: bool = False


def make_operator(ctx: Ctx, sigma: float):
    if inexact(ctx):
        cfg = ctx.cfg
        ref = SolveRefinement(max_steps=INEXACT_REFINE_STEPS,
                              rel_tol=cfg["tol"] * cfg["inner_tol_factor"] * INEXACT_REFINE_TOL_FACTOR)
        return ShiftInvertOperator(ctx.fac, ctx.ops, sigma, ref, checked=True,
                                   accept_rel=cfg["tol"] * cfg["inner_tol_factor"])
    return ShiftInvertOperator(ctx.fac, ctx.ops, sigma)


def lanczos_run(ctx: Ctx, sigma: float, which: str, nev: int, p: int, deflation: torch.Tensor | None,
                run: DefiniteRun, lead: int = 0):
    """definite.cpp:213-309: one run at the held shift; new vectors appended to run.all, then a
    rank-revealing Rayleigh-Ritz over all of them."""
    cfg = ctx.cfg
    op = make_operator(ctx, sigma)
    ko = KrylovOptions(nev=nev, ncv=cfg["ncv"], max_restarts=cfg["max_restarts"],
                       tol=cfg["tol"] * cfg["inner_tol_factor"], seed=cfg["seed"] + run.restarts,
                       block_size=max(p, 1), lead_nev=lead if lead < nev else 0)
    with ctx.clock.phase("iterate"):
        # without refinement an application only enqueues kernels: the next one may be launched
        # ahead of the host's Rayleigh-Ritz probe
        res = lanczos_krylov_schur(ctx.ops, op.apply, which, ko, deflation, prefetch=op.ref.max_steps == 0)
        # operator applications are counted by the operator object
        res.op_applies = op.applications
    run.restarts += res.restarts
    run.applies += res.op_applies
    run.converged = res.converged
    ctx.event("lanczos", sigma=sigma, which=which, nev=nev, block_size=ko.block_size, nconv=res.nconv,
              restarts=res.restarts, op_applies=res.op_applies,
              deflated=0 if deflation is None else int(deflation.shape[0]))
    for th in res.ritz_all:
        if th != 0.0:
            run.ritz_hints.append(sigma + 1.0 / th)
    with ctx.clock.phase("verify"):
        lam = np.array([sigma + 1.0 / th if th != 0.0 else math.inf for th in res.theta])
        blk = Block(ctx.ops, res.vectors, lam)
        blk.recompute_products()
        if run.all is None or run.all.size == 0:
            run.all = blk
        else:
            run.all.append(blk)
        m0 = run.all.size
        r = rayleigh_ritz_definite(run.all)
        if r < m0:
            ctx.event("dependent_dropped", columns=m0 - r)


def keep_converged(run: DefiniteRun, n2a: float, n2b: float, tol: float):
    """definite.cpp:314-324."""
    m = block_metrics(run.all, n2a, n2b)
    conv = [j for j in range(run.all.size) if math.isfinite(run.all.lam[j]) and m[j].backward_error <= 10.0 * tol]
    run.all = run.all.take(conv)


def rescue_pairs(ctx: Ctx, sel: Block, cols, n2a: float, n2b: float, tol: float) -> int:
    """definite.cpp:152-188: x <- x - K^-1 (A x - rho B x), at most 3 times per failing column."""
    ops = ctx.ops
    rescued = 0
    for j in cols:
        for _ in range(3):
            x = sel.x[j:j + 1]
            ax = sel.ax[j:j + 1]
            bx = sel.bx[j:j + 1]
            rho = float(ops.dots(x, ax).cpu()[0]) / float(ops.dots(x, bx).cpu()[0])
            r = ax.clone()
            ops.axpby(-rho, bx, 1.0, r)
            ctx.fac.solve(r)
            xn = x.clone()
            ops.axpby(-1.0, r, 1.0, xn)
            axn = ops.apply_a(xn)
            bxn = ops.apply_b(xn)
            bn = math.sqrt(max(float(ops.dots(xn, bxn).cpu()[0]), 1e-300))
            sel.x[j:j + 1].copy_(xn / bn)
            sel.ax[j:j + 1].copy_(axn / bn)
            sel.bx[j:j + 1].copy_(bxn / bn)
            sel.lam[j] = float(ops.dots(sel.x[j:j + 1], sel.ax[j:j + 1]).cpu()[0])
            sub = Block(ops, sel.x[j:j + 1], sel.lam[j:j + 1])
            sub.ax, sub.bx = sel.ax[j:j + 1], sel.bx[j:j + 1]
            if block_metrics(sub, n2a, n2b)[0].backward_error <= tol:
                break
        rescued += 1
    return rescued


def verify_and_rescue(ctx: Ctx, sel: Block, n2a: float, n2b: float, prepare):
    """definite.cpp:370-405; returns (sel, metrics)."""
    metrics = block_metrics(sel, n2a, n2b)
    failing = [j for j in range(sel.size) if metrics[j].backward_error > ctx.cfg["tol"]]
    if not failing:
        return sel, metrics
    prepare()
    if not ctx.fac.inertia().valid:
        return sel, metrics
    with ctx.clock.phase("verify"):
        r = rescue_pairs(ctx, sel, failing, n2a, n2b, ctx.cfg["tol"])
        m0 = sel.size
        kept = rayleigh_ritz_definite(sel)
        ctx.event("rescue", pairs=r, dependent_dropped=m0 - kept)
        metrics = block_metrics(sel, n2a, n2b)
    return sel, metrics


def select_nearest(lam: np.ndarray, sigma: float, k: int, zero_tol: float):
    """definite.cpp:409-425: (indices nearest first, zero_dropped)."""
    idx = []
    dropped = 0
    for i in range(lam.size):
        if not math.isfinite(lam[i]):
            continue
        if zero_tol >= 0 and abs(lam[i]) <= zero_tol:
            dropped += 1
            continue
        idx.append(i)
    idx.sort(key=lambda i: abs(lam[i] - sigma))
    return idx[:k], dropped


def nearest_beyond(hints, edge: float, direction: float, eps: float) -> float:
    best = math.nan
    for h in hints:
        if not math.isfinite(h) or direction * (h - edge) <= eps:
            continue
        if math.isnan(best) or direction * (h - best) < 0:
            best = h
    return best
