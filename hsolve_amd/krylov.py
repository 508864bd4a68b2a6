"""Block thick-restart Lanczos (Hermitian Krylov-Schur) in the B inner product.

Restates src/eigen/block_lanczos.cpp:125-417 (solver), :421-462 (layout, block_dimensions) and the
result contract of src/eigen/lanczos_krylov_schur.cpp:506-594 (RitzPairs classify / done /
result), with the thresholds of src/eigen/block_lanczos.hpp:26-38:
  * basis [D | V | V_m] (d deflation columns, ncv Krylov columns, one residual block of p);
  * a block step W = Op V_j, block CGS2 against [D V] in the B inner product and Cholesky-QR2
    within the block (columns that break down, |R(i,i)| <= 64 eps ||Op v||, are replaced by seeded
    random vectors); A_j and the coupling R fill the projected matrix H;
  * Rayleigh-Ritz: eig(H(0:m, 0:m)), estimates ||R S(m-p:m, i)||, converged when
    est <= tol |theta| (|theta| -> max |theta| for theta = 0); early probes at block boundaries
    from first_probe (first cycle) or keep + p (later cycles);
  * thick restart keeping `keep` best-ranked Ritz vectors; a converged pair whose coupling is
    <= 4 eps * (smallest wanted |theta|) is locked (its coupling zeroed);
  * result: the nev best ranked, converged first, vectors = V S(:, sel).
The operator, the orthogonalization and the basis rotation run on the GPU (libhsg kernels); the
small projected eigenproblems (order <= ncv) and Cholesky factors of p x p Gram matrices run on
the host, as in the reference (LAPACK heev / potrf).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
import torch

from .ops import DeviceOps

# Host LAPACK (projected eigenproblems, Cholesky factors of small Gram matrices) is called from the
# threads of concurrent spectrum slices without a lock: numpy's and scipy's bundled LAPACK builds
# are reentrant, and a lock would make one slice's 8 x 8 Cholesky wait behind another slice's
# Rayleigh-Ritz eigensolve.

EPS = float(np.finfo(np.float64).eps)
BREAKDOWN_TOL = 64 * EPS      # block_lanczos.hpp:29
LOCK_TOL = 4 * EPS            # block_lanczos.hpp:32
START_MIN_RATIO = 1e-12       # block_lanczos.hpp:35
MIN_NEW_BLOCKS = 2            # block_lanczos.cpp:24
FREE_SPACE_FACTOR = 2         # block_lanczos.hpp (kFreeSpaceFactor)

LARGEST_MAGNITUDE = "largest_magnitude"
LARGEST_ALGEBRAIC = "largest_algebraic"
SMALLEST_ALGEBRAIC = "smallest_algebraic"


@dataclass
class KrylovOptions:
    nev: int
    ncv: int = 0
    max_restarts: int = 1000
    tol: float = 1e-11
    seed: int = 2026
    block_size: int = 1
    lead_nev: int = 0


@dataclass
class EigResult:
    theta: np.ndarray
    estimate: np.ndarray
    vectors: torch.Tensor          # (nev, n)
    ritz_all: np.ndarray
    nconv: int
    converged: bool
    restarts: int = 0
    op_applies: int = 0
    history: list = field(default_factory=list)


def _round_up(x, p):
    return (x + p - 1) // p * p


def block_dimensions(n_free: int, nev: int, ncv_req: int, p: int, extra_blocks: int = 4, keep_fraction: float = 0.5):
    """Layout of block_lanczos.cpp:421-462: (ncv, keep, first_probe, feasible)."""
    setup_ncv = min(ncv_req, n_free) if ncv_req > 0 else min(n_free, max(2 * nev + 1, nev + 16))
    if p > n_free // ((MIN_NEW_BLOCKS + 2) * FREE_SPACE_FACTOR):
        return 0, 0, 0, False
    spare = _round_up(nev + 1, p)
    base = max(_round_up(setup_ncv, p), spare + MIN_NEW_BLOCKS * p)
    ncv = base
    if ncv_req == 0:
        room = (n_free // FREE_SPACE_FACTOR - p) // p * p
        extra = min(extra_blocks, room // p)
        wanted = base + extra * p
        ncv = max(base, min(wanted, room))
    kept = int(nev + keep_fraction * (ncv - nev))
    keep = min(max(kept // p * p, spare), ncv - p)
    feasible = setup_ncv < n_free and ncv + p <= n_free // FREE_SPACE_FACTOR
    return ncv, keep, base, feasible


class RitzPairs:
    """block_lanczos.hpp:86-117, lanczos_krylov_schur.cpp:506-594."""

    def __init__(self, theta, s, est, tol, which):
        self.theta = theta
        self.s = s
        self.est = est
        m = theta.size
        max_abs = float(np.max(np.abs(theta))) if m else 0.0
        self.ref = np.where(theta != 0.0, np.abs(theta), max_abs)
        self.conv = est <= tol * self.ref
        idx = np.arange(m)
        if which == SMALLEST_ALGEBRAIC:
            self.order = idx
        elif which == LARGEST_ALGEBRAIC:
            self.order = idx[::-1].copy()
        else:  # largest magnitude, ties: the positive value first
            self.order = np.array(sorted(idx, key=lambda i: (-abs(theta[i]), -theta[i])), dtype=np.int64)

    def wanted_converged(self, nev):
        return int(np.count_nonzero(self.conv[self.order[:nev]]))

    def done(self, nev, lead):
        if self.wanted_converged(nev) >= nev:
            return True
        if lead <= 0:
            return False
        return bool(np.all(self.conv[self.order[:lead]]))

    def min_wanted_reference(self, nev):
        return float(np.min(self.ref[self.order[:nev]]))

    def record(self, restart, nev, applies):
        nconv, mx = 0, 0.0
        for i in self.order[:nev]:
            if self.conv[i]:
                nconv += 1
            else:
                rel = self.est[i] / self.ref[i] if self.ref[i] > 0 else math.inf
                mx = max(mx, rel)
        return {"restart": restart, "nconv": nconv, "max_estimate": mx, "op_applies": applies}


class BlockLanczos:
    """One shift-invert Krylov-Schur run; `op(x)` returns Op x for a (p, n) block."""

    def __init__(self, ops: DeviceOps, op, which: str, opt: KrylovOptions, deflation: torch.Tensor | None = None,
                 prefetch: bool = False):
        self.ops = ops
        self.op = op
        # the object counting applications (op is usually a bound method of the operator)
        self._op_obj = getattr(op, "__self__", op)
        # prefetch: op only enqueues device work (no host synchronization), so the image of the
        # next block can be launched before the host's Rayleigh-Ritz probe and dropped if the
        # probe ends the run
        self.prefetch = prefetch
        self.which = which
        self.opt = opt
        n = ops.n
        self.n = n
        self.d = 0 if deflation is None else int(deflation.shape[0])
        self.nev = opt.nev
        if self.nev < 1 or self.nev > n - self.d:
            raise ValueError(f"nev = {self.nev} must be in [1, n - d]")
        self.p = max(1, opt.block_size)
        self.lead = opt.lead_nev if 0 < opt.lead_nev < self.nev else 0
        n_free = n - self.d
        ncv, keep, first_probe, feasible = block_dimensions(n_free, self.nev, opt.ncv, self.p)
        if not feasible:
            # one single-vector cycle over the whole free space (exact), block_lanczos.cpp:478-489
            self.p = 1
            ncv = n_free - 1 if n_free > 1 else 1
            keep = max(1, min(ncv - 1, self.nev + 1))
            first_probe = ncv
            self.exact_cycle = True
        else:
            self.exact_cycle = False
        self.ncv, self.keep, self.first_probe = ncv, keep, first_probe
        cols = self.d + ncv + self.p
        self.V = torch.zeros((cols, n), dtype=ops.dtype, device=ops.dev)
        if self.d:
            self.V[: self.d].copy_(deflation)
        self.dt = ops.np_dtype  # projected matrices: real symmetric or complex Hermitian
        self.H = np.zeros((ncv, ncv), dtype=self.dt)
        self.rlast = np.zeros((self.p, self.p), dtype=self.dt)
        self.restarts = 0
        self.m = 0
        self.rr = None
        self.coupling = None
        self.history = []
        self.next_stream = 0
        self.applies0 = 0
        self.fallbacks = 0

    # ------------------------------------------------------------ helpers
    def _random(self, cols: int) -> torch.Tensor:
        out = np.empty((cols, self.n), dtype=self.dt)
        for c in range(cols):
            rng = np.random.default_rng([self.opt.seed, self.next_stream])
            out[c] = rng.uniform(-1.0, 1.0, self.n)
            if self.ops.is_complex:
                out[c] += 1j * rng.uniform(-1.0, 1.0, self.n)
            self.next_stream += 1
        return self.ops.block(out)

    def _project_out(self, k: int, w: torch.Tensor, coef: np.ndarray | None = None):
        """w -= V[:k] (V[:k]^H B w), twice (CGS2); coefficients accumulated into coef (k x p)."""
        if k == 0:
            return
        vk = self.V[:k]
        for _ in range(2):
            c = self.ops.b_gram(vk, w)
            self.ops.update(vk, c, w, alpha=-1.0, beta=1.0)
            if coef is not None:
                coef += c.cpu().numpy().T

    def _orthonormalize(self, k: int, w: torch.Tensor, breakdown_tol: float):
        """Block CGS2 of w (p, n) against V[:k], then Cholesky-QR2 within the block. Returns
        (coef k x p, R p x p). Breakdown columns (B-norm below breakdown_tol times the norm w had on
        entry) are replaced by random vectors (coupling 0). The common case runs in one library
        call (hsg_block_orthonormalize: the CGS2 products, one copy of the entry norms, both
        coefficient blocks and the Gram matrix, the host Cholesky factors, a second Gram matrix),
        during which the other slice threads keep the interpreter; when the Cholesky-QR declines,
        the column-by-column fallback below finishes the block from where the call left it."""
        native = getattr(self.ops, "block_orthonormalize", None)
        if native is not None:
            status, norms_in, coef, R = native(self.V[:k] if k else None, w, breakdown_tol)
            if status == 0:
                return coef, R
            return self._column_fallback(k, w, coef, norms_in, breakdown_tol)
        return self._orthonormalize_steps(k, w, breakdown_tol)

    def _orthonormalize_steps(self, k: int, w: torch.Tensor, breakdown_tol: float):
        """The same block step through the ops' own products (host ops without the one-call form):
        the entry norms, both CGS2 coefficient blocks and the first Gram matrix land in one buffer
        fetched with one copy, so a block step waits twice (before each Cholesky factor)."""
        ops = self.ops
        p = w.shape[0]
        kp = k * p
        buf = torch.empty(p + 2 * kp + p * p, dtype=ops.dtype, device=ops.dev)
        ops.b_dots(w, w, out=buf[:p])
        if k:
            vk = self.V[:k]
            for i in range(2):
                c = buf[p + i * kp:p + (i + 1) * kp].view(p, k)
                ops.b_gram(vk, w, out=c)
                ops.update(vk, c, w, alpha=-1.0, beta=1.0)
        ops.b_gram(w, w, out=buf[p + 2 * kp:].view(p, p))
        h = buf.cpu().numpy()
        norms_in = np.sqrt(np.maximum(h[:p].real, 0.0))
        coef = np.zeros((k, p), dtype=self.dt)
        for i in range(2 if k else 0):
            coef += h[p + i * kp:p + (i + 1) * kp].reshape(p, k).T
        G = h[p + 2 * kp:].reshape(p, p).T
        G = 0.5 * (G + G.conj().T)
        ok = True
        R = np.eye(p, dtype=self.dt)
        if np.all(np.diag(G).real > (breakdown_tol * norms_in) ** 2) and np.all(np.isfinite(G)):
            for it in range(2):
                try:
                    Rk = sla.cholesky(G, lower=False)
                except np.linalg.LinAlgError:
                    ok = False
                    break
                d = np.abs(np.diag(Rk))
                if d.min() <= 1e-7 * d.max():  # too ill-conditioned for Cholesky-QR
                    ok = False
                    break
                Rinv = np.linalg.inv(Rk)
                tmp = ops.empty(p)
                ops.update(w, ops.coef(Rinv), tmp, 1.0, 0.0)
                w.copy_(tmp)
                R = Rk @ R
                if it == 0:
                    G = ops.b_gram(w, w).cpu().numpy().T
                    G = 0.5 * (G + G.conj().T)
        else:
            ok = False
        if ok:
            return coef, R
        return self._column_fallback(k, w, coef, norms_in, breakdown_tol)

    def _column_fallback(self, k: int, w: torch.Tensor, coef: np.ndarray, norms_in: np.ndarray,
                         breakdown_tol: float):
        """Column-by-column CGS2 of w (block_ortho.cpp column_fallback) after the block Cholesky-QR
        declined; coef (k x p) already holds both CGS2 passes against V[:k]."""
        ops = self.ops
        p = w.shape[0]
        self.fallbacks += 1
        R = np.zeros((p, p), dtype=self.dt)
        for c in range(p):
            col = w[c:c + 1]
            extra = np.zeros((k, 1), dtype=self.dt)
            self._project_out(k, col, extra)
            coef[:, c] += extra[:, 0]
            if c:
                rc = np.zeros((c, 1), dtype=self.dt)
                blk = w[:c]
                for _ in range(2):
                    cc = ops.b_gram(blk, col)
                    ops.update(blk, cc, col, alpha=-1.0, beta=1.0)
                    rc += cc.cpu().numpy().T
                R[:c, c] = rc[:, 0]
            nrm = math.sqrt(max(float(ops.b_dots(col, col).cpu()[0]), 0.0))
            if nrm > breakdown_tol * max(norms_in[c], 1e-300):
                col.mul_(1.0 / nrm)
                R[c, c] = nrm
            else:
                R[c, c] = 0.0
                for _ in range(3):
                    col.copy_(self._random(1))
                    self._project_out(k, col)
                    if c:
                        blk = w[:c]
                        for _ in range(2):
                            cc = ops.b_gram(blk, col)
                            ops.update(blk, cc, col, alpha=-1.0, beta=1.0)
                    nr = math.sqrt(max(float(ops.b_dots(col, col).cpu()[0]), 0.0))
                    if nr > 1e-8:
                        col.mul_(1.0 / nr)
                        break
        return coef, R

    def _set_start_block(self):
        d, p = self.d, self.p
        blk = self.V[d:d + p]
        blk.copy_(self._random(p))
        self._orthonormalize(d, blk, START_MIN_RATIO)

    # ------------------------------------------------------------ steps
    def _applications(self) -> int:
        return getattr(self._op_obj, "applications", 0)

    def _expand(self, first: int) -> bool:
        d, p = self.d, self.p
        pending = None  # Op of the block just orthonormalized, launched ahead of the probe
        for j in range(first, self.ncv, p):
            src = self.V[d + j:d + j + p]
            dst = self.V[d + j + p:d + j + 2 * p]
            dst.copy_(pending if pending is not None else self.op(src))
            pending = None
            coef, R = self._orthonormalize(d + j + p, dst, BREAKDOWN_TOL)
            A = coef[d + j:d + j + p, :]
            A = 0.5 * (A + A.conj().T)
            self.H[j:j + p, j:j + p] = A
            if j + p < self.ncv:
                self.H[j + p:j + 2 * p, j:j + p] = np.triu(R)
                self.H[j:j + p, j + p:j + 2 * p] = np.triu(R).conj().T
            self.m = j + p
            self.rlast = np.triu(R)
            threshold = self.first_probe if self.restarts == 0 else self.keep + p
            if self.m < self.ncv and self.m >= threshold:
                if self.prefetch:  # the next step's source is dst, final now
                    pending = self.op(dst)
                self._rayleigh_ritz()
                if self.rr.done(self.nev, self.lead):
                    discard = getattr(self._op_obj, "discard", None)
                    if pending is not None and discard is not None:
                        discard(pending)  # launched but not used
                    return True
        return False

    def _rayleigh_ritz(self):
        m, p = self.m, self.p
        Hm = self.H[:m, :m]
        Hm = 0.5 * (Hm + Hm.conj().T)
        theta, S = np.linalg.eigh(Hm)
        self.coupling = self.rlast @ S[m - p:m, :]
        est = np.linalg.norm(self.coupling, axis=0)
        self.rr = RitzPairs(theta, S, est, self.opt.tol, self.which)

    def _restart(self):
        m, k, p, d = self.ncv, self.keep, self.p, self.d
        rr = self.rr
        lock_tol = LOCK_TOL * rr.min_wanted_reference(self.nev)
        sel = rr.order[:k]
        sk = rr.s[:, sel]
        b = self.coupling[:, sel].copy()
        for l, i in enumerate(sel):
            if rr.conv[i] and rr.est[i] <= lock_tol:
                b[:, l] = 0.0
        vm = self.V[d:d + m]
        tmp = self.ops.empty(k)
        self.ops.update(vm, self.ops.coef(sk), tmp, 1.0, 0.0)
        resid = self.V[d + m:d + m + p].clone()
        self.V[d:d + k].copy_(tmp)
        self.V[d + k:d + k + p].copy_(resid)
        self.H[:] = 0.0
        for l, i in enumerate(sel):
            self.H[l, l] = rr.theta[i]
            self.H[k:k + p, l] = b[:, l]
            self.H[l, k:k + p] = b[:, l].conj()

    def run(self) -> EigResult:
        self.applies0 = self._applications()
        self._set_start_block()
        first = 0
        while True:
            if not self._expand(first):
                self._rayleigh_ritz()
            applies = self._applications() - self.applies0
            self.history.append(self.rr.record(self.restarts, self.nev, applies))
            if self.rr.done(self.nev, self.lead) or self.restarts >= self.opt.max_restarts:
                break
            self._restart()
            first = self.keep
            self.restarts += 1
        return self._finish()

    def _finish(self) -> EigResult:
        rr = self.rr
        m = self.m
        nev = self.nev
        sel = list(rr.order[:nev])
        sel = [i for i in sel if rr.conv[i]] + [i for i in sel if not rr.conv[i]]
        ssel = rr.s[:, sel]
        vec = self.ops.empty(nev)
        self.ops.update(self.V[self.d:self.d + m], self.ops.coef(ssel), vec, 1.0, 0.0)
        nconv = int(np.count_nonzero(rr.conv[sel]))
        return EigResult(theta=rr.theta[sel].copy(), estimate=rr.est[sel].copy(), vectors=vec,
                         ritz_all=rr.theta.copy(), nconv=nconv, converged=nconv >= nev, restarts=self.restarts,
                         op_applies=self._applications() - self.applies0, history=self.history)


def lanczos_krylov_schur(ops: DeviceOps, op, which: str, opt: KrylovOptions,
                         deflation: torch.Tensor | None = None, prefetch: bool = False) -> EigResult:
    """Entry point (src/eigen/lanczos_krylov_schur.cpp:617)."""
    return BlockLanczos(ops, op, which, opt, deflation, prefetch).run()
