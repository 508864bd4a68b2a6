"""Device kernels of the operator and Krylov layers, and the shift-invert operator.

Blocks of vectors are torch tensors of shape (p, n) on the GPU (row c = column c of the
column-major n x p block of the C ABI), float64 for a real pencil and complex128 for a complex
Hermitian one (the hsg_z* kernels; kernels without a complex variant run on the real view, 2n
doubles per row). Every arithmetic kernel is a libhsg entry point; torch only allocates, copies and
indexes.

ShiftInvertOperator restates src/ops/shift_invert_operator.cpp:83-177: Y = K^-1 B X with
K = A - sigma B factorized by GpuFactorization; with an inexact (single-precision) factor every
apply is refined in FP64 against the exact A and B (r = B (x + sigma y) - A y in one fused sweep),
per column until ||r|| <= rel_tol ||B x|| or the componentwise backward error reaches 8 eps, a
correction that does not reduce the residual is rejected (best iterate kept), at most max_steps
block corrections, only the active columns are solved.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .factor import GpuFactorization, Pencil, device, ptr, stream_ptr

EPS = float(np.finfo(np.float64).eps)
ROUNDING_LEVEL = 8.0 * EPS  # shift_invert_operator.cpp:37


class DeviceOps:
    """Kernels bound to one pencil (A, B on the device)."""

    def __init__(self, pencil: Pencil):
        self.pencil = pencil
        self.n = pencil.n
        self.dev = device()
        self.d = pencil.to_device()
        self.is_complex = pencil.is_complex
        self.dtype = torch.complex128 if self.is_complex else torch.float64
        self.np_dtype = np.complex128 if self.is_complex else np.float64
        self._k = "hsg_z" if self.is_complex else "hsg_"

    # --------------------------------------------------------------- helpers
    def empty(self, p: int) -> torch.Tensor:
        return torch.empty((p, self.n), dtype=self.dtype, device=self.dev)

    def zeros(self, p: int) -> torch.Tensor:
        return torch.zeros((p, self.n), dtype=self.dtype, device=self.dev)

    def block(self, a: np.ndarray) -> torch.Tensor:
        """(p, n) device block of this pencil's scalar type from a host array."""
        return torch.from_numpy(np.ascontiguousarray(a, dtype=self.np_dtype)).to(self.dev)

    def coef(self, m: np.ndarray) -> torch.Tensor:
        """A host k x p coefficient matrix as the (p, k) device tensor update() takes. Staged through
        pinned memory and copied without blocking the host: the copy is ordered on the current
        stream before the kernels that read it, and the caching host allocator keeps the staging
        block until the copy has run."""
        a = np.ascontiguousarray(np.asarray(m).T, dtype=self.np_dtype)
        h = torch.empty(a.shape, dtype=self.dtype, pin_memory=True)
        h.numpy()[...] = a
        return h.to(self.dev, non_blocking=True)

    def _rows(self) -> int:
        """Doubles per row of a block (the real view of a complex block has 2n)."""
        return 2 * self.n if self.is_complex else self.n

    # --------------------------------------------------------------- products
    def apply_a(self, x: torch.Tensor, y: torch.Tensor | None = None, alpha=1.0, beta=0.0) -> torch.Tensor:
        p = x.shape[0]
        if y is None:
            y = self.empty(p)
        d = self.d
        _lib.call(self._k + "csr_spmm", self.n, ptr(d["a_rowptr"]), ptr(d["a_col"]), ptr(d["a_val"]), ptr(x),
                  self.n, ptr(y), self.n, p, float(alpha), float(beta), stream_ptr())
        return y

    def apply_b(self, x: torch.Tensor, y: torch.Tensor | None = None, alpha=1.0, beta=0.0) -> torch.Tensor:
        p = x.shape[0]
        if y is None:
            y = self.empty(p)
        d = self.d
        if self.pencil.b_is_diagonal:
            # B = I (b_diag holds ones) or a diagonal mass matrix
            _lib.call(self._k + "diag_spmm", self.n, ptr(d["b_diag"]), ptr(x), self.n, ptr(y), self.n, p,
                      float(alpha), float(beta), stream_ptr())
        else:
            _lib.call(self._k + "csr_spmm", self.n, ptr(d["b_rowptr"]), ptr(d["b_col"]), ptr(d["b_val"]), ptr(x),
                      self.n, ptr(y), self.n, p, float(alpha), float(beta), stream_ptr())
        return y

    def lanczos_matrix(self, apply):
        """(rowptr, col, values, diag) of the matrix `apply` multiplies by, for
        hsg_lanczos_coefficients: A, or B (CSR, or its diagonal); None for another operator."""
        d = self.d
        if getattr(apply, "__func__", None) is DeviceOps.apply_a and apply.__self__ is self:
            return d["a_rowptr"], d["a_col"], d["a_val"], None
        if getattr(apply, "__func__", None) is DeviceOps.apply_b and apply.__self__ is self:
            if self.pencil.b_is_diagonal:
                return None, None, None, d["b_diag"]
            return d["b_rowptr"], d["b_col"], d["b_val"], None
        return None

    def b_weight(self):
        """Diagonal weight of the B inner product (None for B = I)."""
        if self.pencil.b_is_identity:
            return None
        if self.pencil.b_is_diagonal:
            return self.d["b_diag"]
        raise NotImplementedError

    # --------------------------------------------------------------- dense
    def gram(self, v: torch.Tensor, w: torch.Tensor, weight=None, out: torch.Tensor | None = None) -> torch.Tensor:
        """C = V^H diag(weight) W as a (p, k) tensor (column-major k x p); `out`, when given, is a
        contiguous (p, k) tensor that receives C."""
        k, p = v.shape[0], w.shape[0]
        c = torch.empty((p, k), dtype=self.dtype, device=self.dev) if out is None else out
        if k == 0 or p == 0:
            return c
        _lib.call(self._k + "gram", self.n, ptr(v), self.n, k, ptr(weight), ptr(w), self.n, p, ptr(c), k, stream_ptr())
        return c

    def b_gram(self, v: torch.Tensor, w: torch.Tensor, bw: torch.Tensor | None = None,
               out: torch.Tensor | None = None) -> torch.Tensor:
        """V^H B W; bw = B W when already formed (general B)."""
        if self.pencil.b_is_diagonal:
            return self.gram(v, w, self.b_weight(), out)
        if bw is None:
            bw = self.apply_b(w)
        return self.gram(v, bw, None, out)

    def update(self, v: torch.Tensor, c: torch.Tensor, w: torch.Tensor, alpha=1.0, beta=1.0) -> torch.Tensor:
        """W = beta W + alpha V C with C a (p, k) tensor (column-major k x p)."""
        k, p = v.shape[0], w.shape[0]
        if p == 0:
            return w
        _lib.call(self._k + "block_update", self.n, ptr(v), self.n, k, ptr(c), k, ptr(w), self.n, p, float(alpha),
                  float(beta), stream_ptr())
        return w

    def block_orthonormalize(self, v: torch.Tensor | None, w: torch.Tensor, breakdown_tol: float):
        """Block CGS2 of w (p, n) against v[:k] in the B inner product, then two Cholesky-QR passes,
        in one library call that leaves the interpreter free while it runs (hsg_block_orthonormalize).
        Returns (status, norms_in, coef k x p, R p x p): status 0 when w is orthonormalized; 1 or 2
        when the Cholesky-QR declined before or after its first pass (w then holds the block as
        it was at that point, coef and norms_in are final, R is not)."""
        k = 0 if v is None else v.shape[0]
        p = w.shape[0]
        sw = 2 if self.is_complex else 1
        # work comes from torch's stream-ordered caching allocator on the current stream, so the
        # last update still queued there when the call returns reads it before anyone reuses it
        general = not self.pencil.b_is_diagonal
        need = sw * (p + 2 * k * p + 2 * p * p) + sw * self.n * p * (2 if general else 1)
        work = torch.empty(need, dtype=torch.float64, device=self.dev)
        norms = np.zeros(p, dtype=np.float64)
        coef = np.zeros((k, p), dtype=self.np_dtype)
        R = np.zeros((p, p), dtype=self.np_dtype)
        status = C.c_int32(0)
        d = self.d
        brow, bcol, bval = (d["b_rowptr"], d["b_col"], d["b_val"]) if general else (None, None, None)
        bdiag = None if general else self.b_weight()
        _lib.call("hsg_block_orthonormalize", int(self.is_complex), self.n, ptr(v) if k else None, self.n, k, ptr(w),
                  self.n, p, ptr(brow), ptr(bcol), ptr(bval), ptr(bdiag), float(breakdown_tol), ptr(work),
                  norms.ctypes.data, coef.ctypes.data if k else None, R.ctypes.data, C.byref(status), stream_ptr())
        return int(status.value), norms, coef, R

    def dots(self, x: torch.Tensor, y: torch.Tensor, weight=None, out: torch.Tensor | None = None) -> torch.Tensor:
        """x_c^H diag(weight) y_c per row pair; real parts (the values the driver uses: norms and the
        Rayleigh quotients of Hermitian pencils, whose imaginary parts vanish). `out`, when given,
        is a (p,) tensor of the block's scalar type that receives the (complex) sums."""
        p = x.shape[0]
        if out is None:
            out = torch.empty(p, dtype=self.dtype, device=self.dev)
        if p:
            _lib.call(self._k + "column_dots", self.n, ptr(x), self.n, ptr(y), self.n, p, ptr(weight), ptr(out),
                      stream_ptr())
        return out.real if self.is_complex else out

    def b_dots(self, x, y, by=None, out: torch.Tensor | None = None):
        if self.pencil.b_is_diagonal:
            return self.dots(x, y, self.b_weight(), out)
        if by is None:
            by = self.apply_b(y)
        return self.dots(x, by, None, out)

    def pair_sums(self, ax: torch.Tensor, bx: torch.Tensor, x: torch.Tensor, lam: np.ndarray) -> np.ndarray:
        """(p, 3) host array: ||ax - lambda bx||^2, ||bx||^2, ||x||^2 per column."""
        p = x.shape[0]
        if p == 0:
            return np.zeros((0, 3))
        ld = torch.from_numpy(np.ascontiguousarray(lam, dtype=np.float64)).to(self.dev)
        out = torch.empty(3 * p, dtype=torch.float64, device=self.dev)
        nr = self._rows()  # real lambda: the sums of a complex block are those of its real view
        _lib.call("hsg_pair_sums", nr, ptr(ax), ptr(bx), ptr(x), nr, p, ptr(ld), ptr(out), stream_ptr())
        return out.cpu().numpy().reshape(p, 3)

    def axpby(self, a: float, x: torch.Tensor, b: float, y: torch.Tensor) -> torch.Tensor:
        p = x.shape[0]
        if p:
            nr = self._rows()  # real a, b: a complex block is updated through its real view
            _lib.call("hsg_axpby", nr, float(a), ptr(x), nr, float(b), ptr(y), nr, p, stream_ptr())
        return y

    def shifted_residual(self, x: torch.Tensor, y: torch.Tensor, sigma: float, r: torch.Tensor | None = None):
        """R = B (X + sigma Y) - A Y, with ||R_c||^2 and omega_c per column (host arrays)."""
        p = x.shape[0]
        if r is None:
            r = self.empty(p)
        rn = torch.empty(p, dtype=torch.float64, device=self.dev)
        om = torch.empty(p, dtype=torch.float64, device=self.dev)
        d = self.d
        _lib.call(self._k + "shifted_residual", self.n, ptr(d["a_rowptr"]), ptr(d["a_col"]), ptr(d["a_val"]),
                  ptr(d["b_rowptr"]), ptr(d["b_col"]), ptr(d["b_val"]), ptr(d["b_diag"]), float(sigma), ptr(x),
                  self.n, ptr(y), self.n, ptr(r), self.n, p, ptr(rn), ptr(om), stream_ptr())
        return r, rn, om


@dataclass
class SolveRefinement:
    """include/hsolve/operator.hpp:51-57."""
    max_steps: int = 0
    rel_tol: float = 0.0
    stop_at_rounding_level: bool = True


class InexactFactorFailure(RuntimeError):
    """A refined apply with an inexact factor did not converge (src/driver/steps.cpp:132-143)."""


class ShiftInvertOperator:
    """Op = K^{-1} B, K = A - sigma B (src/ops/shift_invert_operator.cpp:83-177)."""

    def __init__(self, factor: GpuFactorization, ops: DeviceOps, sigma: float, refinement: SolveRefinement | None = None,
                 checked: bool = False, accept_rel: float = math.inf):
        if not factor.inertia().valid:
            from .errors import InvalidInputError
            raise InvalidInputError("ShiftInvertOperator: the factor holds no usable factorization")
        self.factor = factor
        self.ops = ops
        self.sigma = float(sigma)
        self.ref = refinement or SolveRefinement()
        self.checked = checked
        self.accept_rel = accept_rel
        self.applications = 0
        self.refinement_steps = 0
        self.last_rel_residual = 0.0
        self.last_backward_error = 0.0
        self.last_converged = True
        self._factorizations = factor.stats.factorizations

    def discard(self, y: torch.Tensor) -> None:
        """Undoes the counts of an application whose result goes unused (launched ahead of a host
        decision that then ended the run), so that the counters report the applications the
        algorithm used."""
        p = y.shape[0]
        self.applications -= p
        self.factor.stats.solves -= 1
        self.factor.stats.rhs_solved -= p

    def apply(self, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
        if self.factor.stats.factorizations != self._factorizations:
            from .errors import InternalError
            raise InternalError("ShiftInvertOperator: the factorization was replaced after the operator was built")
        ops = self.ops
        p = x.shape[0]
        self.applications += p
        y = ops.apply_b(x, y)
        if self.ref.max_steps == 0:
            self.factor.solve(y)
            self.last_converged = True
            return y
        bxn = torch.sqrt(ops.dots(y, y)).cpu().numpy()
        self.factor.solve(y)
        r, rn2, om = ops.shifted_residual(x, y, self.sigma)
        rn2 = rn2.cpu().numpy()
        omega = om.cpu().numpy()

        def rel_of(rr2, k):
            rnk = math.sqrt(max(rr2[k], 0.0))
            if rnk == 0.0:
                return 0.0
            return rnk / bxn[k] if bxn[k] > 0 else math.inf

        rel = np.array([rel_of(rn2, k) for k in range(p)])
        floor = ROUNDING_LEVEL if self.ref.stop_at_rounding_level else -1.0
        active = (rel > self.ref.rel_tol) & (omega > floor)
        for _ in range(self.ref.max_steps):
            if not active.any():
                break
            cols = np.flatnonzero(active)
            idx = torch.from_numpy(cols).to(ops.dev)
            c = r.index_select(0, idx).contiguous()
            self.factor.solve(c)
            self.refinement_steps += 1
            cand = y.index_select(0, idx).contiguous()
            ops.axpby(1.0, c, 1.0, cand)  # cand = y + correction
            xs = x.index_select(0, idx).contiguous()
            rc, rc2, omc = ops.shifted_residual(xs, cand, self.sigma)
            rc2 = rc2.cpu().numpy()
            omc = omc.cpu().numpy()
            accept = []
            for j, k in enumerate(cols):
                cr = rel_of(rc2, j) if bxn[k] > 0 else math.inf
                if cr < rel[k]:
                    accept.append(j)
                    rel[k] = cr
                    omega[k] = omc[j]
                    if cr > self.ref.rel_tol and omega[k] > floor:
                        continue
                active[k] = False
            if accept:
                a_idx = torch.from_numpy(np.asarray(accept)).to(ops.dev)
                dst = torch.from_numpy(cols[np.asarray(accept)]).to(ops.dev)
                y.index_copy_(0, dst, cand.index_select(0, a_idx))
                r.index_copy_(0, dst, rc.index_select(0, a_idx))
        self.last_rel_residual = float(rel.max()) if p else 0.0
        self.last_backward_error = float(omega.max()) if p else 0.0
        self.last_converged = bool(np.all((rel <= self.ref.rel_tol) | (omega <= floor)))
        if self.checked and not self.last_converged and not (self.last_rel_residual <= self.accept_rel):
            raise InexactFactorFailure(
                f"the refinement of an operator application at sigma = {self.sigma:.6g} did not converge "
                f"(relative residual {self.last_rel_residual:.3g}, backward error {self.last_backward_error:.3g}): "
                "eps_factor * cond(K) is too large at this shift")
        return y
