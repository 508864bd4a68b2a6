"""The pencil K(alpha, beta) = alpha A + beta B and its GPU factorization.

Pencil mirrors ShiftedPencil<T> (include/hsolve/sparse.hpp:176-223, src/sparse/pencil.cpp:67-170):
one union pattern of A and B serves every shift. GpuFactorization mirrors the factorization
contract SparseFactorization<T> (include/hsolve/factor.hpp:341-443): analyze once, factorize
for any (alpha, beta), a singular K makes factorize() return False (never raises for null
pivots), count() gives the Sylvester inertia, solve() works in place on column-major blocks.
The numeric work runs in libhsg on the GPU (hsg_factorize / hsg_factor_solve); device buffers are torch
tensors. Blocks of vectors are torch tensors of shape (p, n): row c is column c of the n x p
column-major block the C ABI sees (leading dimension n).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .errors import InvalidInputError


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("hsolve_amd needs a ROCm GPU: torch.cuda.is_available() is False (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> C.c_void_p:
    if t is None:
        return C.c_void_p(0)
    return C.c_void_p(t.data_ptr())


def host_ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _csr(m, name: str) -> sp.csr_matrix:
    m = sp.csr_matrix(m)
    m.sum_duplicates()
    m.sort_indices()
    if m.shape[0] != m.shape[1]:
        raise InvalidInputError(f"{name} must be square, got shape {m.shape}")
    return m


@dataclass
class Inertia:
    """Inertia of K (include/hsolve/factor.hpp:295-300): valid <=> the factorization is usable."""
    neg: int = -1
    zero: int = -1
    pos: int = -1
    valid: bool = False
    two_by_two: int = 0


@dataclass
class FactorStats:
    """FactorStats (include/hsolve/factor.hpp:312-336), the fields this engine measures."""
    analyses: int = 0
    analyze_s: float = 0.0
    factorize_s: float = 0.0
    solve_s: float = 0.0
    factorizations: int = 0
    solves: int = 0
    rhs_solved: int = 0
    factor_entries: int = 0
    flops: float = 0.0
    null_pivots: int = 0
    delayed_pivots: int = 0
    two_by_two_pivots: int = 0
    supernodes: int = 0
    max_front: int = 0
    extra: dict = field(default_factory=dict)


class Pencil:
    """A and B (B = I when None) on the host and on the device, with the union pattern.

    Complex Hermitian pencils (complex128 A or B) are factorized through their real embedding of
    order 2n (the reference's MUMPS input, src/factor/mumps_input.cpp:43-150), here with the real
    and imaginary parts of unknown i interleaved as rows 2i, 2i+1:
        E(2i, 2j) = E(2i+1, 2j+1) = Re K(i, j),  E(2i, 2j+1) = -Im K(i, j),  E(2i+1, 2j) = Im K(i, j),
    so that E acting on the real view of a complex block (re, im interleaved, as complex128 is
    stored) is K acting on the block: solves run in place on that view. Every eigenvalue of K
    appears twice in E. Positions whose real (or, off the diagonal, imaginary) part is zero in A
    and in B are left out of E (alpha and beta are real, so they stay zero for every shift). The
    values of E are gathered from the extended arrays [Re v, Im v, -Im v] of A's and B's values.
    f_n, f_rowptr, f_col, f_a_index, f_b_index describe the pattern the factorization analyses
    (the union pattern itself for a real pencil)."""

    def __init__(self, A, B=None):
        A = _csr(A, "A")
        Bc = None if B is None else _csr(B, "B")
        cplx = bool(np.iscomplexobj(A.data) or (Bc is not None and np.iscomplexobj(Bc.data)))
        dt = np.complex128 if cplx else np.float64
        A = A.astype(dt)
        n = A.shape[0]
        self.n = n
        self.is_complex = cplx
        self.b_is_identity = B is None
        if B is None:
            Bm = sp.identity(n, format="csr", dtype=dt)
        else:
            Bm = Bc.astype(dt)
            if Bm.shape != A.shape:
                raise InvalidInputError(f"A and B differ in shape: {A.shape} vs {Bm.shape}")
        self.A = A
        self.B = None if B is None else Bm
        self._Bfull = Bm
        diag_b = Bm.nnz == n and np.array_equal(Bm.indices, np.arange(n)) and np.array_equal(
            Bm.indptr, np.arange(n + 1))
        # a diagonal B of a Hermitian pencil is real (its imaginary parts are zero)
        self.b_is_diagonal = bool(diag_b) and not (cplx and np.any(Bm.data.imag != 0.0))
        self._build_union(A, Bm)
        if cplx:
            self._build_embedding(A, Bm)
        else:
            self.f_n, self.f_rowptr, self.f_col = n, self.u_rowptr, self.u_col
            self.f_a_index, self.f_b_index = self.a_index, self.b_index
        self._dev = None

    def _build_union(self, A, Bm):
        n = self.n
        ra = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.indptr))
        rb = np.repeat(np.arange(n, dtype=np.int64), np.diff(Bm.indptr))
        ka = ra * n + A.indices.astype(np.int64)
        kb = rb * n + Bm.indices.astype(np.int64)
        keys = np.union1d(ka, kb)
        self.u_rowptr = np.zeros(n + 1, dtype=np.int64)
        rows = keys // n
        np.add.at(self.u_rowptr, rows + 1, 1)
        np.cumsum(self.u_rowptr, out=self.u_rowptr)
        self.u_col = (keys % n).astype(np.int32)
        self.a_index = np.full(keys.size, -1, dtype=np.int64)
        self.b_index = np.full(keys.size, -1, dtype=np.int64)
        self.a_index[np.searchsorted(keys, ka)] = np.arange(ka.size, dtype=np.int64)
        self.b_index[np.searchsorted(keys, kb)] = np.arange(kb.size, dtype=np.int64)

    def _build_embedding(self, A, Bm):
        n = self.n
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(self.u_rowptr))
        cols = self.u_col.astype(np.int64)
        ia, ib = self.a_index, self.b_index
        na, nb = A.nnz, Bm.nnz

        def part_nonzero(data, idx, f):
            out = np.zeros(idx.size, dtype=bool)
            has = idx >= 0
            out[has] = f(data[idx[has]]) != 0.0
            return out

        diag = rows == cols
        re_nz = diag | part_nonzero(A.data, ia, np.real) | part_nonzero(Bm.data, ib, np.real)
        im_nz = ~diag & (part_nonzero(A.data, ia, np.imag) | part_nonzero(Bm.data, ib, np.imag))

        def src(idx, offset):
            return np.where(idx >= 0, idx + offset, -1)

        r, c = rows[re_nz], cols[re_nz]
        ra, rb = ia[re_nz], ib[re_nz]
        q, d = rows[im_nz], cols[im_nz]
        qa, qb = ia[im_nz], ib[im_nz]
        er = np.concatenate((2 * r, 2 * r + 1, 2 * q, 2 * q + 1))
        ec = np.concatenate((2 * c, 2 * c + 1, 2 * d + 1, 2 * d))
        ea = np.concatenate((ra, ra, src(qa, 2 * na), src(qa, na)))      # Re, Re, -Im, +Im
        eb = np.concatenate((rb, rb, src(qb, 2 * nb), src(qb, nb)))
        n2 = 2 * n
        key = er * n2 + ec
        order = np.argsort(key, kind="stable")
        er = er[order]
        self.f_n = n2
        self.f_rowptr = np.zeros(n2 + 1, dtype=np.int64)
        np.add.at(self.f_rowptr, er + 1, 1)
        np.cumsum(self.f_rowptr, out=self.f_rowptr)
        self.f_col = ec[order].astype(np.int32)
        self.f_a_index = np.ascontiguousarray(ea[order], dtype=np.int64)
        self.f_b_index = np.ascontiguousarray(eb[order], dtype=np.int64)

    @property
    def nnz(self) -> int:
        return int(self.u_col.size)

    def to_device(self):
        """Device copies: A (CSR), B (CSR values, and its diagonal when B is diagonal), and the value
        arrays the factorization gathers from (f_a, f_b)."""
        if self._dev is None:
            dev = device()
            d = {
                "a_rowptr": torch.from_numpy(self.A.indptr.astype(np.int64)).to(dev),
                "a_col": torch.from_numpy(self.A.indices.astype(np.int32)).to(dev),
                "a_val": torch.from_numpy(self.A.data).to(dev),
                "b_val": torch.from_numpy(self._Bfull.data).to(dev),
            }
            if self.b_is_diagonal:
                d["b_diag"] = torch.from_numpy(np.ascontiguousarray(self._Bfull.data.real)).to(dev)
                d["b_rowptr"] = d["b_col"] = None
            else:
                d["b_diag"] = None
                d["b_rowptr"] = torch.from_numpy(self._Bfull.indptr.astype(np.int64)).to(dev)
                d["b_col"] = torch.from_numpy(self._Bfull.indices.astype(np.int32)).to(dev)
            if self.is_complex:
                def ext(v):
                    return torch.cat((v.real, v.imag, -v.imag)).contiguous()
                d["f_a"], d["f_b"] = ext(d["a_val"]), ext(d["b_val"])
            else:
                d["f_a"], d["f_b"] = d["a_val"], d["b_val"]
            self._dev = d
        return self._dev


def same_pattern(p: Pencil, q: Pencil) -> bool:
    """Whether one analysis serves both pencils: the same factorized pattern and value maps."""
    return (p.f_n == q.f_n and p.is_complex == q.is_complex and np.array_equal(p.f_rowptr, q.f_rowptr)
            and np.array_equal(p.f_col, q.f_col) and np.array_equal(p.f_a_index, q.f_a_index)
            and np.array_equal(p.f_b_index, q.f_b_index))


class GpuFactorization:
    """Sparse LDL^T of the pencil on the GPU (SparseFactorization contract)."""

    DOUBLE, SINGLE = 0, 1
    NULL_PIVOT_TOL = {0: 1e-8, 1: 1e-4}  # factor.hpp:279-283

    def __init__(self, pencil: Pencil, precision: str = "double", leaf_size: int = 0, panel_width: int = 0,
                 small_front_max: int = 0, order: np.ndarray | None = None):
        self.pencil = pencil
        self.precision = precision
        self.stats = FactorStats()
        self.detection = True
        self._inertia = Inertia()
        self._handle = C.c_void_p(0)
        self._opts = _lib.AnalysisOptions(leaf_size, panel_width, small_front_max)
        self._given_order = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        self.sigma_held = None
        self._analyze()

    # ------------------------------------------------------------------ analysis
    def _analyze(self):
        p = self.pencil
        t0 = time.perf_counter()
        h = C.c_void_p(0)
        order_ptr = host_ptr(self._given_order) if self._given_order is not None else C.c_void_p(0)
        _lib.call("hsg_analyze", p.f_n, host_ptr(p.f_rowptr), host_ptr(p.f_col), host_ptr(p.f_a_index),
                  host_ptr(p.f_b_index), order_ptr, C.byref(self._opts), C.byref(h))
        self._handle = h
        self.stats.analyses += 1
        self.stats.analyze_s += time.perf_counter() - t0
        info = _lib.AnalysisInfo()
        _lib.call("hsg_analysis_get_info", self._handle, C.byref(info))
        self.info = info
        self.stats.factor_entries = int(info.factor_entries)
        self.stats.flops = float(info.flops)
        self.stats.supernodes = int(info.supernodes)
        self.stats.max_front = int(info.max_front)

    def order(self) -> np.ndarray:
        """Elimination order (new -> old) of the analysed pattern (of the embedding, 2n entries, for a
        complex pencil)."""
        out = np.empty(self.pencil.f_n, dtype=np.int32)
        _lib.call("hsg_analysis_get_order", self._handle, host_ptr(out))
        return out

    def fronts(self) -> dict:
        """Supernodal front structure (see hsg_analysis_get_fronts)."""
        ns = int(self.info.supernodes)
        out = {k: np.empty(ns, dtype=np.int32) for k in ("first", "ncol", "m", "parent")}
        out["rowoff"] = np.empty(ns + 1, dtype=np.int64)
        out["rows"] = np.empty(int(self.info.front_rows), dtype=np.int32)
        _lib.call("hsg_analysis_get_fronts", self._handle, host_ptr(out["first"]), host_ptr(out["ncol"]),
                  host_ptr(out["m"]), host_ptr(out["parent"]), host_ptr(out["rowoff"]), host_ptr(out["rows"]))
        return out

    def estimate(self, precision: str | None = None) -> dict:
        """Estimates of the analysis (FactorEstimate, include/hsolve/factor.hpp:302-310): factor entries,
        flops and the device memory (MB) of one factorization in `precision` (default: the
        object's): panels plus contribution blocks."""
        bytes_per = 8 if (precision or self.precision) == "double" else 4
        return {"factor_entries": int(self.info.factor_entries), "flops": float(self.info.flops),
                "incore_mb": (self.info.panel_bytes_double + self.info.cb_bytes_double) / 8 * bytes_per / 1e6,
                "ooc_mb": 0.0}

    # ------------------------------------------------------------------ numeric
    def set_null_pivot_detection(self, on: bool):
        self.detection = bool(on)

    def factorize(self, alpha: float, beta: float, precision: str | None = None) -> bool:
        """K = alpha A + beta B. False when K is (numerically) singular."""
        prec = self.precision if precision is None else precision
        code = self.DOUBLE if prec == "double" else self.SINGLE
        dev = self.pencil.to_device()
        tol = self.NULL_PIVOT_TOL[code] if self.detection else 0.0
        inert = _lib.Inertia()
        sing = C.c_int32(0)
        t0 = time.perf_counter()
        _lib.call("hsg_factorize", self._handle, float(alpha), float(beta), ptr(dev["f_a"]), ptr(dev["f_b"]),
                  code, tol, stream_ptr(), C.byref(inert), C.byref(sing))
        self.stats.factorize_s += time.perf_counter() - t0
        self.stats.factorizations += 1
        self.stats.null_pivots = int(inert.zero)
        self.stats.two_by_two_pivots = int(inert.two_by_two)
        neg, zero, pos, valid = int(inert.neg), int(inert.zero), int(inert.pos), bool(inert.valid)
        if self.pencil.is_complex:
            # mumps_factorization.cpp:367-380: every eigenvalue of K appears twice in E; an odd
            # negative count means the two copies of one were split by the shift (K singular to
            # working precision), treated like a null pivot
            split = neg % 2 != 0
            neg, zero = neg // 2, (zero + 1) // 2
            pos = max(0, self.pencil.n - neg - zero)
            valid = valid and not split
        self._inertia = Inertia(neg, zero, pos, valid, int(inert.two_by_two))
        self.factored_precision = prec
        self.sigma_held = (alpha, beta, prec, self.detection) if valid else None
        return valid

    def inertia(self) -> Inertia:
        return self._inertia

    def count(self, alpha: float, beta: float) -> Inertia:
        """Sylvester inertia of alpha A + beta B from a double factorization (the factors are kept
        and usable for solves when valid)."""
        self.factorize(alpha, beta, precision="double")
        return self._inertia

    MAX_BATCH = 64  # HSG_MAX_BATCH

    def count_batch(self, items, memory_bytes: float | None = None) -> list[Inertia]:
        """Sylvester inertia of alpha A_z + beta B_z for items [(pencil_z, alpha_z, beta_z), ...],
        every pencil of this object's factorized pattern (same_pattern), from FP64 factorizations
        made MAX_BATCH at a time in the same kernel launches (the pencil index is the launch's z
        dimension), null-pivot detection as set. Each count equals count(alpha_z, beta_z) on that
        pencil alone. memory_bytes bounds the device storage of one batch (panels, contribution
        blocks and workspace of every factorization in it; default 1/4 of the free device memory).
        The batched factors are released afterwards: this object then holds no usable
        factorization."""
        out: list[Inertia] = []
        one = float(self.info.panel_bytes_double + self.info.cb_bytes_double + self.info.workspace_bytes_double)
        if memory_bytes is None:
            memory_bytes = 0.25 * torch.cuda.mem_get_info()[0]
        per_launch = int(max(1, min(self.MAX_BATCH, memory_bytes // max(one, 1.0))))
        for p, _, _ in items:
            if p is not self.pencil and not same_pattern(self.pencil, p):
                raise InvalidInputError("count_batch: a pencil's factorization pattern differs from the analysed one")
        tol = self.NULL_PIVOT_TOL[self.DOUBLE] if self.detection else 0.0
        self._inertia = Inertia(0, 0, 0, False)
        self.sigma_held = None
        try:
            for at in range(0, len(items), per_launch):
                part = items[at:at + per_launch]
                z = len(part)
                devs = [p.to_device() for p, _, _ in part]
                alphas = (C.c_double * z)(*[float(a) for _, a, _ in part])
                betas = (C.c_double * z)(*[float(b) for _, _, b in part])
                av = (C.c_void_p * z)(*[ptr(d["f_a"]).value for d in devs])
                bv = (C.c_void_p * z)(*[ptr(d["f_b"]).value for d in devs])
                res = (_lib.Inertia * z)()
                t0 = time.perf_counter()
                _lib.call("hsg_factorize_batch", self._handle, z, alphas, betas, av, bv, self.DOUBLE, tol, stream_ptr(),
                          res)
                self.stats.factorize_s += time.perf_counter() - t0
                self.stats.factorizations += z
                for r in res:
                    neg, zero, pos, valid = int(r.neg), int(r.zero), int(r.pos), bool(r.valid)
                    if self.pencil.is_complex:  # as in factorize: each eigenvalue of K twice in E
                        split = neg % 2 != 0
                        neg, zero = neg // 2, (zero + 1) // 2
                        pos = max(0, self.pencil.n - neg - zero)
                        valid = valid and not split
                    out.append(Inertia(neg, zero, pos, valid, int(r.two_by_two)))
        finally:
            _lib.call("hsg_release_factors", self._handle)
        return out

    def solve(self, x: torch.Tensor) -> torch.Tensor:
        """In place: x (p, n) contiguous on the device (float64, or complex128 for a complex pencil,
        solved through its real view with the embedding), rows = right-hand sides."""
        if not self._inertia.valid:
            raise RuntimeError("solve without a usable factorization")
        if x.dim() == 1:
            x2 = x.view(1, -1)
        else:
            x2 = x
        want = torch.complex128 if self.pencil.is_complex else torch.float64
        if x2.dtype != want or not x2.is_contiguous() or x2.shape[1] != self.pencil.n:
            raise InvalidInputError(f"solve expects a contiguous {want} (p, n) device tensor")
        t0 = time.perf_counter()
        _lib.call("hsg_factor_solve", self._handle, ptr(x2), x2.shape[0], self.pencil.f_n, stream_ptr())
        self.stats.solve_s += time.perf_counter() - t0
        self.stats.solves += 1
        self.stats.rhs_solved += int(x2.shape[0])
        return x

    def clone(self) -> "GpuFactorization":
        """A factorization object over the same analysis (not recomputed) with numeric storage of
        its own: independent factorizations (spectrum slices, counts) can then be held at the same
        time and run on different streams. Keeps this object alive while the clone exists."""
        other = GpuFactorization.__new__(GpuFactorization)
        other.pencil = self.pencil
        other.precision = self.precision
        other.stats = FactorStats(factor_entries=self.stats.factor_entries, flops=self.stats.flops,
                                  supernodes=self.stats.supernodes)
        other.stats.max_front = getattr(self.stats, "max_front", 0)
        other.detection = True
        other._inertia = Inertia()
        other._opts = self._opts
        other._given_order = self._given_order
        other.sigma_held = None
        other.info = self.info
        other._parent = self
        h = C.c_void_p(0)
        _lib.call("hsg_factor_clone", self._handle, C.byref(h))
        other._handle = h
        return other

    def clone_for(self, pencil: Pencil) -> "GpuFactorization":
        """A factorization object for another pencil of the same pattern (e.g. the Bloch k-points of
        one cell): this object's analysis, the other pencil's values. The factorized pattern and
        the value maps must be identical (same_pattern)."""
        if not same_pattern(self.pencil, pencil):
            raise InvalidInputError("clone_for: the pencil's factorization pattern differs from the analysed one")
        other = self.clone()
        other.pencil = pencil
        return other

    def release_factors(self) -> bool:
        _lib.call("hsg_release_factors", self._handle)
        self._inertia = Inertia(self._inertia.neg, self._inertia.zero, self._inertia.pos, False)
        self.sigma_held = None
        return True

    def close(self) -> None:
        """Frees the factorization object and its device storage now (hipFree waits for the
        device, so solves free their objects when they end rather than leaving it to the garbage
        collector, which could run in the middle of a later solve). Idempotent."""
        self.__del__()
        self._inertia = Inertia(self._inertia.neg, self._inertia.zero, self._inertia.pos, False)
        self.sigma_held = None

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().hsg_factor_destroy(h)
            except Exception:
                pass
            self._handle = C.c_void_p(0)
