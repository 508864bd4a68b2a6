"""Batched factorizations (hsg_factorize_batch, GpuFactorization.count_batch): several pencils of
one factorized pattern, or several shifts of one pencil, factorized in the same kernel launches
(the batch index is the launch's z dimension). Every count must be the one a factorization of that
pencil at that shift alone gives: same pivots, hence identical neg / zero / pos / 2x2 counts, and
the same singular verdict. After a batch the object holds no solvable factorization; the next
ordinary factorization of the same object solves as before."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
import torch

from hsolve_amd import _lib
from hsolve_amd.factor import GpuFactorization, Pencil, ptr, stream_ptr
from hsolve_amd.gen import yee

pytestmark = pytest.mark.gpu


def _single(F, pencil, alpha, beta, precision="double"):
    G = F.clone_for(pencil)
    try:
        G.factorize(alpha, beta, precision=precision)
        return G.inertia()
    finally:
        G.close()


def test_bloch_kpoints_batch_equals_one_at_a_time(gpu):
    # k-points complex in the same two directions: one factorized pattern
    kpts = [(kx, ky, 0.0) for kx in (math.pi / 4, math.pi / 3, math.pi / 2) for ky in (0.4, 1.1)]
    pencils = [Pencil(*yee.bloch_cell(6, k)[:2]) for k in kpts]
    F = GpuFactorization(pencils[0])
    try:
        shifts = [12.37, 30.21, 55.13]
        items = [(p, 1.0, -s) for p in pencils for s in shifts]
        got = F.count_batch(items)
        print(f"batch of {len(items)} (n = {pencils[0].n}, complex)", flush=True)
        # the same batch again on the same object (its counters reused, not freshly allocated)
        again = F.count_batch(items)
        assert [(g.neg, g.zero, g.pos, g.valid) for g in again] == [(g.neg, g.zero, g.pos, g.valid) for g in got]
        for (p, a, b), g in zip(items, got):
            want = _single(F, p, a, b)
            assert (g.neg, g.zero, g.pos, g.valid, g.two_by_two) == \
                   (want.neg, want.zero, want.pos, want.valid, want.two_by_two), (b, g, want)
            assert g.neg + g.zero + g.pos == p.n
        assert sum(g.valid for g in got) >= len(got) - 2
        # counts grow with the shift for every k-point
        for i in range(len(pencils)):
            negs = [got[i * len(shifts) + j].neg for j in range(len(shifts))]
            assert negs == sorted(negs) and negs[0] < negs[-1]
    finally:
        F.close()


def test_shift_batch_over_64_splits_and_matches_dense(gpu):
    import scipy.linalg as la
    import scipy.sparse as sp
    rng = np.random.default_rng(11)
    n = 240  # a tridiagonal pencil with a random diagonal B: well separated eigenvalues
    A = sp.diags([-np.ones(n - 1), rng.uniform(0.0, 10.0, n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    B = sp.diags(rng.uniform(0.5, 2.0, n), 0, format="csr")
    p = Pencil(A, B)
    w = la.eigvalsh(A.toarray(), B.toarray())
    gaps = [0.5 * (w[i] + w[i + 1]) for i in range(len(w) - 1) if w[i + 1] - w[i] > 1e-3]
    shifts = gaps[:: max(1, len(gaps) // 70)][:70]
    assert len(shifts) > 64
    F = GpuFactorization(p)
    try:
        got = F.count_batch([(p, 1.0, -s) for s in shifts])
        print(f"{len(shifts)} shifts in 2 batches (n = {p.n})", flush=True)
        for s, g in zip(shifts, got):
            assert g.valid and g.neg == int(np.count_nonzero(w < s)), s
        # the object has no solvable factorization now; an ordinary one works again
        x = torch.randn((2, p.n), dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
        with pytest.raises(RuntimeError):
            F.solve(x.clone())
        F.factorize(1.0, -shifts[3], precision="double")
        y = x.clone()
        F.solve(y)
        K = (A - shifts[3] * B).toarray()
        ref = np.linalg.solve(K, x.cpu().numpy().T).T
        assert np.allclose(y.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    finally:
        F.close()


def test_single_precision_batch_equals_one_at_a_time(gpu):
    A, B, _ = yee.bloch_cell(5, (math.pi / 3, 0.0, 0.0))
    p = Pencil(A, B)
    F = GpuFactorization(p)
    try:
        shifts = [8.0, 21.0, 47.5]
        z = len(shifts)
        d = p.to_device()
        alphas = (C.c_double * z)(*([1.0] * z))
        betas = (C.c_double * z)(*[-s for s in shifts])
        av = (C.c_void_p * z)(*([ptr(d["f_a"]).value] * z))
        bv = (C.c_void_p * z)(*([ptr(d["f_b"]).value] * z))
        res = (_lib.Inertia * z)()
        tol = F.NULL_PIVOT_TOL[F.SINGLE]
        _lib.call("hsg_factorize_batch", F._handle, z, alphas, betas, av, bv, F.SINGLE, tol, stream_ptr(), res)
        for s, r in zip(shifts, res):
            G = F.clone()
            try:
                inert = _lib.Inertia()
                sing = C.c_int32(0)
                _lib.call("hsg_factorize", G._handle, 1.0, -s, ptr(d["f_a"]), ptr(d["f_b"]), F.SINGLE, tol, stream_ptr(),
                          C.byref(inert), C.byref(sing))
                assert (r.neg, r.zero, r.pos, r.valid, r.two_by_two) == \
                       (inert.neg, inert.zero, inert.pos, inert.valid, inert.two_by_two), s
            finally:
                G.close()
    finally:
        F.close()


def test_batch_argument_errors(gpu):
    A, B, _ = yee.bloch_cell(3, (0.0, 0.0, 0.0))
    p = Pencil(A, B)
    F = GpuFactorization(p)
    try:
        d = p.to_device()
        one = (C.c_double * 65)(*([1.0] * 65))
        av = (C.c_void_p * 65)(*([ptr(d["f_a"]).value] * 65))
        res = (_lib.Inertia * 65)()
        from hsolve_amd.errors import HsolveError
        for count in (0, 65):
            with pytest.raises(HsolveError, match="count must be in"):
                _lib.call("hsg_factorize_batch", F._handle, count, one, one, av, None, F.DOUBLE, 0.0, stream_ptr(), res)
        other = Pencil(*yee.bloch_cell(3, (0.5, 0.0, 0.0))[:2])  # complex: another pattern
        with pytest.raises(Exception, match="pattern"):
            F.count_batch([(other, 1.0, -1.0)])
    finally:
        F.close()
