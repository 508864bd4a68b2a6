"""Layer-2 kernels (spmm, diagonal spmm, shifted residual, Gram, block update, column dots) against
plain PyTorch references of the same operations, real FP64 and complex128, called through the C ABI.

Tolerances: the kernels sum in a different order than torch, so results agree to a few ulps of the
summed magnitudes (relative 1e-12 of the operand norms)."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

N, K, P = 3001, 70, 11   # odd sizes: partial tiles in every kernel


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _ptr(t):
    return None if t is None else t.data_ptr()


def _rand(rng, shape, cplx):
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
    return x


def _hermitian(rng, n, cplx):
    m = sp.random(n, n, density=8.0 / n, random_state=rng, format="csr")
    if cplx:
        m = m + 1j * sp.random(n, n, density=8.0 / n, random_state=rng, format="csr")
    m = (m + m.conj().T) * 0.5 + sp.identity(n) * 3.0
    m = sp.csr_matrix(m)
    m.sort_indices()
    return m


def _name(base, cplx):
    return ("hsg_z" if cplx else "hsg_") + base


@pytest.mark.parametrize("cplx", [False, True])
def test_spmm_and_diag(gpu, cplx):
    import torch
    from hsolve_amd import _lib
    from hsolve_amd.factor import stream_ptr
    rng = np.random.default_rng(5)
    A = _hermitian(rng, N, cplx)
    X = _rand(rng, (P, N), cplx)
    Y0 = _rand(rng, (P, N), cplx)
    d = rng.uniform(0.5, 2.0, N)
    rp, ci, av = _t(A.indptr.astype(np.int64), gpu), _t(A.indices.astype(np.int32), gpu), _t(A.data, gpu)
    x, y = _t(X, gpu), _t(Y0, gpu)
    _lib.call(_name("csr_spmm", cplx), N, _ptr(rp), _ptr(ci), _ptr(av), _ptr(x), N, _ptr(y), N, P, 2.0, -0.5,
              stream_ptr())
    want = 2.0 * (A @ X.T).T - 0.5 * Y0
    got = y.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    dd = _t(d, gpu)
    _lib.call(_name("diag_spmm", cplx), N, _ptr(dd), _ptr(x), N, _ptr(y), N, P, 1.5, 0.0, stream_ptr())
    assert np.abs(y.cpu().numpy() - 1.5 * d * X).max() <= 1e-13 * np.abs(X).max()
    torch.cuda.synchronize()


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("weighted", [False, True])
def test_gram_update_dots(gpu, cplx, weighted):
    import torch
    from hsolve_amd import _lib
    from hsolve_amd.factor import stream_ptr
    rng = np.random.default_rng(7)
    V = _rand(rng, (K, N), cplx)
    W = _rand(rng, (P, N), cplx)
    w = rng.uniform(0.5, 2.0, N) if weighted else None
    v, wt, wd = _t(V, gpu), _t(W, gpu), (_t(w, gpu) if weighted else None)
    dt = torch.complex128 if cplx else torch.float64
    c = torch.empty((P, K), dtype=dt, device=gpu)
    _lib.call(_name("gram", cplx), N, _ptr(v), N, K, _ptr(wd), _ptr(wt), N, P, _ptr(c), K, stream_ptr())
    Ww = W * w if weighted else W
    want = V.conj() @ Ww.T                   # (K, P): C = V^H diag(w) W
    got = c.cpu().numpy().T
    scale = np.linalg.norm(V, axis=1)[:, None] * np.linalg.norm(Ww, axis=1)[None, :]
    assert np.all(np.abs(got - want) <= 1e-12 * scale)
    # W2 = 0.5 W2 - V C
    C = _rand(rng, (P, K), cplx)          # row c = column c of the K x P matrix
    w2, cd = _t(W, gpu), _t(C, gpu)
    _lib.call(_name("block_update", cplx), N, _ptr(v), N, K, _ptr(cd), K, _ptr(w2), N, P, -1.0, 0.5,
              stream_ptr())
    want_u = 0.5 * W - (V.T @ C.T).T
    assert np.abs(w2.cpu().numpy() - want_u).max() <= 1e-12 * np.abs(want_u).max()
    # column dots x_c^H diag(w) y_c
    out = torch.empty(P, dtype=dt, device=gpu)
    _lib.call(_name("column_dots", cplx), N, _ptr(wt), N, _ptr(w2), N, P, _ptr(wd), _ptr(out), stream_ptr())
    want_d = np.sum(W.conj() * (w2.cpu().numpy() * (w if weighted else 1.0)), axis=1)
    sc = np.linalg.norm(W, axis=1) * np.linalg.norm(w2.cpu().numpy(), axis=1) * (2.0 if weighted else 1.0)
    assert np.all(np.abs(out.cpu().numpy() - want_d) <= 1e-12 * sc)
    torch.cuda.synchronize()


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("b_kind", ["csr", "diag"])
def test_shifted_residual(gpu, cplx, b_kind):
    import torch
    from hsolve_amd import _lib
    from hsolve_amd.factor import stream_ptr
    rng = np.random.default_rng(11)
    A = _hermitian(rng, N, cplx)
    if b_kind == "csr":
        B = _hermitian(rng, N, cplx)
        bd = None
    else:
        bd = rng.uniform(0.5, 2.0, N)
        B = sp.diags(bd, format="csr")
    X = _rand(rng, (P, N), cplx)
    Y = _rand(rng, (P, N), cplx)
    sigma = 1.75
    d = {k: None for k in ("brp", "bci", "bv", "bd")}
    if b_kind == "csr":
        d["brp"], d["bci"], d["bv"] = (_t(B.indptr.astype(np.int64), gpu), _t(B.indices.astype(np.int32), gpu),
                                       _t(B.data, gpu))
    else:
        d["bd"] = _t(bd, gpu)
    dt = torch.complex128 if cplx else torch.float64
    r = torch.empty((P, N), dtype=dt, device=gpu)
    rn = torch.empty(P, dtype=torch.float64, device=gpu)
    om = torch.empty(P, dtype=torch.float64, device=gpu)
    # every device array stays referenced until the kernel has run (a temporary freed while the
    # call's arguments are built could be reused by the next one before the launch)
    arp, aci, av = _t(A.indptr.astype(np.int64), gpu), _t(A.indices.astype(np.int32), gpu), _t(A.data, gpu)
    x, y = _t(X, gpu), _t(Y, gpu)
    _lib.call(_name("shifted_residual", cplx), N, _ptr(arp), _ptr(aci), _ptr(av), _ptr(d["brp"]), _ptr(d["bci"]),
              _ptr(d["bv"]), _ptr(d["bd"]), sigma, _ptr(x), N, _ptr(y), N, _ptr(r), N, P, _ptr(rn), _ptr(om),
              stream_ptr())
    torch.cuda.synchronize()
    Wv = X + sigma * Y
    R = (B @ Wv.T).T - (A @ Y.T).T
    den = (abs(B) @ np.abs(Wv).T).T + (abs(A) @ np.abs(Y).T).T
    got = r.cpu().numpy()
    assert np.abs(got - R).max() <= 1e-12 * np.abs(den).max()
    np.testing.assert_allclose(rn.cpu().numpy(), np.sum(np.abs(R) ** 2, axis=1), rtol=1e-11)
    np.testing.assert_allclose(om.cpu().numpy(), np.max(np.abs(R) / den, axis=1), rtol=1e-9)
    torch.cuda.synchronize()
