"""hsg_gram_noalloc: the allocation-free variant of hsg_gram (fixed device scratch array) gives the
same Gram matrices bit for bit as hsg_gram through the C ABI, on two calls in a row (the second
smaller, so it reads an array holding the first call's partial sums), with and without the
diagonal weight, and rejects a shape whose partial sums exceed the array before launching."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gram(name, V, W, weight):
    import torch
    from hsolve_amd import _lib
    from hsolve_amd.factor import ptr, stream_ptr
    k, n = V.shape
    p = W.shape[0]
    out = torch.full((p, k), float("nan"), dtype=torch.float64, device=V.device)
    rc = getattr(_lib.lib(), name)(n, ptr(V), n, k, ptr(weight), ptr(W), n, p, ptr(out), k, stream_ptr())
    assert rc == 0, _lib.lib().hsg_last_error_message()
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("weighted", [False, True])
def test_noalloc_gram_equals_gram_on_consecutive_calls(gpu, weighted):
    import torch
    rng = np.random.default_rng(11)
    for n, k, p in ((60000, 200, 8), (3001, 37, 3)):
        V = torch.from_numpy(rng.standard_normal((k, n))).to(gpu)
        W = torch.from_numpy(rng.standard_normal((p, n))).to(gpu)
        wd = torch.from_numpy(rng.uniform(0.5, 2.0, n)).to(gpu) if weighted else None
        ref = _gram("hsg_gram", V, W, wd)
        got = _gram("hsg_gram_noalloc", V, W, wd)
        print("n", n, "k", k, "p", p, "weighted", weighted, "max |diff|", float(np.max(np.abs(ref - got))), flush=True)
        assert np.array_equal(ref, got)
        # and both agree with the FP64 product on the host
        Vh, Wh = V.cpu().numpy(), W.cpu().numpy()
        if weighted:
            Wh = Wh * wd.cpu().numpy()[None, :]
        np.testing.assert_allclose(got, Wh @ Vh.T, rtol=1e-12, atol=1e-9)


def test_noalloc_gram_rejects_oversized_shapes(gpu):
    import torch
    from hsolve_amd import _lib
    from hsolve_amd.factor import ptr, stream_ptr
    # min(32, n / 256) = 32 row groups x 16 x 8192 doubles = 32 MiB of partial sums (> 16 MiB)
    n, k, p = 8192, 8192, 16
    V = torch.zeros((k, n), dtype=torch.float64, device=gpu)
    W = torch.zeros((p, n), dtype=torch.float64, device=gpu)
    out = torch.zeros((p, k), dtype=torch.float64, device=gpu)
    rc = _lib.lib().hsg_gram_noalloc(n, ptr(V), n, k, None, ptr(W), n, p, ptr(out), k, stream_ptr())
    assert rc == 1  # HSG_INVALID_INPUT, nothing launched
    assert b"scratch" in _lib.lib().hsg_last_error_message()
