"""The norm estimate of the solve set-up (estimate_norm2: 30 Lanczos steps, max |Ritz value|, as
the reference's norms.cpp:60-95): the recurrence run by the library in one call
(hsg_lanczos_coefficients) gives the estimate the step-by-step loop gives with the same kernels,
for real and complex pencils, CSR and diagonal matrices; and the estimate is a lower bound of the
2-norm, reached to a few digits on these small matrices (30 steps span most of their spectrum
ends)."""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.linalg as la

from hsolve_amd.driver_core import estimate_norm2
from hsolve_amd.gen import yee
from hsolve_amd.ops import DeviceOps

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kvec", [(0.0, 0.0, 0.0), (math.pi / 3, 0.5, 0.0)])
def test_native_recurrence_matches_python_loop(gpu, kvec):
    from hsolve_amd.factor import Pencil
    A, B, _ = yee.bloch_cell(4, kvec)
    p = Pencil(A, B)
    ops = DeviceOps(p)
    for apply, M in ((ops.apply_a, A), (ops.apply_b, B)):
        native = estimate_norm2(ops, apply, 30, 2026)
        loop = estimate_norm2(ops, lambda x, _f=apply: _f(x), 30, 2026)  # not A or B itself: the Python loop
        print(f"k={kvec} native {native:.17g} loop {loop:.17g}", flush=True)
        assert abs(native - loop) <= 1e-12 * loop
        true = float(np.max(np.abs(la.eigvalsh(M.toarray()))))
        assert native <= true * (1 + 1e-12) and native >= 0.9 * true


def test_native_recurrence_breakdown_truncates(gpu):
    # B = 2 I: the Krylov space of a multiple of the identity breaks down after one step
    import scipy.sparse as sp
    from hsolve_amd.factor import Pencil
    n = 50
    A = sp.diags(np.linspace(1.0, 3.0, n), 0, format="csr") + sp.diags([np.ones(n - 1) * 0.1] * 2, [-1, 1])
    B = sp.identity(n, format="csr") * 2.0
    ops = DeviceOps(Pencil(A, B))
    assert estimate_norm2(ops, ops.apply_b, 30, 2026) == pytest.approx(2.0, rel=1e-14)
