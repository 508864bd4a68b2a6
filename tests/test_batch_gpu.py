"""Batched pencils (hsolve_amd.batch): Bloch k-points of one photonic cell solved together share
one analysis per factorized pattern and run on concurrent streams; every pencil's result is
bitwise the one a standalone one-stream solve of that pencil gives (the shared analysis is the
analysis the pencil would have computed: the ordering depends on the pattern only)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from hsolve_amd.gen import yee
from oracle import evaluate as ev

pytestmark = pytest.mark.gpu


def test_bloch_batch_matches_standalone(gpu):
    import scipy.linalg as la
    from hsolve_amd import batch, solver
    ks = [0.0, math.pi / 3, math.pi]
    kpts = [(a, b, c) for a in ks for b in ks[:2] for c in ks[1:2]]      # 6 k-points, 4 pattern classes
    pencils = [yee.bloch_cell(5, k)[:2] for k in kpts]
    cfg = {"target": {"mode": "interval", "lower": 15.0, "upper": 70.0}, "tol": 1e-10}
    out = batch.solve_batch(pencils, cfg, workers=3)
    notes = out[0][2]["parallel"]["notes"][-1]
    print(notes, flush=True)
    assert "6 pencils" in notes and "4 analyses" in notes
    for (A, B), (lam, X, rep, _) in zip(pencils, out):
        l1, X1, r1, _ = solver.solve(A, B, cfg, streams=1)
        assert np.array_equal(lam, l1) and np.array_equal(X, X1)
        assert rep["certificate"]["passed"] and rep["success"]
        w = la.eigvalsh(A.toarray(), B.toarray())
        res = ev.evaluate(A, B, lam, X, cfg["target"], w)
        assert res["passed"], res
