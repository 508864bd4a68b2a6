"""The CPU oracle's solver (oracle/eigs_cpu.py, the bench's CPU baseline path) against the
reference's own golden vectors: examples/refs/*/eigenvalues.npy are complete discrete-exact
spectra of the stored matrices (ref.json kind 'discrete-exact'). CPU only."""
import json
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from oracle import eigs_cpu
from oracle import evaluate as ev

EX = Path(__file__).resolve().parent / "golden" / "examples"


def _case(name):
    d = EX / name
    A = sp.load_npz(d / "A.npz").tocsr()
    B = sp.load_npz(d / "B.npz").tocsr() if (d / "B.npz").exists() else None
    return A, B, json.loads((d / "config.json").read_text())["target"], np.real(np.load(d / "eigenvalues.npy"))


def test_nearest_matches_golden_spectrum():
    A, B, t, ref = _case("near-singular-shift")
    out = eigs_cpu.solve_nearest(A, B, t["sigma"], t["k"], tol=1e-11)
    want = ref[np.argsort(np.abs(ref - t["sigma"]), kind="stable")[: t["k"]]]
    assert out["passed"]
    assert np.allclose(np.sort(out["eigenvalues"]), np.sort(want), rtol=0, atol=1e-9 * np.abs(ref).max())
    res = ev.evaluate(A, B, out["eigenvalues"], out["eigenvectors"], {"mode": "nearest", **t}, ref)
    assert res["passed"], res


def test_slice_matches_golden_interval():
    A, B, t, ref = _case("interval-slicing")
    out = eigs_cpu.solve_slice(A, B, t["lower"], t["upper"], tol=1e-11)
    want = ref[(ref >= t["lower"]) & (ref <= t["upper"])]
    assert out["passed"] and out["count_inertia"] == want.size
    got = np.sort(out["eigenvalues"][(out["eigenvalues"] >= t["lower"]) & (out["eigenvalues"] <= t["upper"])])
    assert np.allclose(got, np.sort(want), rtol=0, atol=1e-9 * np.abs(ref).max())


def test_factor_solve_sample_solves_the_permuted_pencil():
    A, B, t, ref = _case("near-singular-shift")
    n = A.shape[0]
    perm = np.random.default_rng(3).permutation(n)
    out = eigs_cpu.factor_solve_sample(A, B, 0.5 * (ref[3] + ref[4]), perm, p=4)
    assert out["relative_residual"] < 1e-10 and out["lu_entries"] >= A.nnz
    assert out["factor_s"] > 0 and out["solve_s"] > 0 and out["p"] == 4
