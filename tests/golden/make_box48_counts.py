"""Writes tests/golden/box48_counts.json: Sylvester counts nu(A - s B) of the 48 x 24 x 36 PEC box
pencils (n = 117,036) from SuperLU (oracle/inertia.py's elimination: symmetric mode, diagonal
pivots), run on the CPU.

* Unjittered FIT pencil (B = edge mass): the counts at gap points of the closed-form spectrum
  (oracle/yee_exact.py) are computed and must equal the closed form; this pins the SuperLU count
  itself at this size before its jittered counts are trusted.
* Jittered pencil (B = edge mass x (1 + 0.05 U(-1, 1)), numpy default_rng(2026)): counts at fixed
  shifts, the fixture the GPU engine's counts are compared with.

The elimination order handed to SuperLU is a fill-reducing nested dissection of the pattern (the
engine's host analysis, which needs no GPU); the counts do not depend on the order (Sylvester's
law of inertia), only the elimination's size does. About 10 s per count on one core.

    python tests/golden/make_box48_counts.py
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from hsolve_amd.factor import GpuFactorization, Pencil  # noqa: E402  (host analysis only)
from hsolve_amd.gen import yee  # noqa: E402
from oracle import yee_exact  # noqa: E402

CELLS = (48, 24, 36)
LENGTHS = (2 * np.pi, np.pi, 1.5 * np.pi)
SEED = 2026
JITTERED_SHIFTS = [40.0, 42.5, 45.0, 50.0, 55.0]


def superlu_count(A, B, s, perm):
    K = (sp.csc_matrix(A) - s * sp.csc_matrix(B)).tocsc()[perm][:, perm].tocsc()
    lu = sla.splu(K, permc_spec="NATURAL", diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))
    if not np.array_equal(lu.perm_r, lu.perm_c):
        raise RuntimeError("SuperLU left the diagonal")
    d = lu.U.diagonal()
    return int(np.sum(d < 0)), int(np.sum(d == 0))


def main():
    out = {"cells": list(CELLS), "lengths": list(LENGTHS), "seed": SEED, "jitter": 0.05, "unjittered": [],
           "jittered": []}
    A0, B0, g = yee.jittered_box(CELLS, LENGTHS, seed=SEED, jitter=0.0)
    perm = GpuFactorization(Pencil(A0, B0)).order().astype(np.int64)
    ref = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    pts = yee_exact.gap_points(ref, 40.0, 56.0)
    pts = pts[np.linspace(0, pts.size - 1, 5).astype(int)]
    for s in pts:
        t0 = time.perf_counter()
        neg, zero = superlu_count(A0, B0, float(s), perm)
        exact = yee_exact.count_below(ref, float(s))
        print(f"unjittered s={s:.10g} superlu={neg} zero={zero} closed form={exact} ({time.perf_counter() - t0:.1f} s)",
              flush=True)
        if neg != exact or zero:
            raise SystemExit("SuperLU count differs from the closed form: fixture not written")
        out["unjittered"].append({"s": float(s), "nu": exact})
    A, B, _ = yee.jittered_box(CELLS, LENGTHS, seed=SEED)
    for s in JITTERED_SHIFTS:
        t0 = time.perf_counter()
        neg, zero = superlu_count(A, B, s, perm)
        print(f"jittered s={s} superlu={neg} zero={zero} ({time.perf_counter() - t0:.1f} s)", flush=True)
        out["jittered"].append({"s": s, "nu": neg, "zero": zero})
    (Path(__file__).parent / "box48_counts.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
