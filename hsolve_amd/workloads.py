"""Benchmark workloads (the five configurations the engine is measured on).

  cavity20   PEC cube [0, pi]^3, 20^3 Yee cells, standard form (n = 21,660), sigma = 10, k = 21
  box48      PEC box [0, 2pi] x [0, pi] x [0, 1.5pi], 48 x 24 x 36 cells (n = 117,036), FIT
             A = C^T M_nu C, B = edge mass x (1 + 0.05 U(-1, 1)) seeded; all eigenvalues in
             [40, b] with b set by inertia bisection so that the window holds 500
  bloch16    periodic unit cell, 16^3 cells (n = 12,288), eps = 12 inside a centred sphere of
             radius 0.3 a, Bloch phases of a 4 x 4 x 4 grid over [0, pi/a]^3 (endpoints included),
             window [0.01, 1.0] (2 pi / a)^2 per k-point
  box80      PEC box on 80 x 70 x 64 cells (n = 1,045,014), jittered mass, sigma mid-spectrum, k = 20
  serving    a seeded request stream with Poisson arrivals: PEC boxes of aspect ratios U(1, 2) with
             n log-uniform over [n_min, n_max] (12..60 cells per side), B = I or jittered diagonal
             mass, k uniform in {5, ..., 40}, sigma uniform inside the closed-form spectrum; a
             fraction of Bloch requests (the bloch16 cell at a random k-point of its grid, its window)
The inputs are generated by hsolve_amd.gen.yee (seed 2026 unless stated).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .gen import yee


@dataclass
class Workload:
    name: str
    A: object
    B: object
    target: dict
    meta: dict


def cavity20() -> Workload:
    A, B, g = yee.cube_cavity(20)
    return Workload("cavity20", A, B, {"mode": "nearest", "sigma": 10.0, "k": 21},
                    {"n": A.shape[0], "cells": [20, 20, 20]})


def count_below(A, B, s: float) -> int:
    """nu(A - s B) on the GPU engine (double factorization, null-pivot detection off)."""
    from .factor import GpuFactorization, Pencil
    F = GpuFactorization(Pencil(A, B))
    F.set_null_pivot_detection(False)
    F.factorize(1.0, -s, precision="double")
    return F.inertia().neg


def window_upper(A, B, lower: float, count: int, guess: float, tol: float = 1e-9) -> float:
    """b > lower with nu(b) - nu(lower) == count, by bisection on exact GPU counts; returns the
    middle of the gap found between the count-th and the next eigenvalue."""
    from .factor import GpuFactorization, Pencil
    F = GpuFactorization(Pencil(A, B))
    F.set_null_pivot_detection(False)

    def nu(s):
        F.factorize(1.0, -s, precision="double")
        return F.inertia().neg

    base = nu(lower)
    target = base + count
    lo, hi = lower, max(guess, lower * 1.01)
    while nu(hi) < target + 1:
        lo, hi = hi, hi + 2.0 * (hi - lower)
    # narrow to the bracket (l, h) with nu(l) <= target < target + 1 <= nu(h)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        c = nu(mid)
        if c <= target:
            lo = mid
        else:
            hi = mid
        if hi - lo <= tol * hi:
            break
    # lo: nu(lo) == target (count-th eigenvalue below lo); the next eigenvalue lies in (lo, hi]
    if nu(lo) != target:
        raise RuntimeError("window_upper: bisection did not isolate the count")
    return lo


def box48(seed: int = 2026, lower: float = 40.0, count: int = 500, upper: float | None = None) -> Workload:
    A, B, g = yee.jittered_box(seed=seed)
    if upper is None:
        upper = window_upper(A, B, lower, count, guess=55.0)
    return Workload("box48", A, B, {"mode": "interval", "lower": lower, "upper": upper},
                    {"n": A.shape[0], "cells": [48, 24, 36], "seed": seed, "count": count})


def bloch_kpoints(per_dim: int = 4):
    ks = np.linspace(0.0, math.pi, per_dim)
    return [(a, b, c) for a in ks for b in ks for c in ks]


def bloch16(kpoint=(0.0, 0.0, 0.0)) -> Workload:
    A, B, g = yee.bloch_cell(16, kpoint)
    scale = (2 * math.pi) ** 2
    return Workload("bloch16", A, B, {"mode": "interval", "lower": 0.01 * scale, "upper": 1.0 * scale},
                    {"n": A.shape[0], "kpoint": list(kpoint)})


def box80(seed: int = 2026, k: int = 20) -> Workload:
    A, B, g = yee.jittered_box(cells=(80, 70, 64), lengths=(80 / 64 * math.pi, 70 / 64 * math.pi, math.pi), seed=seed)
    # mid-spectrum shift: a quarter of the spectrum's top 12 / h^2 (about 4 x 10^5 eigenvalues
    # below), moved by 0.0137 off the homogeneous box's eigenvalue that sits exactly there
    hmin = min(g.spacing)
    sigma = 0.25 * 12.0 / hmin ** 2 + 0.0137
    return Workload("box80", A, B, {"mode": "nearest", "sigma": sigma, "k": k},
                    {"n": A.shape[0], "cells": [80, 70, 64], "seed": seed})


def _box_spectrum_range(cells, lengths) -> tuple[float, float]:
    """Smallest nonzero and largest eigenvalue of the homogeneous PEC box pencil: per dimension the
    1-D terms (2 / h sin(m pi / 2N))^2, m = 0..N-1; the smallest mode has one index 0 and two 1's
    (a mode needs two nonzero indices), the largest all indices N - 1."""
    t1 = sorted((2.0 * c / L * math.sin(math.pi / (2 * c))) ** 2 for c, L in zip(cells, lengths))
    top = sum((2.0 * c / L * math.sin((c - 1) * math.pi / (2 * c))) ** 2 for c, L in zip(cells, lengths))
    return t1[0] + t1[1], top


def serving_stream(count: int = 2000, seed: int = 2026, rate: float = 4.0, n_min: float = 5e3, n_max: float = 6e5,
                   bloch_fraction: float = 0.2, bloch_cells: int = 16) -> list[dict]:
    """Requests of the serving workload, sorted by arrival: dicts with A, B, config (target,
    tol 1e-10), arrival (seconds; exponential gaps of mean 1 / rate), kind and n. Box cells per side
    are the lengths (pi times U(1, 2) each) over one spacing chosen so that 3 Nx Ny Nz is the drawn
    n, clipped to 12..60 cells per side. Matrices are built here, so keep count modest (a 6e5 box
    takes about a second to assemble)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    kpts = bloch_kpoints(4)
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        if rng.uniform() < bloch_fraction:
            kp = kpts[int(rng.integers(len(kpts)))]
            w = bloch16(kp) if bloch_cells == 16 else None
            if w is None:
                A, B, _ = yee.bloch_cell(bloch_cells, kp)
                scale = (2 * math.pi) ** 2
                w = Workload("bloch", A, B, {"mode": "interval", "lower": 0.01 * scale, "upper": scale}, {})
            out.append({"A": w.A, "B": w.B, "config": {"target": dict(w.target), "tol": 1e-10}, "arrival": t,
                        "kind": "bloch", "n": int(w.A.shape[0]), "kpoint": list(kp)})
            continue
        aspect = rng.uniform(1.0, 2.0, 3)
        lengths = tuple(float(math.pi * a) for a in aspect)
        n_want = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        h = (float(np.prod(lengths)) / (n_want / 3.0)) ** (1.0 / 3.0)
        cells = tuple(int(min(60, max(12, round(L / h)))) for L in lengths)
        jitter = bool(rng.uniform() < 0.5)
        A, B, g = yee.jittered_box(cells, lengths, seed=int(rng.integers(1 << 31)), jitter=0.05 if jitter else 0.0)
        if not jitter:
            A, B = yee.standard_form(A, B), None
        lo, hi = _box_spectrum_range(cells, lengths)
        sigma = float(rng.uniform(lo, hi))
        k = int(rng.integers(5, 41))
        out.append({"A": A, "B": B, "config": {"target": {"mode": "nearest", "sigma": sigma, "k": k}, "tol": 1e-10},
                    "arrival": t, "kind": "box-ghep" if jitter else "box-hep", "n": int(A.shape[0]),
                    "cells": list(cells)})
    return out
