"""A-posteriori bounds on the eigenvector subspace angles of a certified result (report.subspace,
the per-pair subspace_sin_bound), following the reference's attach_subspace_bound /
subspace_bounds (src/driver/subspace.cpp:512-848, report fields include/hsolve/solver.hpp:349-374).

The bound is the Davis-Kahan sin(Theta) theorem in the B inner product: for a group of returned
pairs (X_g, Lambda_g) with R_g = A X_g - B X_g Lambda_g,

    || sin Theta(span X_g, E_g) ||_F <= || B^{-1/2} R_g ||_F / (delta_g sqrt(1 - eps_g)),

E_g the exact spectral subspace the group approximates, delta_g the distance from the group's
Ritz values to every eigenvalue outside E_g, eps_g = || X_g^H B X_g - I ||_2. The gaps come from
the completeness certificate: the certified window [a, b] holds exactly `count` eigenvalues
(Sylvester counts), so every other eigenvalue lies outside [a, b].

  * When the returned pairs in the window number `count` (all of the window's eigenvalues are
    returned), the pairs are grouped: neighbours closer than cluster_rtol (|lambda| + scale) or than
    their residuals join; each group's exact eigenvalues lie in its Kahan interval
    [lo - eta_g, hi + eta_g], eta_g = rho_g / s + sqrt(1 + eps_g) (hi - lo) eps_g / (s (1 + s)),
    s = sqrt(1 - eps_g); groups whose intervals meet merge. If every interval stays inside the
    window, a group's gap is the distance to the window ends and to its neighbours' intervals
    ("cluster" groups, E_g of dimension = the group's size).
  * Otherwise (fewer pairs returned than the window holds: a nearest target's k cutting a
    multiple eigenvalue; or an interval reaching a window end) all pairs of the window form one
    "window" group, compared with the count-dimensional spectral subspace of the window, gap =
    the distance from the pairs' extreme values to the window ends.

||B^{-1/2} r|| is exact for B = I and diagonal B (sum r_i^2 / b_ii); for a general B the bound is
not computed (noted). Rounding of the sums is covered by a factor (1 + gamma) with
gamma = 2 k u / (1 - 2 k u) (Higham's gamma_k with a factor 2 of slack), as in the reference.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

U = 2.0 ** -53
CLUSTER_RTOL = 1e-6        # pairs closer than this (relative to |lambda| + scale) share a group
MAX_ORTHOGONALITY_DEFECT = 0.5


def gamma(k: float) -> float:
    t = 2.0 * k * U
    return t / (1.0 - t) if t < 0.5 else math.inf


def _down(x: float) -> float:
    return math.nextafter(x, -math.inf) if math.isfinite(x) else x


def _up(x: float) -> float:
    return math.nextafter(x, math.inf) if math.isfinite(x) else x


class _Group:
    def __init__(self, idx):
        self.idx = list(idx)
        self.lo = self.hi = 0.0
        self.rho = 0.0
        self.eps = 0.0
        self.eta = 0.0
        self.left = -math.inf
        self.right = math.inf


def residual_norms(ops, x: torch.Tensor, lam: np.ndarray) -> np.ndarray:
    """||B^{-1/2} (A x_j - lambda_j B x_j)||_2 per row of the (m, n) device block x (B = I or
    diagonal), computed on the device."""
    m = x.shape[0]
    if m == 0:
        return np.zeros(0)
    ax = ops.apply_a(x)
    bx = ops.apply_b(x)
    lam_d = torch.as_tensor(np.asarray(lam, dtype=np.float64), device=x.device)
    r = ax - lam_d[:, None].to(ax.dtype) * bx
    w = None if ops.pencil.b_is_identity else (1.0 / ops.d["b_diag"])
    r2 = ops.dots(r, r, weight=w)
    return np.sqrt(np.maximum(r2.cpu().numpy().astype(np.float64), 0.0))


def gram_defect(ops, x: torch.Tensor, idx, bnorm2: np.ndarray) -> float:
    """|| X_g^H B X_g - I ||_2 of the rows idx of x (rounding of the Gram entries included);
    bnorm2[j] = x_j^H B x_j serves single-pair groups without another device call."""
    rnd = gamma(ops.n + 2.0)
    if len(idx) == 1:
        d = float(bnorm2[idx[0]])
        return abs(d - 1.0) * (1.0 + rnd) + rnd * abs(d)
    xs = x[list(idx)].contiguous()
    g = ops.b_gram(xs, xs).cpu().numpy().T
    g = 0.5 * (g + g.conj().T)
    d = np.linalg.eigvalsh(g - np.eye(g.shape[0]))
    return float(np.max(np.abs(d))) * (1.0 + rnd) + rnd * float(np.max(np.abs(np.diag(g))))


def bounds(ops, x: torch.Tensor, lam, windows, scale: float) -> tuple[dict, list]:
    """report.subspace and the per-pair bounds (None where no bound) for the returned pairs
    (rows of x, eigenvalues lam) and the certified windows [(lower, upper, count), ...]."""
    t0 = time.perf_counter()
    lam = np.asarray(lam, dtype=np.float64)
    m = lam.size
    rep = {"performed": True, "max_sin_bound": None, "rigorous": True, "failure_probability": 0.0,
           "residual_norm": "2" if ops.pencil.b_is_identity else "b_inverse_diagonal",
           "b_min": 1.0, "b_min_source": "identity", "groups": [], "note": "", "seconds": 0.0}
    if not ops.pencil.b_is_identity:
        bd = ops.d["b_diag"]
        rep["b_min"] = float(bd.min().item())
        rep["b_min_source"] = "diagonal"
    rho = residual_norms(ops, x, lam) * (1.0 + gamma(64.0))
    bnorm2 = ops.b_dots(x, x).cpu().numpy().astype(np.float64)
    per_pair = [None] * m
    issues = []
    placed = np.zeros(m, dtype=bool)

    def describe(g: _Group, gap: float, kind: str, lower: float, upper: float, dim: int):
        bound = g.rho / (gap * math.sqrt(1.0 - g.eps)) * (1.0 + gamma(4.0)) if gap > 0 else math.inf
        ok = gap > 0 and g.eps < MAX_ORTHOGONALITY_DEFECT and math.isfinite(bound)
        sin_b = min(bound, 1.0) if ok else None
        for j in g.idx:
            per_pair[j] = sin_b
        rep["groups"].append({"kind": kind, "pairs": [int(j) for j in g.idx], "unreturned": max(0, dim - len(g.idx)),
                              "lower": lower, "upper": upper, "dimension": int(dim), "gap": gap,
                              "residual": g.rho, "orthogonality": g.eps, "sin_bound": sin_b})

    def measure(g: _Group):
        g.lo = float(np.min(lam[g.idx]))
        g.hi = float(np.max(lam[g.idx]))
        g.rho = math.sqrt(float(np.sum(rho[g.idx] ** 2))) * (1.0 + gamma(len(g.idx) + 2.0))
        g.eps = gram_defect(ops, x, g.idx, bnorm2)
        if not g.eps < MAX_ORTHOGONALITY_DEFECT:
            g.eta, g.left, g.right = math.inf, -math.inf, math.inf
            return
        s = math.sqrt(1.0 - g.eps)
        excess = g.eps / (s * (1.0 + s))
        g.eta = (g.rho / s + math.sqrt(1.0 + g.eps) * (g.hi - g.lo) * excess) * (1.0 + gamma(8.0))
        g.left, g.right = _down(g.lo - g.eta), _up(g.hi + g.eta)

    for lower, upper, count in windows:
        inside = [j for j in range(m) if not placed[j] and lower <= lam[j] < upper]
        if not inside:
            continue
        placed[inside] = True
        inside.sort(key=lambda j: lam[j])
        rng = f"[{lower:.10g}, {upper:.10g})"
        if len(inside) > count:
            issues.append(f"{len(inside)} pairs in the window {rng} of {count} eigenvalues")
            continue

        def window_group():
            g = _Group(inside)
            measure(g)
            describe(g, min(_down(g.lo - lower), _down(upper - g.hi)), "window", lower, upper, count)

        if len(inside) < count:
            window_group()
            continue
        groups: list[_Group] = []
        for j in inside:
            if groups:
                prev = groups[-1].idx[-1]
                close = CLUSTER_RTOL * (max(abs(lam[j]), abs(lam[prev])) + scale)
                if lam[j] - lam[prev] <= close or lam[j] - lam[prev] <= rho[j] + rho[prev]:
                    groups[-1].idx.append(j)
                    continue
            groups.append(_Group([j]))
        for g in groups:
            measure(g)
        while True:  # merge neighbours whose intervals meet
            meet = next((q for q in range(len(groups) - 1) if groups[q].right >= groups[q + 1].left), None)
            if meet is None:
                break
            groups[meet].idx.extend(groups[meet + 1].idx)
            del groups[meet + 1]
            measure(groups[meet])
        if not (math.isfinite(groups[0].eta) and groups[0].left > lower and groups[-1].right < upper):
            window_group()
            continue
        for q, g in enumerate(groups):
            gap = min(_down(g.lo - lower), _down(upper - g.hi))
            if q > 0:
                gap = min(gap, _down(g.lo - groups[q - 1].right))
            if q + 1 < len(groups):
                gap = min(gap, _down(groups[q + 1].left - g.hi))
            describe(g, gap, "cluster", g.left, g.right, len(g.idx))
    without = sum(1 for v in per_pair if v is None)
    if without:
        issues.insert(0, f"{without} of {m} pair(s) without a bound (outside the certified windows, no gap, or "
                         "vectors not numerically independent)")
    elif m:
        rep["max_sin_bound"] = float(max(per_pair))
    rep["note"] = "; ".join(issues)
    rep["seconds"] = time.perf_counter() - t0
    return rep, per_pair


def attach(ctx, rep: dict, sel, cfg: dict) -> None:
    """Fills rep['subspace'] and rep['pairs'][j]['subspace_sin_bound'] (as attach_subspace_bound,
    src/driver/subspace.cpp:848-905): computed when subspace_bound is not off, the certificate
    passed and B is the identity or diagonal."""
    out = rep["subspace"]
    cert = rep["certificate"]
    t0 = time.perf_counter()
    out.update({"performed": False, "max_sin_bound": None, "rigorous": False, "failure_probability": 1.0,
                "groups": [], "note": ""})
    if cfg["subspace_bound"] == "off":
        out["note"] = "disabled by configuration"
    elif not cert["performed"]:
        out["note"] = "no completeness certificate" + (f" ({cert['note']})" if cert["note"] else "")
    elif not cert["passed"]:
        out["note"] = "the completeness certificate did not pass"
    elif not (ctx.ops.pencil.b_is_identity or ctx.ops.pencil.b_is_diagonal):
        out["note"] = "not computed for a non-diagonal B (the B^-1 norm of the residuals needs B's inverse)"
    elif sel is None or sel.size == 0 or sel.x is None:
        out["note"] = "no pairs"
    else:
        lower, upper = cert["lower"], cert["upper"]
        if not (math.isfinite(lower) and math.isfinite(upper)):
            sl = rep.get("slices") or []
            if sl:
                lower = min(lower, min(s["lower"] for s in sl)) if math.isfinite(lower) else min(s["lower"] for s in sl)
                upper = max(upper, max(s["upper"] for s in sl)) if math.isfinite(upper) else max(s["upper"] for s in sl)
        if not (math.isfinite(lower) and math.isfinite(upper)):
            out["note"] = "the certified window has an open end"
        else:
            res, per_pair = bounds(ctx.ops, sel.x, np.asarray(sel.lam, dtype=np.float64),
                                   [(lower, upper, int(cert["count_inertia"]))], rep["scale"])
            out.update(res)
            out["rigorous"] = bool(out["rigorous"] and cert["rigorous"])
            out["failure_probability"] = cert["failure_probability"]
            for j, v in enumerate(per_pair):
                if j < len(rep["pairs"]):
                    rep["pairs"][j]["subspace_sin_bound"] = v
    out["seconds"] = time.perf_counter() - t0
