"""Closed-form discrete spectra of homogeneous Yee/FIT pencils (oracle).

Restates the formula of the reference's generator, tools/gen/yee.py:231-260 (module docstring
:24-31): for a PEC box the eigenvalues are sum_d (2/h_d sin(m_d pi / (2 N_d)))^2, m_d = 0..N_d-1,
twice when all three m_d >= 1, once when exactly one vanishes, never when two vanish, plus the
gradient null space of dimension (Nx-1)(Ny-1)(Nz-1); for a Bloch box sum_d (2/h_d sin(q_d h_d / 2))^2
with q_d = k_d + 2 pi m_d / L_d, twice per wave vector, plus one zero per wave vector (three for
the constant fields at k = 0). The per-dimension terms are summed in sorted order so that equal
index triples give bitwise-equal eigenvalues.

Pinned by tests/test_oracle_yee.py against dense symmetric eigenvalues of the assembled matrices
on small grids (mathematics: the assembled operator's spectrum), the reference's own check
(tests/py/test_gen_yee.py:30-77 compares the same formula with dense solvers to 1e-13 * scale).
"""
from __future__ import annotations

import numpy as np


def pec_box_spectrum(cells, lengths, eps: float = 1.0, mu: float = 1.0) -> np.ndarray:
    cells = [int(c) for c in cells]
    h = [float(L) / c for L, c in zip(lengths, cells)]
    terms = [(2.0 / h[d] * np.sin(np.arange(cells[d]) * np.pi / (2 * cells[d]))) ** 2 for d in range(3)]
    idx = np.meshgrid(*[np.arange(c) for c in cells], indexing="ij")
    s = np.sort(np.stack([terms[d][idx[d].ravel()] for d in range(3)], axis=1), axis=1)
    g = (s[:, 0] + s[:, 1]) + s[:, 2]
    active = np.count_nonzero(np.stack([terms[d][idx[d].ravel()] > 0 for d in range(3)], axis=1), axis=1)
    mult = np.where(active == 3, 2, np.where(active == 2, 1, 0))
    null = (cells[0] - 1) * (cells[1] - 1) * (cells[2] - 1)
    w = np.concatenate((np.repeat(g, mult), np.zeros(null)))
    n_edges = cells[0] * (cells[1] - 1) * (cells[2] - 1) + (cells[0] - 1) * cells[1] * (cells[2] - 1) + \
        (cells[0] - 1) * (cells[1] - 1) * cells[2]
    assert w.size == n_edges, "mode count does not match the edge count"
    return np.sort(w / (eps * mu))


def bloch_box_spectrum(cells, lengths, phases, eps: float = 1.0, mu: float = 1.0) -> np.ndarray:
    cells = [int(c) for c in cells]
    h = [float(L) / c for L, c in zip(lengths, cells)]
    terms = []
    for d in range(3):
        m = np.arange(cells[d])
        arg = float(phases[d]) / (2 * cells[d]) + m * np.pi / cells[d]
        terms.append((2.0 / h[d] * np.sin(arg)) ** 2)
    idx = np.meshgrid(*[np.arange(c) for c in cells], indexing="ij")
    s = np.sort(np.stack([terms[d][idx[d].ravel()] for d in range(3)], axis=1), axis=1)
    g = (s[:, 0] + s[:, 1]) + s[:, 2]
    mult = np.where(g > 0, 2, 0)
    null = cells[0] * cells[1] * cells[2] + (2 if all(float(p) % (2 * np.pi) == 0.0 for p in phases) else 0)
    w = np.concatenate((np.repeat(g, mult), np.zeros(null)))
    return np.sort(w / (eps * mu))


def count_below(spectrum: np.ndarray, s: float) -> int:
    """nu(s) = #{lambda < s} (the Sylvester count of A - s B for an HPD B)."""
    return int(np.searchsorted(np.sort(spectrum), s, side="left"))


def gap_points(spectrum: np.ndarray, lo: float, hi: float, rel: float = 1e-6) -> np.ndarray:
    """Mid-points of the gaps between distinct eigenvalues in [lo, hi] (shifts far from eigenvalues)."""
    w = np.sort(spectrum)
    w = w[(w >= lo) & (w <= hi)]
    if w.size < 2:
        return np.empty(0)
    d = np.unique(np.round(w, 12))
    mids = 0.5 * (d[1:] + d[:-1])
    gaps = d[1:] - d[:-1]
    return mids[gaps > rel * np.maximum(1.0, np.abs(mids))]
