"""Yee-grid (finite integration) curl-curl pencils for the benchmark workloads.

Restates the discretisation of the reference's generator (tools/gen/yee.py:1-40 for the scheme,
:167-193 for the curl operator, :206-228 for the FIT assembly and the standard form) in this
package's own code, plus the workload definitions the benchmarks use:

  * unknowns are line integrals on primal edges; PEC walls drop the edges lying in them, a
    Bloch-periodic box keeps all 3 Nx Ny Nz edges and multiplies the wrapped-around entries of
    the 1-D difference by exp(i k_d L_d);
  * C (faces x edges) is the topological curl, A = C^H M_nu C, B = M_eps with diagonal Hodge
    matrices M_eps = diag(eps_e |e~|/|e|), M_nu = diag(nu_f |f~|/|f|);
  * edge blocks x, y, z in C order of (cell, node, node), (node, cell, node), (node, node, cell);
    face blocks x, y, z in C order of (node, cell, cell), (cell, node, cell), (cell, cell, node).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

__all__ = ["YeeGrid", "curl", "assemble", "standard_form", "cube_cavity", "jittered_box",
           "bloch_cell", "edge_midpoints"]


@dataclass(frozen=True)
class YeeGrid:
    cells: tuple
    lengths: tuple = (1.0, 1.0, 1.0)
    boundary: str = "pec"          # "pec" or "bloch"
    phases: tuple = (0.0, 0.0, 0.0)  # k_d * L_d, Bloch only

    @property
    def spacing(self):
        return tuple(float(L) / int(N) for L, N in zip(self.lengths, self.cells))

    def nodes(self, d):
        return int(self.cells[d]) - 1 if self.boundary == "pec" else int(self.cells[d])

    @property
    def edge_counts(self):
        c = [int(x) for x in self.cells]
        n = [self.nodes(d) for d in range(3)]
        return (c[0] * n[1] * n[2], n[0] * c[1] * n[2], n[0] * n[1] * c[2])

    @property
    def n(self):
        return int(sum(self.edge_counts))

    @property
    def is_complex(self):
        return any(not _is_real_phase(p) for p in self.phases)


def _is_real_phase(p):
    r = float(p) / np.pi
    return abs(r - round(r)) == 0.0


def _wrap(p):
    r = float(p) / np.pi
    if r == round(r):
        return 1.0 if int(round(r)) % 2 == 0 else -1.0
    return complex(np.exp(1j * float(p)))


def _diff(grid: YeeGrid, d: int) -> sp.csr_matrix:
    """cells x unknown nodes: (D u)_i = u_{i+1} - u_i (PEC: u_0 = u_N = 0; Bloch: u_N = phase u_0)."""
    N = int(grid.cells[d])
    if grid.boundary == "pec":
        i = np.arange(N)
        rows = np.concatenate([i[:-1], i[1:]])
        cols = np.concatenate([i[:-1], i[1:] - 1])
        vals = np.concatenate([np.ones(N - 1), -np.ones(N - 1)])
        return sp.csr_matrix((vals, (rows, cols)), shape=(N, N - 1))
    dtype = complex if grid.is_complex else float
    i = np.arange(N)
    rows = np.concatenate([i, i])
    cols = np.concatenate([i, (i + 1) % N])
    vals = np.concatenate([-np.ones(N), np.ones(N)]).astype(dtype)
    vals[-1] = _wrap(grid.phases[d])
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def _k3(a, b, c):
    return sp.kron(sp.kron(a, b, format="csr"), c, format="csr")


def curl(grid: YeeGrid) -> sp.csr_matrix:
    """Topological curl (faces x edges); (C u)_x = d_y u_z - d_z u_y and cyclic."""
    D = [_diff(grid, d) for d in range(3)]
    In = [sp.identity(grid.nodes(d), format="csr") for d in range(3)]
    Ic = [sp.identity(int(grid.cells[d]), format="csr") for d in range(3)]
    return sp.bmat([
        [None, -_k3(In[0], Ic[1], D[2]), _k3(In[0], D[1], Ic[2])],
        [_k3(Ic[0], In[1], D[2]), None, -_k3(D[0], In[1], Ic[2])],
        [-_k3(Ic[0], D[1], In[2]), _k3(D[0], Ic[1], In[2]), None],
    ], format="csr")


def edge_midpoints(grid: YeeGrid) -> np.ndarray:
    """(n, 3) coordinates of the edge midpoints, box [0, L]^3, in the unknown order."""
    h = grid.spacing
    pts = []
    for d in range(3):
        axes = []
        for e in range(3):
            if e == d:
                axes.append((np.arange(int(grid.cells[e])) + 0.5) * h[e])
            else:
                off = 1 if grid.boundary == "pec" else 0
                axes.append((np.arange(grid.nodes(e)) + off) * h[e])
        X, Y, Z = np.meshgrid(*axes, indexing="ij")
        pts.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1))
    return np.concatenate(pts, axis=0)


def assemble(grid: YeeGrid, eps_edge=1.0, nu_face=1.0):
    """FIT pencil (A, B) = (C^H M_nu C, M_eps); eps_edge / nu_face: scalars or per-edge/face arrays."""
    hx, hy, hz = grid.spacing
    ne = grid.edge_counts
    edge_geom = np.concatenate([np.full(ne[0], hy * hz / hx), np.full(ne[1], hx * hz / hy),
                                np.full(ne[2], hx * hy / hz)])
    C = curl(grid)
    nf = C.shape[0]
    c = [int(x) for x in grid.cells]
    n = [grid.nodes(d) for d in range(3)]
    fcount = (n[0] * c[1] * c[2], c[0] * n[1] * c[2], c[0] * c[1] * n[2])
    assert sum(fcount) == nf
    face_geom = np.concatenate([np.full(fcount[0], hx / (hy * hz)), np.full(fcount[1], hy / (hx * hz)),
                                np.full(fcount[2], hz / (hx * hy))])
    m_eps = edge_geom * np.asarray(eps_edge, dtype=float)
    m_nu = face_geom * np.asarray(nu_face, dtype=float)
    A = (C.conj().T @ sp.diags(m_nu) @ C).tocsr()
    A = ((A + A.conj().T) * 0.5).tocsr()
    A.sum_duplicates()
    A.eliminate_zeros()
    A.sort_indices()
    B = sp.diags(m_eps, format="csr")
    return A, B


def standard_form(A: sp.csr_matrix, B: sp.csr_matrix) -> sp.csr_matrix:
    s = sp.diags(1.0 / np.sqrt(B.diagonal().real))
    M = (s @ A @ s).tocsr()
    M = ((M + M.conj().T) * 0.5).tocsr()
    M.sort_indices()
    return M


# ------------------------------------------------------------------ benchmark workloads

def cube_cavity(cells: int = 20):
    """PEC cube [0, pi]^3, cells^3 Yee cells: standard form A = C^T C / h^2, B = I."""
    g = YeeGrid((cells, cells, cells), (np.pi, np.pi, np.pi), "pec")
    A, B = assemble(g)
    return standard_form(A, B), None, g


def jittered_box(cells=(48, 24, 36), lengths=(2 * np.pi, np.pi, 1.5 * np.pi), seed: int = 2026,
                 jitter: float = 0.05):
    """PEC box GHEP: A = C^T M_nu C (FIT), B = edge mass times (1 + jitter U(-1, 1)) (seeded)."""
    g = YeeGrid(tuple(cells), tuple(lengths), "pec")
    A, B = assemble(g)
    rng = np.random.default_rng(seed)
    d = B.diagonal() * (1.0 + jitter * rng.uniform(-1.0, 1.0, size=g.n))
    return A, sp.diags(d, format="csr"), g


def bloch_cell(cells: int = 16, kvec=(0.0, 0.0, 0.0), eps_in: float = 12.0, radius: float = 0.3):
    """Periodic unit cell (a = 1): eps = eps_in on edges whose midpoint lies within radius of the
    centre, else 1; Bloch phases k_d a on the wrapped edges. Returns (A, B, grid)."""
    g = YeeGrid((cells, cells, cells), (1.0, 1.0, 1.0), "bloch", tuple(float(k) for k in kvec))
    mid = edge_midpoints(g)
    r = np.linalg.norm(mid - 0.5, axis=1)
    eps = np.where(r <= radius, eps_in, 1.0)
    A, B = assemble(g, eps_edge=eps)
    return A, B, g
