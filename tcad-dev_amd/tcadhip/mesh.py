"""Tensor-product meshes for the structured 2D/3D drift-diffusion path.

Node ordering follows the reference: k = j*Nx + i in 2D
(pytcad/pytcad/mesh2d.py:1-8) and k = kk*Nx*Ny + j*Nx + i in 3D
(pytcad/pytcad/mesh3d.py:1-9), x fastest.

The workload meshes have fixed node counts (60x40, 200x120,
64x48x48), so instead of the reference's h_min/h_max-driven
`graded_mesh` (pytcad/pytcad/mesh.py:30-114, whose node count is an
output) the builders here take the node count as an input:
`geometric_mesh` grades from one end with an exact first cell, and
`focused_mesh` places nodes at equal increments of the arc length
t(x) = int dx / s(x) of the same gradient-limited size function
s(x) = min(h_max, h_min + (ratio - 1) * dist(x, foci)) that
`graded_mesh` uses, with the cell count fixed.
"""
from dataclasses import dataclass

import numpy as np


def control_volume_widths(h):
    """Box widths: interior nodes take half of each adjacent interval,
    endpoints one half-interval (pytcad/pytcad/mesh2d.py:15-28)."""
    h = np.asarray(h, dtype=float)
    dv = np.zeros(h.size + 1)
    dv[1:-1] = 0.5 * (h[:-1] + h[1:])
    dv[0] = 0.5 * h[0]
    dv[-1] = 0.5 * h[-1]
    return dv


def uniform_mesh(L, n_nodes):
    return np.linspace(0.0, float(L), int(n_nodes))


def geometric_mesh(L, n_nodes, h0):
    """n_nodes nodes on [0, L] whose first cell is exactly h0 and whose
    cells grow by a constant ratio r (found by bisection)."""
    n_cells = int(n_nodes) - 1
    L = float(L)
    if h0 * n_cells >= L:
        return uniform_mesh(L, n_nodes)

    def total(r):
        return h0 * (r ** n_cells - 1.0) / (r - 1.0)

    lo, hi = 1.0 + 1e-12, 2.0
    while total(hi) < L:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) < L:
            lo = mid
        else:
            hi = mid
    r = 0.5 * (lo + hi)
    x = np.concatenate(([0.0], np.cumsum(h0 * r ** np.arange(n_cells))))
    x *= L / x[-1]
    x[0], x[-1] = 0.0, L
    return x


def focused_mesh(L, foci, n_nodes, h_min, h_max, ratio=1.15):
    """n_nodes nodes on [0, L], refined around `foci` (equal arc-length
    increments of 1/s(x); see the module docstring)."""
    L = float(L)
    foci = np.clip(np.atleast_1d(np.asarray(foci, dtype=float)), 0.0, L)
    m = 200001
    xs = np.linspace(0.0, L, m)
    d = np.min(np.abs(xs[:, None] - foci[None, :]), axis=1)
    s = np.minimum(h_max, h_min + (ratio - 1.0) * d)
    inv = 1.0 / s
    t = np.concatenate(([0.0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(xs))))
    t *= (n_nodes - 1) / t[-1]
    x = np.interp(np.arange(n_nodes, dtype=float), t, xs)
    x[0], x[-1] = 0.0, L
    return x


@dataclass
class Mesh2D:
    """pytcad/pytcad/mesh2d.py:31-56."""
    x: np.ndarray
    y: np.ndarray

    def __post_init__(self):
        self.x = np.asarray(self.x, dtype=float)
        self.y = np.asarray(self.y, dtype=float)
        self.Nx, self.Ny = self.x.size, self.y.size
        self.N = self.Nx * self.Ny

    def idx(self, i, j):
        return j * self.Nx + i


@dataclass
class Mesh3D:
    """pytcad/pytcad/mesh3d.py:20-52."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray

    def __post_init__(self):
        self.x = np.asarray(self.x, dtype=float)
        self.y = np.asarray(self.y, dtype=float)
        self.z = np.asarray(self.z, dtype=float)
        self.Nx, self.Ny, self.Nz = self.x.size, self.y.size, self.z.size
        self.N = self.Nx * self.Ny * self.Nz

    def idx(self, i, j, k):
        return k * self.Nx * self.Ny + j * self.Nx + i
