"""Synthetic jittered structured meshes and manufactured diffusion problems.

The reference reads gmsh meshes (z02/config.py:53); there are none in its
repository, so the workloads are generated here:

* 2D: the unit square split into n x n cells, each cut into two counter-clockwise
  triangles along the same diagonal;
* 3D: the unit cube split into n^3 hexahedra, each cut into the six Kuhn tetrahedra
  that share the main diagonal (conforming across cells).

Interior vertices are jittered uniformly by +-jitter*h per coordinate from a seed.
P3 nodes sit at their affine (straight-sided) positions, in the reference's local
node order, and every element is oriented so that det J > 0 with J rows x_k - x_D
(the orientation the reference enforces in z02/mesh_init.py:276-289 for tets and
emits counter-clockwise for triangles, z02/mesh_init.py:19-31).

All boundary faces are Dirichlet faces (``glb_bcface_type`` 0).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np

# node multi-indices (barycentric, sum 3) in the reference's local order; the same
# tables appear in the product's C++ element tables and in the oracle.
TET_P3 = np.array([
    (3, 0, 0, 0), (0, 3, 0, 0), (0, 0, 3, 0), (0, 0, 0, 3),
    (2, 0, 1, 0), (1, 0, 2, 0), (2, 1, 0, 0), (1, 2, 0, 0),
    (0, 2, 1, 0), (0, 1, 2, 0), (2, 0, 0, 1), (1, 0, 0, 2),
    (0, 2, 0, 1), (0, 1, 0, 2), (0, 0, 2, 1), (0, 0, 1, 2),
    (0, 1, 1, 1), (1, 1, 1, 0), (1, 0, 1, 1), (1, 1, 0, 1),
], dtype=np.float64) / 3.0
TRI_P3 = np.array([
    (3, 0, 0), (0, 3, 0), (0, 0, 3),
    (2, 1, 0), (1, 2, 0), (0, 2, 1), (0, 1, 2),
    (1, 0, 2), (2, 0, 1), (1, 1, 1),
], dtype=np.float64) / 3.0


@dataclass
class Mesh:
    """A P3 DG mesh in the reference's layouts.

    x_ref_in : (nele, D, nloc) float64 -- P3 node coordinates per element
    cells    : (nele, D+1) int64       -- P1 (vertex) connectivity, global vertex ids
    verts    : (nverts, D) float64     -- vertex coordinates
    """
    D: int
    x_ref_in: np.ndarray
    cells: np.ndarray
    verts: np.ndarray

    @property
    def nele(self) -> int:
        return self.cells.shape[0]

    @property
    def nloc(self) -> int:
        return self.x_ref_in.shape[2]

    @property
    def nverts(self) -> int:
        return self.verts.shape[0]


def _jittered_grid(n: int, D: int, jitter: float, seed: int) -> np.ndarray:
    h = 1.0 / n
    axes = [np.linspace(0.0, 1.0, n + 1)] * D
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, D)
    idx = np.stack(np.meshgrid(*([np.arange(n + 1)] * D), indexing="ij"), axis=-1).reshape(-1, D)
    interior = np.all((idx > 0) & (idx < n), axis=1)
    rng = np.random.default_rng(seed)
    delta = rng.uniform(-jitter * h, jitter * h, size=grid.shape)
    grid[interior] += delta[interior]
    return grid


def _p3_nodes(verts: np.ndarray, cells: np.ndarray, table: np.ndarray) -> np.ndarray:
    v = verts[cells]                      # (nele, D+1, D)
    x = np.einsum("nk,ekd->edn", table, v)  # (nele, D, nloc)
    return np.ascontiguousarray(x)


def _orient(verts: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """Swap the first two vertices where det(x_k - x_D) < 0."""
    v = verts[cells]
    D = verts.shape[1]
    J = v[:, :D, :] - v[:, D:D + 1, :]
    det = np.linalg.det(J)
    flip = det < 0
    cells = cells.copy()
    cells[flip, 0], cells[flip, 1] = cells[flip, 1].copy(), cells[flip, 0].copy()
    return cells


def square_mesh(n: int, jitter: float = 0.2, seed: int = 0) -> Mesh:
    """n x n cells of the unit square, two triangles per cell (2 n^2 elements)."""
    verts = _jittered_grid(n, 2, jitter, seed)
    vid = lambda i, j: i * (n + 1) + j  # noqa: E731  (x index i, y index j)
    cells = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            cells.append((a, b, c))
            cells.append((a, c, d))
    cells = _orient(verts, np.array(cells, dtype=np.int64))
    return Mesh(D=2, x_ref_in=_p3_nodes(verts, cells, TRI_P3), cells=cells, verts=verts)


def _morton_order(idx: np.ndarray) -> np.ndarray:
    """Order of integer grid cells (rows of idx) along the Z-order curve (bits of the
    coordinates interleaved)."""
    idx = np.asarray(idx, dtype=np.int64)
    code = np.zeros(idx.shape[0], dtype=np.int64)
    d = idx.shape[1]
    for bit in range(21):
        for a in range(d):
            code |= ((idx[:, a] >> bit) & 1) << (bit * d + a)
    return np.argsort(code, kind="stable")


def cube_mesh(n: int, jitter: float = 0.15, seed: int = 0) -> Mesh:
    """n^3 cells of the unit cube, six Kuhn tetrahedra per cell (6 n^3 elements)."""
    verts = _jittered_grid(n, 3, jitter, seed)
    m = n + 1
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)  # (ncell, 3)
    cells = []
    for perm in itertools.permutations(range(3)):
        p = np.zeros((base.shape[0], 4, 3), dtype=np.int64)
        p[:, 0] = base
        step = np.zeros(3, dtype=np.int64)
        for s, ax in enumerate(perm):
            step = step.copy()
            step[ax] += 1
            p[:, s + 1] = base + step
        cells.append(p)
    # cell-major, 6 tets per cell, cells in Morton (Z-curve) order, so that the face
    # neighbours of an element lie close to it in memory in all three directions
    p = np.stack(cells, axis=1)[_morton_order(base)].reshape(-1, 4, 3)
    ids = (p[..., 0] * m + p[..., 1]) * m + p[..., 2]
    cells = _orient(verts, ids.astype(np.int64))
    return Mesh(D=3, x_ref_in=_p3_nodes(verts, cells, TET_P3), cells=cells, verts=verts)


# --------------------------------------------------------------------------------------
# manufactured problems
# --------------------------------------------------------------------------------------
@dataclass
class Problem:
    """Right-hand-side data in the reference's per-DG-node layout
    (z02/bc_f.py:892-958): nodal source ``f`` and Dirichlet data ``u_bc``,
    plus the per-element diffusivity ``kappa`` (z05/volume_mf_um_diff_varprop.py:567)."""
    mesh: Mesh
    kappa: np.ndarray   # (nele,)
    f: np.ndarray       # (nele*nloc,)
    u_bc: np.ndarray    # (nele*nloc,)
    u_exact: np.ndarray  # (nele*nloc,) manufactured solution at the DG nodes


def sine_problem(mesh: Mesh, sigma: float = 0.0, kappa_seed: int = 1,
                 amplitude: float = 1.0) -> Problem:
    """u = A prod sin(pi x_d); kappa_E = exp(sigma N(0,1)) per element (sigma = 0 gives
    kappa = 1); f = D pi^2 kappa_E u at the DG nodes; u_bc = u at the DG nodes."""
    D = mesh.D
    x = mesh.x_ref_in  # (nele, D, nloc)
    u = amplitude * np.prod(np.sin(np.pi * x), axis=1)  # (nele, nloc)
    if sigma > 0:
        rng = np.random.default_rng(kappa_seed)
        kappa = np.exp(sigma * rng.standard_normal(mesh.nele))
    else:
        kappa = np.ones(mesh.nele)
    f = D * np.pi ** 2 * kappa[:, None] * u
    return Problem(mesh=mesh, kappa=kappa, f=f.reshape(-1).copy(),
                   u_bc=u.reshape(-1).copy(), u_exact=u.reshape(-1).copy())
