"""Packed solves: many independent diffusion problems in one library context.

The reference solves one mesh per process (z02/main.py:202-227).  For batches and for
serving, the meshes of several requests of the same dimension are packed into one union
mesh whose components are the requests: every operator apply, smoothing step, DG<->P1CG
transfer and V-cycle level is one kernel launch over all of them, while each component
keeps its own SFC hierarchy (built from its own vertices), its own Krylov space and its
own convergence test.  A component's result is the one it gets when solved alone, up to
the solver tolerance.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .diffusion import GMRES_ITS, GMRES_M, DGDiffusion


@dataclass
class Packed:
    """Union of meshes of one dimension; component c owns elements
    [elem_off[c], elem_off[c+1]) and vertices [vert_off[c], vert_off[c+1]).  The P3 node
    coordinates stay in the components' own arrays (``x_parts``); ``x_ref_in`` concatenates
    them on first use."""
    D: int
    x_parts: list
    cells: np.ndarray
    verts: np.ndarray
    elem_off: np.ndarray
    vert_off: np.ndarray

    @property
    def x_ref_in(self):
        return np.concatenate(self.x_parts)


def pack(meshes) -> Packed:
    """Concatenate meshes (``meshgen.Mesh``-like: D, x_ref_in, cells, verts) of one
    dimension, renumbering each mesh's vertices after the previous ones."""
    if not meshes:
        raise ValueError("no meshes to pack")
    D = meshes[0].D
    if any(m.D != D for m in meshes):
        raise ValueError("packed meshes must share one dimension")
    eo = np.zeros(len(meshes) + 1, dtype=np.int64)
    vo = np.zeros(len(meshes) + 1, dtype=np.int64)
    for i, m in enumerate(meshes):
        eo[i + 1] = eo[i] + m.cells.shape[0]
        vo[i + 1] = vo[i] + m.verts.shape[0]
    cells = np.concatenate([np.asarray(m.cells, dtype=np.int64) + vo[i]
                            for i, m in enumerate(meshes)])
    parts = [np.ascontiguousarray(m.x_ref_in, dtype=np.float64) for m in meshes]
    return Packed(D, parts, cells, np.concatenate([m.verts for m in meshes]), eo, vo)


class DGBatch:
    """Independent P3 DG diffusion solves of one dimension packed into one context.
    The default smoother is the exact block inverse that the reference's
    variable-diffusivity solver uses (z05/volume_mf_um_diff_varprop.py:384-392)."""

    def __init__(self, meshes, kappas, device="cuda:0", face_kappa="harmonic",
                 blk_solver="direct", graph_vcycle=True, alloc_limit=0, params=None):
        """``params``: solver parameters (``dgmi_set_param`` names and values) set on the
        context before its multigrid set-up."""
        p = pack(meshes)
        kappa = np.concatenate([np.asarray(k, dtype=np.float64).reshape(-1) for k in kappas])
        self.space = DGDiffusion(p.x_parts, p.cells, p.verts, kappa, device=device,
                                 face_kappa=face_kappa, blk_solver=blk_solver,
                                 components=(p.elem_off, p.vert_off), alloc_limit=alloc_limit)
        self.packed = p
        self.ncomp = len(meshes)
        self.nloc = self.space.nloc
        self.dof_off = p.elem_off * self.nloc
        try:
            # replay the packed V-cycle from a captured graph, or launch its kernels one by one
            self.space.lib.dgmi_set_param(self.space.ctx, b"graph_vcycle",
                                          1.0 if graph_vcycle else 0.0)
            for name, value in (params or {}).items():
                self.space.set_param(name, value)
            self.space._get_RAR_and_sfc_data_Um()
        except BaseException:
            self.space.close()  # the device memory goes back at once, not at collection
            raise

    def rhs(self, fs, u_bcs):
        """b = M f + Dirichlet terms of every component (nodal f and u_bc per mesh)."""
        # each mesh's data goes straight to its offset in the packed device vector
        f = self.space._upload(fs, "f")
        u = self.space._upload(u_bcs, "u_bc")
        return self.space._get_rhs(self.space.new_vector(), f, u)

    def solve(self, b, rtol=1e-10, inner=torch.float32, m=GMRES_M, max_its=GMRES_ITS, x=None):
        """Solve every component to ||r_c|| <= rtol ||b_c||.  Returns (x, stats) with one
        stats dict per component."""
        x = self.space.new_vector() if x is None else x
        self.space._gmres_mg_solver(x, b, rtol=rtol, inner=inner, m=m, max_its=max_its)
        st = self.space.last_stats
        return x, (st if isinstance(st, list) else [st])

    def split(self, v):
        """Per-component views of a packed DOF vector."""
        return [v[self.dof_off[c]:self.dof_off[c + 1]] for c in range(self.ncomp)]

    def close(self, synced=False):
        """Give the device memory back (``synced``: the caller's queued work has finished,
        see ``DGDiffusion.close``)."""
        self.space.close(synced=synced)
