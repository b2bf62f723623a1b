"""Shared helpers for the GPU parity tests (problem builders, oracle comparisons)."""
import numpy as np
import torch

from dg_nn_public_amd import meshgen
from dg_nn_public_amd.diffusion import DGDiffusion
from oracle import dg, mg

DEV = "cuda:0"


def build(D, n, sigma=0.0, seed=0, blk_solver="jacobian", face_kappa="mean"):
    m = meshgen.square_mesh(n, seed=seed) if D == 2 else meshgen.cube_mesh(n, seed=seed)
    pb = meshgen.sine_problem(m, sigma=sigma)
    g = DGDiffusion.from_mesh(m, pb.kappa, blk_solver=blk_solver, device=DEV,
                              face_kappa=face_kappa)
    sp = dg.make_space(m, pb.kappa, face_kappa=face_kappa)
    return m, pb, g, sp


def to_dev(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=DEV).contiguous()


def to_np(t):
    return t.detach().double().cpu().numpy()


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
