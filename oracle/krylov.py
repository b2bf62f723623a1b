"""Krylov solvers with the SFC-multigrid preconditioner (oracle).

* ``gmres_mg_solver``: left-preconditioned restarted GMRES(m) with modified
  Gram-Schmidt and a QR least-squares solve, stopping on the true residual checked at
  the end of each restart cycle (z02/volume_mf_diff.py:1757-1841).  The reference
  compares the absolute residual with ``tol``; the workloads here use a relative
  tolerance, so ``tol`` is multiplied by ||b||.  Saad, Iterative Methods for Sparse
  Linear Systems, 2nd ed., Alg. 6.9 (left preconditioning) is the textbook form.
* ``cg_mg_solver``: MG-preconditioned CG (z02/volume_mf_diff.py:1875-1931).
* ``mg_solver``: stand-alone V-cycle iteration (z02/volume_mf_diff.py:1844-1873).
"""
from __future__ import annotations

import numpy as np

from . import dg, mg


def gmres_mg_solver(sp, mgd, x, b, rtol=1e-10, m=30, max_its=400, precond=None,
                    history=None):
    """Returns (x, its).  ``precond(v)`` defaults to one V-cycle."""
    if precond is None:
        precond = lambda v: mg.um_left_precond(sp, mgd, v)  # noqa: E731
    x = np.array(x, dtype=np.float64)
    bnorm = np.linalg.norm(b)
    tol = rtol * bnorm
    its = 0
    r = b - dg.apply_A(sp, x)
    rnorm = np.linalg.norm(r)
    while rnorm > tol and its < max_its:
        z = precond(r)
        beta = np.linalg.norm(z)
        V = np.zeros((m + 1, sp.ndof))
        H = np.zeros((m + 1, m))
        V[0] = z / beta
        for j in range(m):
            w = precond(dg.apply_A(sp, V[j]))
            for i in range(j + 1):
                H[i, j] = np.dot(w, V[i])
                w -= H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            V[j + 1] = w / H[j + 1, j]
            its += 1
        q, rr = np.linalg.qr(H, mode="complete")
        e1 = np.zeros(m + 1)
        e1[0] = beta
        y = np.linalg.solve(rr[:m, :m], q[:, :m].T @ e1)
        x += V[:m].T @ y
        r = b - dg.apply_A(sp, x)
        rnorm = np.linalg.norm(r)
        if history is not None:
            history.append((its, rnorm / bnorm))
    return x, its


def cg_mg_solver(sp, mgd, x, b, rtol=1e-10, max_its=400):
    x = np.array(x, dtype=np.float64)
    tol = rtol * np.linalg.norm(b)
    r = b - dg.apply_A(sp, x)
    z = mg.um_left_precond(sp, mgd, r)
    p = z.copy()
    its = 0
    rnorm = np.linalg.norm(r)
    while rnorm > tol and its < max_its:
        Ap = dg.apply_A(sp, p)
        rz = r @ z
        alpha = rz / (Ap @ p)
        x += alpha * p
        r -= alpha * Ap
        z = mg.um_left_precond(sp, mgd, r)
        rnorm = np.linalg.norm(r)
        beta = (r @ z) / rz
        p = z + beta * p
        its += 1
    return x, its


def mg_solver(sp, mgd, x, b, rtol=1e-10, max_its=400):
    x = np.array(x, dtype=np.float64)
    tol = rtol * np.linalg.norm(b)
    r = b - dg.apply_A(sp, x)
    rnorm = np.linalg.norm(r)
    its = 0
    while rnorm > tol and its < max_its:
        x += mg.um_left_precond(sp, mgd, r)
        r = b - dg.apply_A(sp, x)
        rnorm = np.linalg.norm(r)
        its += 1
    return x, its
