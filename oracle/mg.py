"""DG -> P1CG -> SFC multigrid, FP64 NumPy/SciPy restatement (oracle).

Restates:
* P1CG sparsity (all vertex pairs sharing an element) -- z02/sparsity.py:102-122 and
  getfin_p1cg (z00-fortran-src/get_neighbouring.f90:133-156);
* I_cf / I_fc with 1/multiplicity weights -- z02/sparsity.py:125-165;
* P3DG <-> P1DG element transfers -- z02/multigrid_linearelastic.py:286-297;
* distance-2 colouring and coloured probing of RAR -- z02/sparsity.py:235-306,
  z02/volume_mf_diff.py:1696-1754;
* SFC level operators A_l[I,J] = sum A_0[i,j] over SFC positions i>>l = I, j>>l = J --
  best_sfc_mapping_to_sfc_matrix_unstructured (z00-fortran-src/compact_fcode.f90:7-225),
  level count calculate_nlevel_sfc (:229-250), get_a_diaga
  (z02/multigrid_linearelastic.py:42-109);
* the V-cycle ``mg_on_P1CG`` (z02/multigrid_linearelastic.py:112-172) including its
  behaviour that the residual evaluation after the pre-smoothing sweeps applies a
  fourth sweep in place, and that the coarse initial guess is the pairwise-summed fine
  error;
* the preconditioner ``_um_left_precond`` (z02/volume_mf_diff.py:1934-2029).

The space-filling curve itself is not reproducible (the reference calls a Fortran graph
heuristic with annealing, combined_fortran_source.f90:10575); here it is a Hilbert
order of the P1 vertex coordinates, defined exactly in ``hilbert_order`` so that the
product's C++ implementation can be compared bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sps

from . import dg

JAC_WEI = dg.JAC_WEI


# --------------------------------------------------------------------------------------
# P1CG structures
# --------------------------------------------------------------------------------------
def p1cg_sparsity(cells: np.ndarray, nverts: int):
    """CSR (fina, cola) of the P1CG graph: (i, j) for every pair of vertices of an
    element, i == j included, columns sorted (z02/sparsity.py:102-122)."""
    nele, nv = cells.shape
    rows = np.repeat(cells, nv, axis=1).reshape(-1)
    cols = np.tile(cells, (1, nv)).reshape(-1)
    m = sps.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(nverts, nverts)).tocsr()
    m.sort_indices()
    return m.indptr.astype(np.int64), m.indices.astype(np.int64)


def transfer_matrices(cells: np.ndarray, nverts: int):
    """I_cf (cg x p1dg) and I_fc = I_cf^T with 1/multiplicity weights
    (z02/sparsity.py:125-165)."""
    p1dg = cells.size
    cols = np.arange(p1dg)
    rows = cells.reshape(-1)
    mult = np.bincount(rows, minlength=nverts).astype(np.float64)
    val = 1.0 / mult[rows]
    I_cf = sps.csr_matrix((val, (rows, cols)), shape=(nverts, p1dg))
    return I_cf, I_cf.T.tocsr()


def restrict_p3_to_p1(P: np.ndarray, r: np.ndarray) -> np.ndarray:
    """vel_pndg_to_p1dg_restrictor: y[e,k] = sum_n P[n,k] r[e,n]
    (z02/multigrid_linearelastic.py:286-290)."""
    nloc = P.shape[0]
    return np.einsum("nk,en->ek", P, r.reshape(-1, nloc)).reshape(-1)


def prolong_p1_to_p3(P: np.ndarray, e: np.ndarray) -> np.ndarray:
    """vel_p1dg_to_pndg_prolongator: y[e,n] = sum_k P[n,k] e[e,k]
    (z02/multigrid_linearelastic.py:293-297)."""
    nk = P.shape[1]
    return np.einsum("nk,ek->en", P, e.reshape(-1, nk)).reshape(-1)


def color2(fina, cola, n):
    """Greedy distance-2 colouring, colours 1..ncolor (z02/sparsity.py:235-269)."""
    whichc = np.zeros(n, dtype=np.int64)
    ncolor = 0
    need = np.ones(n, dtype=bool)
    while need.any():
        ncolor += 1
        for i in np.nonzero(need)[0]:
            nbrs = cola[fina[i]:fina[i + 1]]
            d2 = np.concatenate([cola[fina[j]:fina[j + 1]] for j in nbrs])
            if not np.any(whichc[d2] == ncolor):
                whichc[i] = ncolor
        need = whichc == 0
    return whichc, ncolor


def rar_colored(sp, fina, cola, I_cf, I_fc, whichc, ncolor):
    """RAR values on the P1CG pattern by coloured probing: for each colour c,
    RARm = I_cf R^T A R I_fc mask_c, scattered to the row entries whose column has
    colour c (z02/volume_mf_diff.py:1696-1754, z02/sparsity.py:272-306)."""
    P = sp.el.p1_prolong
    value = np.zeros(cola.size)
    rows = np.repeat(np.arange(fina.size - 1), np.diff(fina))
    for c in range(1, ncolor + 1):
        mask = (whichc == c).astype(np.float64)
        Rm = prolong_p1_to_p3(P, I_fc @ mask)
        ARm = dg.apply_A(sp, Rm)
        RARm = I_cf @ restrict_p3_to_p1(P, ARm)
        sel = whichc[cola] == c
        value[sel] += RARm[rows[sel]]
    return value


def rar_galerkin(sp, fina, cola, I_cf, I_fc):
    """RAR by an explicit sparse triple product (independent check of the probing)."""
    P = sp.el.p1_prolong
    nele, nloc = sp.nele, sp.nloc
    A = sps.csr_matrix(dg.assemble_dense(sp)) if sp.ndof <= 4000 else None
    if A is None:
        raise ValueError("rar_galerkin is for small meshes")
    Pblk = sps.block_diag([sps.csr_matrix(P)] * nele).tocsr()
    R = I_cf @ Pblk.T
    RAR = (R @ A @ R.T).tocsr()
    n = fina.size - 1
    rows = np.repeat(np.arange(n), np.diff(fina))
    return np.asarray(RAR[rows, cola]).reshape(-1)


# --------------------------------------------------------------------------------------
# space-filling curve
# --------------------------------------------------------------------------------------
HILBERT_BITS = {2: 16, 3: 10}


def quantize(verts: np.ndarray, bits: int) -> np.ndarray:
    """Integer grid coordinates: floor((x - lo)/(hi - lo) * (2^bits - 1) + 0.5) per axis,
    evaluated in that operation order in IEEE double."""
    lo = verts.min(axis=0)
    hi = verts.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    scale = float((1 << bits) - 1)
    t = (verts - lo) / span
    t = t * scale
    t = t + 0.5
    return np.floor(t).astype(np.int64)


def hilbert_index(q: np.ndarray, bits: int) -> np.ndarray:
    """Hilbert key of integer points q (npt, n) -- Skilling's transpose algorithm
    (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004),
    then bit interleaving with axis 0 most significant."""
    X = [q[:, i].copy() for i in range(q.shape[1])]
    n = len(X)
    M = 1 << (bits - 1)
    Q = M
    while Q > 1:
        Pm = Q - 1
        for i in range(n):
            hit = (X[i] & Q) != 0
            X[0] = np.where(hit, X[0] ^ Pm, X[0])
            t = np.where(hit, 0, (X[0] ^ X[i]) & Pm)
            X[0] = X[0] ^ t
            X[i] = X[i] ^ t
        Q >>= 1
    for i in range(1, n):
        X[i] = X[i] ^ X[i - 1]
    t = np.zeros_like(X[0])
    Q = M
    while Q > 1:
        t = np.where((X[n - 1] & Q) != 0, t ^ (Q - 1), t)
        Q >>= 1
    for i in range(n):
        X[i] = X[i] ^ t
    key = np.zeros_like(X[0])
    for j in range(bits - 1, -1, -1):
        for i in range(n):
            key = (key << 1) | ((X[i] >> j) & 1)
    return key


def hilbert_order(verts: np.ndarray):
    """SFC numbering: sfc[node] = 1-based position along the curve (the reference's
    ``space_filling_curve_numbering[:, 0]`` convention); ties broken by node id."""
    D = verts.shape[1]
    bits = HILBERT_BITS[D]
    key = hilbert_index(quantize(verts, bits), bits)
    order = np.lexsort((np.arange(key.size), key))  # node at each position
    sfc = np.empty(key.size, dtype=np.int64)
    sfc[order] = np.arange(1, key.size + 1)
    return sfc


def nlevel_sfc(n: int) -> int:
    """calculate_nlevel_sfc (z00-fortran-src/compact_fcode.f90:229-250)."""
    if n == 1:
        return 1
    lev = 1
    while n > 1:
        n = (n - 1) // 2 + 1
        lev += 1
    return lev


@dataclass
class SfcLevels:
    sfc: np.ndarray                 # 1-based SFC position per P1CG node
    nlevel: int
    nodes_per_level: list
    A: list = field(default_factory=list)     # scipy CSR per level (SFC order)
    diag: list = field(default_factory=list)


def mg_on_P1CG_prep(fina, cola, rar_values, sfc):
    """Level operators in SFC order (z02/multigrid_linearelastic.py:175-283 with
    z00-fortran-src/compact_fcode.f90:7-225)."""
    n = fina.size - 1
    rows = np.repeat(np.arange(n), np.diff(fina))
    pos_r = sfc[rows] - 1
    pos_c = sfc[cola] - 1
    nlev = nlevel_sfc(n)
    lv = SfcLevels(sfc=sfc, nlevel=nlev, nodes_per_level=[])
    for l in range(nlev):
        nl = ((n - 1) >> l) + 1
        Al = sps.coo_matrix((rar_values, (pos_r >> l, pos_c >> l)), shape=(nl, nl)).tocsr()
        Al.sum_duplicates()
        Al.sort_indices()
        lv.A.append(Al)
        lv.diag.append(Al.diagonal().copy())
        lv.nodes_per_level.append(nl)
    return lv


def _pairsum(v):
    n = v.size
    m = (n - 1) // 2 + 1
    out = np.zeros(m)
    out += v[0::2]
    out[: n // 2] += v[1::2]
    return out


def mg_smooth_one_level(lv: SfcLevels, level, e, b, omega=JAC_WEI):
    """rr = b - A_l e; e += omega rr / diag (z02/multigrid_linearelastic.py:25-39).
    Updates e in place and returns (e, rr)."""
    rr = b - lv.A[level] @ e
    e += rr / lv.diag[level] * omega
    return e, rr


def mg_on_P1CG(r0, lv: SfcLevels, pre=3, post=3, omega=JAC_WEI, exact_level=None):
    """One SFC V-cycle (z02/multigrid_linearelastic.py:112-172).  With exact_level = c the
    cycle stops at level c and solves it exactly (the FP32 cycles of the HIP path do that
    at one level, which the context reports); None is the reference's cycle."""
    L = lv.nlevel if exact_level is None else exact_level + 1
    r_s = [np.asarray(r0, dtype=np.float64).copy()]
    e_s = [np.zeros(lv.nodes_per_level[0])]
    for level in range(L - 1):
        for _ in range(pre):
            mg_smooth_one_level(lv, level, e_s[level], r_s[level], omega)
        # residual evaluation: the reference's call also updates e in place (4th sweep)
        _, rr = mg_smooth_one_level(lv, level, e_s[level], r_s[level], omega)
        r_s.append(_pairsum(rr))
        e_s.append(_pairsum(e_s[level]))
    for level in range(L - 1, -1, -1):
        if level == L - 1:
            A = lv.A[level].toarray()
            e_s[level] = np.linalg.solve(A, r_s[level])
        else:
            e_s[level] += np.repeat(e_s[level + 1], 2)[: lv.nodes_per_level[level]]
            for _ in range(post):
                mg_smooth_one_level(lv, level, e_s[level], r_s[level], omega)
    return e_s[0]


# --------------------------------------------------------------------------------------
# multigrid preconditioner data
# --------------------------------------------------------------------------------------
@dataclass
class MGData:
    I_cf: object
    I_fc: object
    fina: np.ndarray
    cola: np.ndarray
    rar: np.ndarray
    levels: SfcLevels
    bdiag: np.ndarray
    diag: np.ndarray


def build_mg(sp, verts, rar_method="colored", sfc=None):
    """Set-up as _get_RAR_and_sfc_data_Um (z02/volume_mf_diff.py:1642-1693) with the
    SFC multigrid option, plus get_bdiag_diag."""
    nverts = verts.shape[0]
    fina, cola = p1cg_sparsity(sp.cells, nverts)
    I_cf, I_fc = transfer_matrices(sp.cells, nverts)
    if rar_method == "colored":
        whichc, ncolor = color2(fina, cola, nverts)
        rar = rar_colored(sp, fina, cola, I_cf, I_fc, whichc, ncolor)
    else:
        rar = rar_galerkin(sp, fina, cola, I_cf, I_fc)
    if sfc is None:
        sfc = hilbert_order(verts)
    levels = mg_on_P1CG_prep(fina, cola, rar, sfc)
    diag, bdiag = dg.get_bdiag_diag(sp)
    return MGData(I_cf=I_cf, I_fc=I_fc, fina=fina, cola=cola, rar=rar, levels=levels,
                  bdiag=bdiag, diag=diag)


def um_left_precond(sp, mgd: MGData, x_rhs, pre=3, post=3, blk_solver="jacobian",
                    omega=JAC_WEI, exact_level=None):
    """One V-cycle from x = 0 (z02/volume_mf_diff.py:1934-2029).  The post-smoothing
    loop runs ``pre`` times, as the reference's loop bound does (:2021)."""
    P = sp.el.p1_prolong
    x = np.zeros(sp.ndof)
    for _ in range(pre):
        r = np.zeros(sp.ndof)
        dg.get_residual_or_smooth(sp, r, x, x_rhs, True, mgd.bdiag, blk_solver, omega)
    r = np.array(x_rhs, dtype=np.float64) - dg.apply_A(sp, x)
    r1 = mgd.I_cf @ restrict_p3_to_p1(P, r)
    sfc = mgd.levels.sfc
    inv = np.argsort(sfc)
    e_sfc = mg_on_P1CG(r1[inv], mgd.levels, omega=omega, exact_level=exact_level)
    e = e_sfc[sfc - 1]
    x += prolong_p1_to_p3(P, mgd.I_fc @ e)
    for _ in range(pre):
        r = np.zeros(sp.ndof)
        dg.get_residual_or_smooth(sp, r, x, x_rhs, True, mgd.bdiag, blk_solver, omega)
    return x
