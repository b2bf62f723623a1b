"""SIPG diffusion operator on P3 DG, FP64 NumPy restatement (oracle).

Restates, for a scalar unknown with per-element diffusivity kappa:

* mesh topology ``nbf``/``alnmt``/``glb_bcface_type``  -- z02/mesh_init.py:293-390
  (3D) and :40-100 (2D), boundary typing as in z02/mesh_init.py:697-741;
* iso-parametric geometry at every quadrature point -- z02/shape_function.py:1321-1384
  (Jacobians), :1650-1725 (volume), :1730-1934 (faces, normal from the cross product
  of J . drst_duv);
* the residual ``get_residual_or_smooth`` (z02/volume_mf_diff.py:403-483) with the
  volume term ``_k_res_one_batch`` (:487-617), interior faces ``_s_res_fi_all_face``
  (:909-1369; this side :1016-1174 masked to interior faces, other side :1301-1369),
  Dirichlet faces ``_s_res_fb_all_face`` (:1460-1556); kappa enters as in
  z05/volume_mf_um_diff_varprop.py (:421-426 volume, :567 mean on interior faces,
  :682 owner on boundary faces);
* ``get_bdiag_diag`` (z02/get_bdiag_diag.py:30-62) and the block-Jacobi smoother
  (z02/volume_mf_diff.py:454-480 'jacobian'; z05/volume_mf_um_diff_varprop.py:384-392
  exact block inverse);
* the right-hand side ``_get_rhs`` (z02/volume_mf_diff.py:153-357, Dirichlet part
  :298-357 scaled by the owner's kappa as z05 :320-342 does);
* ``get_l2_error`` (z02/volume_mf_diff.py:2032-2056), including its definition that
  squares each basis contribution separately.

Arrays follow the reference: DOF vectors are flat (nele*nloc), element-major.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import tables

ETA_E = 36.0          # z02/config.py:245
JAC_WEI = 2.0 / 3.0   # z02/config.py:108


# --------------------------------------------------------------------------------------
# topology
# --------------------------------------------------------------------------------------
def topology(cells: np.ndarray, D: int):
    """nbf, alnmt, glb_bcface_type for a conforming simplex mesh.

    Faces are the vertex lists of z02/mesh_init.py:305-320 (tets) / :43-45
    (triangles).  Two faces match when their vertex sets agree; ``alnmt`` is 0 when
    both list the same first vertex, 1 when the neighbour starts at this side's second
    vertex, else 2 (z02/mesh_init.py:379-387).  Unmatched faces are boundary faces,
    all Dirichlet (type 0); interior faces are type -1.
    """
    nface = D + 1
    fv = tables.TET_FACE_VERTS if D == 3 else tables.TRI_FACE_VERTS
    nele = cells.shape[0]
    faces = cells[:, fv]  # (nele, nface, D) global vertex ids
    nbf = -np.ones(nele * nface, dtype=np.int64)
    alnmt = -np.ones(nele * nface, dtype=np.int64)
    table = {}
    flat = faces.reshape(-1, D)
    for gf in range(nele * nface):
        key = tuple(sorted(flat[gf].tolist()))
        if key in table:
            other = table.pop(key)
            nbf[gf], nbf[other] = other, gf
            a, b = flat[gf], flat[other]
            if a[0] == b[0]:
                al = 0
            elif a[1] == b[0]:
                al = 1
            else:
                al = 2
            alnmt[gf] = alnmt[other] = al
        else:
            table[key] = gf
    bctype = np.where(nbf >= 0, -1, 0).astype(np.int64)
    return nbf, alnmt, bctype


# --------------------------------------------------------------------------------------
# function space: tables + topology + geometry at quadrature points
# --------------------------------------------------------------------------------------
@dataclass
class Space:
    el: tables.Element
    nele: int
    x_ref_in: np.ndarray   # (nele, D, nloc)
    kappa: np.ndarray      # (nele,)
    nbf: np.ndarray
    alnmt: np.ndarray
    bctype: np.ndarray
    cells: np.ndarray
    # geometry (iso-parametric, per quadrature point)
    nx: np.ndarray = None        # (nele, ngi, D, nloc) physical gradients
    detwei: np.ndarray = None    # (nele, ngi)
    snx: np.ndarray = None       # (nele, nface, sngi, D, nloc)
    snormal: np.ndarray = None   # (nele, nface, sngi, D)
    sdetwei: np.ndarray = None   # (nele, nface, sngi)
    gamma: np.ndarray = None     # (nele, nface) eta/h
    eta_e: float = ETA_E
    face_kappa: str = "mean"

    @property
    def D(self):
        return self.el.D

    @property
    def nloc(self):
        return self.el.nloc

    @property
    def nface(self):
        return self.el.nface

    @property
    def ndof(self):
        return self.nele * self.el.nloc


def _inv_det(J):
    return np.linalg.inv(J), np.linalg.det(J)


def face_kappa_value(kE, kN, mode="mean"):
    """Interior-face diffusivity.  "mean": (kE + kN)/2, the reference's choice
    (z05/volume_mf_um_diff_varprop.py:567).  "harmonic": 2 kE kN/(kE + kN), which is the
    symmetric weighted interior penalty of Ern, Stephansen & Zunino (IMA J. Numer. Anal.
    29, 2009) -- its diffusivity-weighted average {kappa grad u}_w equals the harmonic
    mean times the plain average -- and keeps the form coercive at any contrast."""
    if mode == "mean":
        return 0.5 * (kE + kN)
    if mode == "harmonic":
        return 2.0 * kE * kN / (kE + kN)
    raise ValueError("face_kappa must be 'mean' or 'harmonic'")


def make_space(mesh, kappa=None, eta_e: float = ETA_E, face_kappa: str = "mean",
               bctype=None) -> Space:
    """``bctype`` (nele*nface) overrides the boundary faces' ``glb_bcface_type`` (0
    Dirichlet, 1 Neumann, other values: no term), as the physical tags of a gmsh mesh do
    (z02/mesh_init.py:697-741); interior faces stay -1."""
    D = mesh.D
    el = tables.make_element(D)
    nbf, alnmt, bct = topology(mesh.cells, D)
    if bctype is not None:
        bctype = np.asarray(bctype, dtype=np.int64).reshape(-1)
        if bctype.shape != bct.shape or np.any((bctype < 0) != (nbf >= 0)):
            raise ValueError("bctype must be -1 exactly on the interior faces")
        bct = bctype.copy()
    bctype = bct
    nele = mesh.nele
    x = mesh.x_ref_in
    if kappa is None:
        kappa = np.ones(nele)
    sp = Space(el=el, nele=nele, x_ref_in=x, kappa=np.asarray(kappa, dtype=np.float64),
               nbf=nbf, alnmt=alnmt, bctype=bctype, cells=mesh.cells, eta_e=eta_e,
               face_kappa=face_kappa)
    # volume: J[i][j] = sum_n nlx[i,n,g] x[j,n]  (z02/volume_mf_diff.py:536-545)
    Jv = np.einsum("ing,ejn->egij", el.nlx, x)
    invJ, det = _inv_det(Jv)
    # nx[i] = sum_j invJ[i][j] nlx[j]   (z02/volume_mf_diff.py:586-591)
    sp.nx = np.einsum("egij,jng->egin", invJ, el.nlx)
    sp.detwei = det * el.weight[None, :]  # detJ without abs (z02/volume_mf_diff.py:607-617)
    # faces
    Js = np.einsum("fing,ejn->efgij", el.snlx, x)  # (nele, nface, sngi, D, D)
    invJs, _ = _inv_det(Js)
    sp.snx = np.einsum("efgij,fjng->efgin", invJs, el.snlx)
    if D == 3:
        # tangents t_k[j] = sum_i J[i][j] drst_duv[f,i,k] (z02/volume_mf_diff.py:1045-1060)
        t = np.einsum("efgij,fik->efgkj", Js, el.drst_duv)
        cr = np.cross(t[..., 0, :], t[..., 1, :])
        sdet = np.linalg.norm(cr, axis=-1)
        sp.snormal = cr / sdet[..., None]
    else:
        # tangent t[j] = sum_i J[i][j] drst_duv[f,i]; n = (t1, -t0)/|t| (:1061-1098)
        t = np.einsum("efgij,fi->efgj", Js, el.drst_duv)
        sdet = np.linalg.norm(t, axis=-1)
        sp.snormal = np.stack([t[..., 1], -t[..., 0]], axis=-1) / sdet[..., None]
    sp.sdetwei = sdet * el.sweight[None, None, :]
    h = sp.sdetwei.sum(axis=-1)
    if D == 3:
        h = np.sqrt(h)
    sp.gamma = eta_e / h
    return sp


# --------------------------------------------------------------------------------------
# operator
# --------------------------------------------------------------------------------------
def _face_cache(sp: Space):
    """Geometry-only face arrays, computed once per space (the reference recomputes
    them per call unless ``is_store_jacobian``, z02/config.py:73):
    dnv[z, n, g]    = grad v_n . n at face point g of global face z;
    sn_nb[k, n, g]  = neighbour basis n at the neighbour point aligned with g;
    dnv_nb[k, n, g] = neighbour grad v_n . n_this at that point (interior faces k)."""
    if getattr(sp, "_fc", None) is not None:
        return sp._fc
    el = sp.el
    nele, nloc, nface = sp.nele, el.nloc, el.nface
    snx = sp.snx.reshape(nele * nface, el.sngi, el.D, nloc)
    nrm = sp.snormal.reshape(nele * nface, el.sngi, el.D)
    dnv = np.einsum("zgin,zgi->zng", snx, nrm)
    gf = np.arange(nele * nface)
    zi = gf[sp.bctype < 0]
    zb = gf[sp.bctype == 0]
    nb = sp.nbf[zi]
    aln = el.gi_align[sp.alnmt[zi]]
    rows = np.arange(zi.size)[:, None]
    sn_nb = np.ascontiguousarray(el.sn[nb % nface][rows, :, aln].transpose(0, 2, 1))
    snx_nb = snx[nb][rows, aln]                       # (ni, sngi, D, nloc)
    dnv_nb = np.einsum("zgin,zgi->zng", snx_nb, nrm[zi])
    E = gf // nface
    F = gf % nface
    kap_i = face_kappa_value(sp.kappa[E[zi]], sp.kappa[nb // nface], sp.face_kappa)
    sp._fc = dict(dnv=dnv, zi=zi, zb=zb, nb_e=nb // nface, sn_nb=sn_nb, dnv_nb=dnv_nb,
                  E=E, F=F, kap_i=kap_i,
                  sdw=sp.sdetwei.reshape(-1, el.sngi), gam=sp.gamma.reshape(-1))
    return sp._fc


def apply_A(sp: Space, x: np.ndarray) -> np.ndarray:
    """A x for the SIPG operator (the negative of the residual with b = 0).

    Volume: kappa_E sum_g detwei nx . nx x (z02/volume_mf_diff.py:577-617).
    Interior face (E, f) with neighbour N (z02/volume_mf_diff.py:1100-1174 this side,
    :1301-1369 other side), scaled by (kappa_E + kappa_N)/2:
      sum_g sdetwei [ v_n (-1/2 (du_E/dn + du_N/dn) + gamma (u_E - u_N))
                      + dv_n/dn (-1/2 (u_E - u_N)) ]
    Dirichlet face (z02/volume_mf_diff.py:1515-1548), scaled by kappa_E:
      sum_g sdetwei [ v_n (-du_E/dn + gamma u_E) - dv_n/dn u_E ]
    with n the outward normal of E and gamma = eta/h of E's face."""
    el = sp.el
    nele, nloc = sp.nele, el.nloc
    c = _face_cache(sp)
    X = x.reshape(nele, nloc)
    # volume
    nx = sp.nx.reshape(nele, -1, nloc)                          # (e, g*D, n)
    gradu = np.matmul(nx, X[:, :, None])                          # (e, g*D, 1)
    wgt = np.repeat(sp.detwei * sp.kappa[:, None], el.D, axis=1)  # (e, g*D)
    out = np.matmul(nx.transpose(0, 2, 1), gradu * wgt[:, :, None])[:, :, 0]
    sn = el.sn
    E, F, dnv, sdw, gam = c["E"], c["F"], c["dnv"], c["sdw"], c["gam"]
    zi, zb = c["zi"], c["zb"]
    # interior faces
    Xi = X[E[zi]][:, None, :]
    Xn = X[c["nb_e"]][:, None, :]
    sn_i = sn[F[zi]]
    u_own = np.matmul(Xi, sn_i)[:, 0]
    dn_own = np.matmul(Xi, dnv[zi])[:, 0]
    u_nb = np.matmul(Xn, c["sn_nb"])[:, 0]
    dn_nb = np.matmul(Xn, c["dnv_nb"])[:, 0]
    w = sdw[zi]
    jump = u_own - u_nb
    # halves written as divisions: the same FP64 values, and exact for Fraction arrays
    q_sn = w * (-(dn_own + dn_nb) / 2 + gam[zi][:, None] * jump) * c["kap_i"][:, None]
    q_dn = w * (-jump / 2) * c["kap_i"][:, None]
    contrib = np.matmul(sn_i, q_sn[:, :, None])[:, :, 0] + np.matmul(dnv[zi], q_dn[:, :, None])[:, :, 0]
    np.add.at(out, E[zi], contrib)
    # Dirichlet faces
    if zb.size:
        Xb = X[E[zb]][:, None, :]
        sn_b = sn[F[zb]]
        u_b = np.matmul(Xb, sn_b)[:, 0]
        dn_b = np.matmul(Xb, dnv[zb])[:, 0]
        kb = sp.kappa[E[zb]][:, None]
        w = sdw[zb]
        q_sn = w * (-dn_b + gam[zb][:, None] * u_b) * kb
        q_dn = w * (-u_b) * kb
        contrib = np.matmul(sn_b, q_sn[:, :, None])[:, :, 0] + np.matmul(dnv[zb], q_dn[:, :, None])[:, :, 0]
        np.add.at(out, E[zb], contrib)
    return out.reshape(-1)


def with_dtype(sp: Space, dtype=np.longdouble) -> Space:
    """A copy of ``sp`` whose operator arrays hold the same FP64 coefficients as ``dtype``
    (``np.longdouble``: the x87 extended format on x86-64, 64-bit significand; ``object``:
    exact ``fractions.Fraction`` values), so that ``apply_A`` evaluates the same operator
    with less (or no) rounding.  The FP64 evaluation of b - A x rounds at about
    eps |A||x|, which reaches 1e-10 ||b|| for strongly varying diffusivity on fine meshes;
    not a step of the reference, which evaluates in FP64 (z02/volume_mf_diff.py:403-483)."""
    import copy
    from fractions import Fraction

    def conv(a):
        a = np.asarray(a)
        if a.dtype != np.float64:
            return a
        if dtype is object:
            return np.array([Fraction(float(v)) for v in a.ravel()],
                            dtype=object).reshape(a.shape)
        return a.astype(dtype)

    fc = _face_cache(sp)
    out = copy.copy(sp)
    for k in ("kappa", "nx", "detwei", "snx", "snormal", "sdetwei", "gamma"):
        setattr(out, k, conv(getattr(sp, k)))
    out._fc = {k: conv(v) for k, v in fc.items()}
    out.el = copy.copy(sp.el)
    out.el.sn = conv(sp.el.sn)
    return out


def residual_extended(sp: Space, b, x, dtype=np.longdouble):
    """b - A x with the FP64 operator of ``sp`` evaluated in ``dtype`` (with_dtype); b and x
    are taken exactly.  Returns the residual in ``dtype``."""
    se = with_dtype(sp, dtype)
    if dtype is object:
        from fractions import Fraction
        xe = np.array([Fraction(float(v)) for v in x], dtype=object)
        be = np.array([Fraction(float(v)) for v in b], dtype=object)
    else:
        xe, be = np.asarray(x).astype(dtype), np.asarray(b).astype(dtype)
    return be - apply_A(se, xe)


def get_residual_or_smooth(sp: Space, r0, x_i, x_rhs, do_smooth=False, bdiag=None,
                           blk_solver="jacobian", omega=JAC_WEI):
    """r0 <- r0 + x_rhs - A x_i; if do_smooth, x_i += omega * c with c from the block
    smoother (z02/volume_mf_diff.py:403-483).  Mutates and returns (r0, x_i)."""
    if isinstance(x_rhs, (int, float)) and x_rhs == 0:
        pass
    else:
        r0 += x_rhs
    r0 -= apply_A(sp, x_i)
    if do_smooth:
        x_i += omega * block_smooth(sp, r0, bdiag, blk_solver, omega)
    return r0, x_i


def block_smooth(sp: Space, r, bdiag, blk_solver="jacobian", omega=JAC_WEI):
    """Correction c ~ bdiagK^-1 r.

    'jacobian': three damped point-Jacobi sweeps c += omega (r - B c)/diag(B) from c = 0
    (z02/volume_mf_diff.py:474-480).  'direct': exact block inverse
    (z05/volume_mf_um_diff_varprop.py:384-392)."""
    nele, nloc = sp.nele, sp.nloc
    R = r.reshape(nele, nloc)
    B = bdiag
    if blk_solver == "direct":
        return np.linalg.solve(B, R[..., None])[..., 0].reshape(-1)
    d = np.einsum("enn->en", B)
    c = np.zeros_like(R)
    for _ in range(3):
        c = c + omega * (R - np.einsum("enm,em->en", B, c)) / d
    return c.reshape(-1)


def get_bdiag_diag(sp: Space):
    """Diagonal blocks A_EE (nele, nloc, nloc) and their diagonals
    (z02/get_bdiag_diag.py:30-62): volume + own-side interior face + Dirichlet terms."""
    el = sp.el
    nele, nloc, nface = sp.nele, el.nloc, el.nface
    B = sp.kappa[:, None, None] * np.einsum("egin,egim,eg->enm", sp.nx, sp.nx, sp.detwei)
    sn = el.sn
    for f in range(nface):
        gfi = np.arange(nele) * nface + f
        snx = sp.snx[:, f]                       # (nele, sngi, D, nloc)
        nrm = sp.snormal[:, f]
        w = sp.sdetwei[:, f]
        dnv = np.einsum("egin,egi->eng", snx, nrm)
        ss = np.einsum("ng,mg,eg->enm", sn[f], sn[f], w)
        sd = np.einsum("ng,emg,eg->enm", sn[f], dnv, w)       # v_n (grad u_m . n)
        interior = sp.bctype[gfi] < 0
        dirich = sp.bctype[gfi] == 0
        nbe = np.where(interior, sp.nbf[gfi] // nface, 0)
        kap_i = face_kappa_value(sp.kappa, sp.kappa[nbe], sp.face_kappa)
        Ki = (-0.5 * sd - 0.5 * sd.transpose(0, 2, 1)
              + sp.gamma[:, f][:, None, None] * ss) * kap_i[:, None, None]
        Kb = (-sd - sd.transpose(0, 2, 1)
              + sp.gamma[:, f][:, None, None] * ss) * sp.kappa[:, None, None]
        B += np.where(interior[:, None, None], Ki, 0.0) + np.where(dirich[:, None, None], Kb, 0.0)
    diag = np.einsum("enn->en", B).copy()
    return diag, B


def get_rhs(sp: Space, f, u_bc, u_neumann=None, alpha_x_i=None, dt=None):
    """b = sum_e M_e f_e - int (grad v . n) u_D kappa + gamma int v u_D kappa on Dirichlet
    faces (z02/volume_mf_diff.py:195-239, 298-357; kappa as z05 :320-342)
    + int v g_N on Neumann faces (glb_bcface_type 1, :360-400: sn sn sdetwei, no kappa)
    + sum_e M_e alpha_e / dt when ``alpha_x_i`` is given (the transient term, :230-238)."""
    el = sp.el
    nele, nloc, nface = sp.nele, el.nloc, el.nface
    F = np.asarray(f).reshape(nele, nloc)
    U = np.asarray(u_bc).reshape(nele, nloc)
    b = np.einsum("mg,ng,eg,en->em", el.n, el.n, sp.detwei, F)
    if alpha_x_i is not None:
        Al = np.asarray(alpha_x_i).reshape(nele, nloc)
        b += np.einsum("mg,ng,eg,en->em", el.n, el.n, sp.detwei, Al) / dt
    if u_neumann is not None:
        G = np.asarray(u_neumann).reshape(nele, nloc)
        for fc in range(nface):
            e = np.nonzero(sp.bctype[np.arange(nele) * nface + fc] == 1)[0]
            if e.size:
                b[e] += np.einsum("mg,ng,eg,en->em", el.sn[fc], el.sn[fc],
                                  sp.sdetwei[e, fc], G[e])
    for fc in range(nface):
        gfi = np.arange(nele) * nface + fc
        dirich = sp.bctype[gfi] == 0
        if not dirich.any():
            continue
        e = np.nonzero(dirich)[0]
        snx = sp.snx[e, fc]
        nrm = sp.snormal[e, fc]
        w = sp.sdetwei[e, fc]
        ub = np.einsum("ng,en->eg", el.sn[fc], U[e])
        dnv = np.einsum("egin,egi->eng", snx, nrm)
        contrib = (-np.einsum("eng,eg,eg->en", dnv, ub, w)
                   + sp.gamma[e, fc][:, None] * np.einsum("ng,eg,eg->en", el.sn[fc], ub, w))
        b[e] += contrib * sp.kappa[e][:, None]
    return b.reshape(-1)


def get_l2_error(sp: Space, x_num, x_ana):
    """(l2, linf) exactly as z02/volume_mf_diff.py:2032-2056 defines them:
    l2^2 = sum_e sum_n sum_g (n[n,g] (x_num - x_ana)[e,n])^2 detwei[e,g]."""
    d = (np.asarray(x_num) - np.asarray(x_ana)).reshape(sp.nele, sp.nloc)
    l2 = np.sqrt(np.einsum("ng,en,eg->", sp.el.n ** 2, d ** 2, sp.detwei))
    linf = np.abs(d).max()
    return l2, linf


def assemble_dense(sp: Space) -> np.ndarray:
    """Dense A by probing columns with unit vectors, as the reference's own check does
    (z05/volume_mf_um_diff_varprop.py:78-99).  Tiny meshes only."""
    n = sp.ndof
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = apply_A(sp, e)
        e[j] = 0.0
    return A
