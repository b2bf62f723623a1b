"""Independent pins of the oracle pieces that the GPU tests compare against: the block
smoother (closed forms), the right-hand side's Dirichlet, Neumann and transient terms
(consistency of the SIPG form on polynomial fields), the DG<->CG transfer weights, the
composition of the V-cycle preconditioner, and convergence on the reference's own 2D
problem.  Each check is built from the mathematics, not from the oracle's own code path."""
import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from dg_nn_public_amd import meshgen
from oracle import dg, mg, tables

OMEGA = dg.JAC_WEI


def _mesh(D, n, jitter=None, seed=0):
    kw = {} if jitter is None else {"jitter": jitter}
    return meshgen.square_mesh(n, seed=seed, **kw) if D == 2 else meshgen.cube_mesh(n, seed=seed, **kw)


# ------------------------------------------------------------------------ block smoother
@pytest.fixture(scope="module", params=[2, 3], ids=["2d", "3d"])
def blocks(request):
    D = request.param
    m = _mesh(D, 3 if D == 2 else 2)
    kap = np.exp(np.random.default_rng(5).standard_normal(m.nele))
    sp = dg.make_space(m, kap)
    d, B = dg.get_bdiag_diag(sp)
    return sp, d, B


def test_direct_block_smoother_solves_each_block(blocks):
    """'direct' (z05/volume_mf_um_diff_varprop.py:384-392): B_e c_e = r_e on every block."""
    sp, d, B = blocks
    r = np.random.default_rng(1).standard_normal(sp.ndof)
    c = dg.block_smooth(sp, r, B, "direct").reshape(sp.nele, sp.nloc)
    R = r.reshape(sp.nele, sp.nloc)
    res = np.einsum("enm,em->en", B, c) - R
    assert np.linalg.norm(res) <= 1e-13 * np.linalg.norm(R) * np.abs(B).max() / np.abs(d).min()


def test_jacobi_block_smoother_is_the_three_term_matrix_polynomial(blocks):
    """'jacobian' (z02/volume_mf_diff.py:474-480): three damped point-Jacobi sweeps from
    c = 0 equal c = sum_{k=0..2} (I - w D^-1 B)^k w D^-1 r, block by block."""
    sp, d, B = blocks
    r = np.random.default_rng(2).standard_normal(sp.ndof)
    c = dg.block_smooth(sp, r, B, "jacobian").reshape(sp.nele, sp.nloc)
    R = r.reshape(sp.nele, sp.nloc)
    nl = sp.nloc
    for e in range(sp.nele):
        Dinv = np.diag(1.0 / np.diag(B[e]))
        G = np.eye(nl) - OMEGA * Dinv @ B[e]
        C = (np.eye(nl) + G + G @ G) @ (OMEGA * Dinv)
        assert np.allclose(c[e], C @ R[e], rtol=1e-12, atol=1e-14 * np.abs(c[e]).max())


def test_smoothing_step_is_x_plus_omega_c(blocks):
    """get_residual_or_smooth with do_smooth: r0 <- b - A x and x <- x + w c(r0)."""
    sp, d, B = blocks
    rng = np.random.default_rng(3)
    x = rng.standard_normal(sp.ndof)
    b = rng.standard_normal(sp.ndof)
    A = dg.assemble_dense(sp)
    r0 = np.zeros(sp.ndof)
    x1 = x.copy()
    dg.get_residual_or_smooth(sp, r0, x1, b, True, B, "direct")
    assert np.allclose(r0, b - A @ x, rtol=0, atol=1e-12 * np.abs(b - A @ x).max())
    Binv = np.zeros((sp.ndof, sp.ndof))
    nl = sp.nloc
    for e in range(sp.nele):
        Binv[e * nl:(e + 1) * nl, e * nl:(e + 1) * nl] = np.linalg.inv(B[e])
    assert np.allclose(x1, x + OMEGA * Binv @ (b - A @ x), rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------------ right-hand side
def _poly(D, x):
    """A quadratic field and its gradient at points x (..., D): u = 1 + 2x - y (+ 0.5 z)
    + x^2 + x y (+ y z), so that -div(kappa grad u) = -2 kappa."""
    if D == 2:
        X, Y = x[..., 0], x[..., 1]
        u = 1 + 2 * X - Y + X ** 2 + X * Y
        g = np.stack([2 + 2 * X + Y, -1 + X], axis=-1)
    else:
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        u = 1 + 2 * X - Y + 0.5 * Z + X ** 2 + X * Y + Y * Z
        g = np.stack([2 + 2 * X + Y, -1 + X + Z, 0.5 + Y], axis=-1)
    return u, g


def _top_face_bctype(m, sp, axis_val=1.0):
    """glb_bcface_type: 1 (Neumann) on boundary faces in the plane x_{D-1} = 1, else 0."""
    D = m.D
    fv = tables.TET_FACE_VERTS if D == 3 else tables.TRI_FACE_VERTS
    bct = sp.bctype.copy()
    for z in np.nonzero(bct >= 0)[0]:
        e, f = divmod(int(z), D + 1)
        pts = m.verts[m.cells[e, list(fv[f])]]
        if np.all(np.abs(pts[:, D - 1] - axis_val) < 1e-12):
            bct[z] = 1
    return bct


@pytest.mark.parametrize("D,n", [(2, 3), (3, 2)])
@pytest.mark.parametrize("neumann", [False, True], ids=["dirichlet", "neumann"])
def test_rhs_is_consistent_with_the_operator_on_quadratic_fields(D, n, neumann):
    """SIPG consistency: a quadratic u lies in the P3 space, so with f = -div(kappa grad u),
    u_D = u and (on Neumann faces) g_N = kappa grad u . n, A I_h u = b holds on every
    element, Dirichlet and Neumann elements included.  Pins the sign and kappa scaling of
    -int (grad v . n) u_D, the penalty term gamma int v u_D, the mass term and the Neumann
    term int v g_N (z02/volume_mf_diff.py:195-239, 298-400)."""
    m = _mesh(D, n)
    kap = 2.5 * np.ones(m.nele)
    sp0 = dg.make_space(m, kap)
    bct = _top_face_bctype(m, sp0) if neumann else None
    sp = dg.make_space(m, kap, bctype=bct)
    if neumann:
        assert np.any(sp.bctype == 1)
    x = np.moveaxis(m.x_ref_in, 1, 2)  # (nele, nloc, D)
    u, g = _poly(D, x)
    f = np.full(u.shape, -2.0 * 2.5)
    # Neumann data: kappa du/dn with the outward normal +e_{D-1} of the top faces
    gn = 2.5 * g[..., D - 1]
    b = dg.get_rhs(sp, f.reshape(-1), u.reshape(-1), u_neumann=gn.reshape(-1) if neumann else None)
    Au = dg.apply_A(sp, u.reshape(-1))
    assert np.abs(Au - b).max() <= 1e-11 * np.abs(b).max()
    if neumann:  # the Neumann term matters: without it the equations do not hold
        b0 = dg.get_rhs(sp, f.reshape(-1), u.reshape(-1))
        assert np.abs(Au - b0).max() > 1e-3 * np.abs(b).max()


def test_transient_term_is_the_mass_matrix_over_dt():
    """M alpha / dt (z02/volume_mf_diff.py:230-238): adding alpha with dt equals adding
    alpha / dt to the source, and the mass matrix integrates constants to the measure."""
    m = _mesh(3, 2)
    sp = dg.make_space(m)
    rng = np.random.default_rng(4)
    f, u, al = (rng.standard_normal(sp.ndof) for _ in range(3))
    dt = 0.05
    b1 = dg.get_rhs(sp, f, u, alpha_x_i=al, dt=dt)
    b2 = dg.get_rhs(sp, f + al / dt, u)
    assert np.allclose(b1, b2, rtol=1e-13, atol=1e-11)
    z = np.zeros(sp.ndof)
    assert abs(dg.get_rhs(sp, z, z, alpha_x_i=np.ones(sp.ndof), dt=dt).sum() - 1.0 / dt) < 1e-11


def test_reference_2d_problem_converges_at_order_4():
    """The reference's own 2D test problem (z02/bc_f.py:950-954): u = sin(pi x) sinh(pi y)
    / sinh(pi), harmonic (f = 0), non-zero Dirichlet data on the whole boundary."""
    errs = []
    for n in (4, 8):
        m = _mesh(2, n, jitter=0.0)
        sp = dg.make_space(m)
        X, Y = m.x_ref_in[:, 0, :], m.x_ref_in[:, 1, :]
        u = (np.sin(np.pi * X) * np.sinh(np.pi * Y) / np.sinh(np.pi)).reshape(-1)
        b = dg.get_rhs(sp, np.zeros(sp.ndof), u)
        uh = np.linalg.solve(dg.assemble_dense(sp), b)
        errs.append(dg.get_l2_error(sp, uh, u)[0])
    assert 13.0 < errs[0] / errs[1] < 19.0


# ------------------------------------------------------------------------ transfers
def test_transfer_weights_average_the_dg_copies():
    """I_cf[i, (e,k)] = I_fc[(e,k), i] = 1/mult(i) (z02/sparsity.py:125-165): I_cf averages
    the P1DG copies of a continuous field back to it; I_fc hands each copy value/mult."""
    m = _mesh(3, 3)
    I_cf, I_fc = mg.transfer_matrices(m.cells, m.nverts)
    v = np.random.default_rng(6).standard_normal(m.nverts)
    copies = v[m.cells].reshape(-1)
    mult = np.bincount(m.cells.reshape(-1), minlength=m.nverts)
    assert np.allclose(I_cf @ copies, v, rtol=1e-14, atol=1e-14)
    assert np.allclose(I_fc @ v, (v / mult)[m.cells].reshape(-1), rtol=1e-14, atol=1e-15)
    assert abs(I_cf - I_fc.T).max() == 0


@pytest.mark.parametrize("D", [2, 3])
def test_p1_to_p3_prolongation_interpolates_linear_fields(D):
    """P[n, k] = barycentric coordinate k of P3 node n (z02/mesh_init.py:224-235, 610-631):
    prolonging the vertex values of a linear field gives its values at the P3 nodes, and
    the restriction is the transpose."""
    m = _mesh(D, 2)
    c = np.array([0.4, -1.1, 0.9][:D])
    lin = lambda x: x @ c + 0.3  # noqa: E731
    P = tables.make_element(D).p1_prolong
    vert_vals = lin(m.verts)[m.cells].reshape(-1)
    node_vals = lin(np.moveaxis(m.x_ref_in, 1, 2)).reshape(-1)
    assert np.allclose(mg.prolong_p1_to_p3(P, vert_vals), node_vals, rtol=0, atol=1e-13)
    r = np.random.default_rng(7).standard_normal(m.nele * P.shape[0])
    y = np.random.default_rng(8).standard_normal(m.nele * P.shape[1])
    assert abs(mg.restrict_p3_to_p1(P, r) @ y - r @ mg.prolong_p1_to_p3(P, y)) < 1e-12


# ------------------------------------------------------------------------ preconditioner
@pytest.mark.parametrize("blk", ["jacobian", "direct"])
def test_preconditioner_equals_its_dense_composition(blk):
    """_um_left_precond (z02/volume_mf_diff.py:1934-2029) as dense matrices: three
    smoothing steps x += w C (b - A x) from x = 0, the residual, restriction R = I_cf P^T in
    SFC order, the V-cycle V (its own closed-form tests pin it), x += P I_fc V R r, three
    more smoothing steps (the reference's post loop also runs pre_smooth_its times)."""
    m = _mesh(2, 3)
    kap = np.exp(0.7 * np.random.default_rng(9).standard_normal(m.nele))
    sp = dg.make_space(m, kap)
    mgd = mg.build_mg(sp, m.verts)
    A = dg.assemble_dense(sp)
    n, nl = sp.ndof, sp.nloc
    C = np.zeros((n, n))
    for e in range(sp.nele):
        Be = mgd.bdiag[e]
        if blk == "direct":
            Ce = np.linalg.inv(Be)
        else:
            Dinv = np.diag(1.0 / np.diag(Be))
            G = np.eye(nl) - OMEGA * Dinv @ Be
            Ce = (np.eye(nl) + G + G @ G) @ (OMEGA * Dinv)
        C[e * nl:(e + 1) * nl, e * nl:(e + 1) * nl] = Ce
    P = sp.el.p1_prolong
    Pblk = sps.block_diag([sps.csr_matrix(P)] * sp.nele).toarray()   # (ndof, p1dg)
    ncg = m.nverts
    perm = np.zeros((ncg, ncg))                     # r_sfc = perm @ r_cg
    perm[mgd.levels.sfc - 1, np.arange(ncg)] = 1.0
    V = np.stack([mg.mg_on_P1CG(col, mgd.levels) for col in np.eye(ncg)], axis=1)
    R = perm @ mgd.I_cf.toarray() @ Pblk.T
    Pr = Pblk @ mgd.I_fc.toarray() @ perm.T
    S = np.eye(n) - OMEGA * C @ A                    # error propagation of one step
    # x after three steps from 0: (I - S^3) A^-1 b written without A^-1:
    pre = sum(np.linalg.matrix_power(S, k) for k in range(3)) @ (OMEGA * C)
    mid = pre + Pr @ V @ R @ (np.eye(n) - A @ pre)
    Minv = np.linalg.matrix_power(S, 3) @ mid + sum(np.linalg.matrix_power(S, k)
                                                    for k in range(3)) @ (OMEGA * C)
    b = np.random.default_rng(10).standard_normal(n)
    out = mg.um_left_precond(sp, mgd, b, blk_solver=blk)
    assert np.allclose(out, Minv @ b, rtol=1e-10, atol=1e-10 * np.abs(out).max())


# ------------------------------------------------------------------------ definiteness
def test_reference_mean_form_is_indefinite_at_sigma_one():
    """With lognormal kappa, sigma = 1 (the 3D workload's contrast) on 4^3 x 6 tets, the
    reference's arithmetic-mean interior-face diffusivity (z05/volume_mf_um_diff_varprop.py
    :567) already gives a negative eigenvalue; the harmonic mean keeps the operator SPD."""
    m = _mesh(3, 4)
    kap = np.exp(np.random.default_rng(1).standard_normal(m.nele))
    lo = {}
    for mode in ("mean", "harmonic"):
        sp = dg.make_space(m, kap, face_kappa=mode)
        op = spla.LinearOperator((sp.ndof, sp.ndof), matvec=lambda v, s=sp: dg.apply_A(s, v),
                                 dtype=np.float64)
        lo[mode] = spla.eigsh(op, k=1, which="SA", tol=1e-6, maxiter=20000,
                              v0=np.ones(sp.ndof))[0][0]
    assert lo["mean"] < 0 < lo["harmonic"]


# -------------------------------------------------------------- extended-precision residual
@pytest.mark.parametrize("D", [2, 3], ids=["2d", "3d"])
def test_extended_residual_equals_exact_rational_evaluation(D):
    """residual_extended evaluates the same FP64 operator with a 64-bit significand: on a
    residual in the cancellation regime (b = A x rounded, so b - A x is of the size of one
    rounding of A x) it agrees with the exact evaluation in rational arithmetic to 1e-18 of
    max |A x|, while the FP64 evaluation is off by about its own size."""
    assert np.finfo(np.longdouble).eps < 1e-18  # x87 extended format
    m = _mesh(D, 2, seed=3)
    p = meshgen.sine_problem(m, sigma=2.0, kappa_seed=4)
    sp = dg.make_space(m, p.kappa, face_kappa="harmonic")
    x = np.random.default_rng(1).standard_normal(sp.ndof)
    ax = dg.apply_A(dg.with_dtype(sp), x.astype(np.longdouble))
    b = ax.astype(np.float64)
    exact = np.array([float(v) for v in dg.residual_extended(sp, b, x, dtype=object)])
    ext = dg.residual_extended(sp, b, x).astype(np.float64)
    r64 = b - dg.apply_A(sp, x)
    scale = np.abs(ax).max()
    err_ext = float(np.abs(ext - exact).max() / scale)
    err_64 = float(np.abs(r64 - exact).max() / scale)
    print("exact |r|", np.abs(exact).max() / scale, "extended err", err_ext, "FP64 err", err_64)
    assert err_ext <= 1e-18 and err_64 >= 1e-17
    # the same operator: on an ordinary residual the three agree to FP64 rounding
    b2 = np.random.default_rng(2).standard_normal(sp.ndof)
    r_ext = dg.residual_extended(sp, b2, x).astype(np.float64)
    assert np.linalg.norm(r_ext - (b2 - dg.apply_A(sp, x))) <= 1e-14 * np.linalg.norm(r_ext)
