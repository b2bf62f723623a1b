"""The rest of the reference's hot-path interface on the GPU, against the FP64 oracle:
Neumann and transient right-hand-side terms, MG-preconditioned CG, the stand-alone SFC
level operations (mg_on_P1CG_prep, mg_smooth_one_level, mg_on_P1CG on caller-held
levels), the element P3DG <-> P1DG transfers, and the reference's own 3D test problem
(exp field, Dirichlet on five faces, Neumann on z = 1)."""
import numpy as np
import pytest
import torch

from dg_nn_public_amd import meshgen
from dg_nn_public_amd.diffusion import DGDiffusion
from dg_nn_public_amd._lib import check
from gpu_common import DEV, build, rel, to_dev, to_np
from oracle import dg, krylov, mg, tables

pytestmark = pytest.mark.gpu


def _top_neumann(m):
    """glb_bcface_type: -1 interior, 1 on boundary faces in the plane x_{D-1} = 1, else 0."""
    D = m.D
    nbf, _, bct = dg.topology(m.cells, D)
    fv = tables.TET_FACE_VERTS if D == 3 else tables.TRI_FACE_VERTS
    for z in np.nonzero(bct >= 0)[0]:
        e, f = divmod(int(z), D + 1)
        if np.all(np.abs(m.verts[m.cells[e, list(fv[f])], D - 1] - 1.0) < 1e-12):
            bct[z] = 1
    return bct


@pytest.mark.parametrize("D,n", [(2, 6), (3, 3)])
def test_rhs_with_neumann_and_transient_terms_matches_oracle(D, n):
    print(f"[case] rhs D={D} n={n}", flush=True)
    m = meshgen.square_mesh(n) if D == 2 else meshgen.cube_mesh(n)
    kap = np.exp(0.5 * np.random.default_rng(1).standard_normal(m.nele))
    bct = _top_neumann(m)
    sp = dg.make_space(m, kap, bctype=bct)
    dt = 0.05
    g = DGDiffusion.from_mesh(m, kap, bctype=bct, device=DEV, is_transient=True, dt=dt)
    rng = np.random.default_rng(2)
    f, ud, un, al = (rng.standard_normal(sp.ndof) for _ in range(4))
    ref = dg.get_rhs(sp, f, ud, u_neumann=un, alpha_x_i=al, dt=dt)
    x_rhs = to_dev(np.ones(sp.ndof))  # the reference accumulates into x_rhs
    out = to_np(g._get_rhs(x_rhs, f, [to_dev(ud), to_dev(un)], to_dev(al)))
    assert rel(out - 1.0, ref) < 1e-13
    with pytest.raises(ValueError):  # Neumann faces need their data
        g._get_rhs(g.new_vector(), f, ud)
    # a steady space ignores alpha_x_i, as the reference does when not isTransient
    g2 = DGDiffusion.from_mesh(m, kap, bctype=bct, device=DEV)
    out2 = to_np(g2._get_rhs(g2.new_vector(), f, [ud, un], al))
    assert rel(out2, dg.get_rhs(sp, f, ud, u_neumann=un)) < 1e-13


@pytest.fixture(scope="module", params=[(2, 6, 0.0), (3, 3, 0.5)], ids=["2d", "3d"])
def prob(request):
    D, n, sig = request.param
    print(f"[case] D={D} n={n} sigma={sig}", flush=True)
    m, pb, g, sp = build(D, n, sigma=sig)
    g._get_RAR_and_sfc_data_Um()
    mgd = mg.build_mg(sp, m.verts)
    return m, pb, g, sp, mgd


def test_cg_mg_solver_matches_oracle(prob):
    m, pb, g, sp, mgd = prob
    b = dg.get_rhs(sp, pb.f, pb.u_bc)
    x_ref, its_ref = krylov.cg_mg_solver(sp, mgd, np.zeros(sp.ndof), b, rtol=1e-10, max_its=200)
    x, its = g._cg_mg_solver(g.new_vector(), to_dev(b), rtol=1e-10, max_its=200)
    xs = to_np(x)
    relres = np.linalg.norm(b - dg.apply_A(sp, xs)) / np.linalg.norm(b)
    print("cg its", its, "oracle its", its_ref, g.last_stats, "relres", relres)
    assert abs(its - its_ref) <= 1
    assert relres <= 1e-9
    assert rel(xs, x_ref) < 1e-8
    assert abs(g.last_stats["relres"] - relres) <= 1e-3 * relres + 1e-15
    # the reference's absolute test
    x2, its2 = g._cg_mg_solver(g.new_vector(), to_dev(b), tol=1e-8 * np.linalg.norm(b))
    assert its2 <= its


def test_solve_for_diff_runs_each_linear_solver(prob):
    m, pb, g, sp, mgd = prob
    b = dg.get_rhs(sp, pb.f, pb.u_bc)
    for name in ("gmres-mg", "cg-mg", "mg"):
        g.linear_solver = name
        x = to_np(g.solve_for_diff(g.new_vector(), pb.f, pb.u_bc, rtol=1e-10, max_its=800))
        relres = np.linalg.norm(b - dg.apply_A(sp, x)) / np.linalg.norm(b)
        print(name, "relres", relres, g.last_stats)
        assert relres <= 2e-10
    g.linear_solver = "gmres-mg"


def test_level_operations_match_oracle(prob):
    m, pb, g, sp, mgd = prob
    fina, cola, rar, sfc = g.mg_export()
    sfc2, vs, nlevel, nodes = g.mg_on_P1CG_prep(torch.as_tensor(fina), torch.as_tensor(cola),
                                                torch.as_tensor(rar), g)
    lv = mg.mg_on_P1CG_prep(fina, cola, rar, sfc)
    assert np.array_equal(sfc2[:, 0], sfc) and nlevel == lv.nlevel
    assert nodes == lv.nodes_per_level == g.nodes_per_level
    for l in range(nlevel):
        a, diag, nl = vs[l]
        A = lv.A[l]
        assert np.array_equal(to_np(a.crow).astype(np.int64), A.indptr)
        assert np.array_equal(to_np(a.col).astype(np.int64), A.indices)
        assert rel(to_np(a.values), A.data) < 1e-13 and rel(to_np(diag)[0], lv.diag[l]) < 1e-13
    rng = np.random.default_rng(3)
    for l in (0, 1, nlevel // 2):
        e0 = rng.standard_normal(nodes[l])
        bl = rng.standard_normal(nodes[l])
        e_ref, rr_ref = mg.mg_smooth_one_level(lv, l, e0.copy(), bl)
        e = to_dev(e0)
        e_out, rr = g.mg_smooth_one_level(l, e, to_dev(bl), vs)
        assert e_out is e
        assert rel(to_np(e), e_ref) < 1e-13 and rel(to_np(rr), rr_ref) < 1e-13
        # the context's own levels through sfc_data_Um give the same sweep
        e2 = to_dev(e0)
        g.mg_smooth_one_level(l, e2, to_dev(bl), g.sfc_data_Um.variables_sfc)
        assert rel(to_np(e2), e_ref) < 1e-13
    r = rng.standard_normal(nodes[0])
    ref = mg.mg_on_P1CG(r, lv)
    out = to_np(g.mg_on_P1CG(to_dev(r), vs, nlevel, nodes, nodes[0]))
    fused = to_np(g.mg_on_P1CG(to_dev(r), g.sfc_data_Um.variables_sfc, nlevel, nodes, nodes[0]))
    print("composed cycle", rel(out, ref), "fused", rel(fused, ref))
    assert rel(out, ref) < 1e-12 and rel(fused, ref) < 1e-11
    out32 = to_np(g.mg_on_P1CG(to_dev(r, torch.float32), vs, nlevel, nodes, nodes[0]))
    assert rel(out32, ref) < 1e-4
    with pytest.raises(ValueError):  # a CPU vector never reaches a kernel
        g.mg_on_P1CG(torch.as_tensor(r), vs, nlevel, nodes, nodes[0])
    with pytest.raises(ValueError):
        g.prolong_add_from_p1cg(to_dev(r[:-1]), g.new_vector())
    with pytest.raises(ValueError):
        g.prolong_add_from_p1cg(to_dev(r, torch.float32), g.new_vector())


def test_element_transfers_match_oracle(prob):
    m, pb, g, sp, mgd = prob
    P = sp.el.p1_prolong
    rng = np.random.default_rng(4)
    r = rng.standard_normal(sp.ndof)
    y = rng.standard_normal(m.nele * (m.D + 1))
    for dt, tol in ((torch.float64, 1e-14), (torch.float32, 1e-6)):
        assert rel(to_np(g.vel_pndg_to_p1dg_restrictor(to_dev(r, dt))),
                   mg.restrict_p3_to_p1(P, r)) < tol
        assert rel(to_np(g.vel_p1dg_to_pndg_prolongator(to_dev(y, dt))),
                   mg.prolong_p1_to_p3(P, y)) < tol


def _exp_problem(m):
    """The reference's 3D diffusion test (z02/bc_f.py:898-923): u = exp(-x-y-z),
    f = -3 exp(-x-y-z) (kappa = 1), Dirichlet data u on the faces x = 0, 1, y = 0, 1,
    z = 0 and Neumann flux -exp(-x-y-1) on z = 1 (z21-cube-mesh/cube_ho_poi.geo)."""
    X = m.x_ref_in
    u = np.exp(-X[:, 0] - X[:, 1] - X[:, 2]).reshape(-1)
    f = -3.0 * u
    gn = -np.exp(-X[:, 0] - X[:, 1] - 1.0).reshape(-1)
    return u, f, gn


def test_reference_3d_problem_solves_and_converges():
    errs = []
    for n in (2, 4, 8):
        print(f"[case] exp problem n={n}", flush=True)
        m = meshgen.cube_mesh(n, jitter=0.0)
        bct = _top_neumann(m)
        u, f, gn = _exp_problem(m)
        g = DGDiffusion.from_mesh(m, None, bctype=bct, device=DEV)
        x = g.solve_for_diff(g.new_vector(), f, [u, gn], rtol=1e-12)
        if n == 2:  # the oracle's solve on the same mesh
            sp = dg.make_space(m, bctype=bct)
            b = dg.get_rhs(sp, f, u, u_neumann=gn)
            x_ref = np.linalg.solve(dg.assemble_dense(sp), b)
            assert rel(to_np(x), x_ref) < 1e-9
        errs.append(g.get_l2_error(x, to_dev(u))[0])
        g.close()
    print("L2 errors", errs)
    assert errs[1] / errs[2] > 13.0 and errs[0] / errs[1] > 10.0


def test_solve_from_a_p3_gmsh_file(tmp_path):
    """The reference's workflow: a P3 gmsh file with physical groups "diri" and "neu"
    (z21-cube-mesh/cube_ho_poi.geo) is read (tetra20 cells renumbered into the reference's
    local order, groups numbered 0 and 1), the exp problem is posed on it and solved on
    the GPU; the result equals the oracle's solve on the same mesh and is O(h^4)-close to
    the exact solution."""
    from dg_nn_public_amd import io
    m0 = meshgen.cube_mesh(3, jitter=0.1, seed=3)
    bct0 = _top_neumann(m0)
    p = tmp_path / "cube_ho_poi.msh"
    io.write_msh(p, m0, bctype=bct0, tag_of_bc={0: 1, 1: 2}, order=3,
                 physical_names={1: (2, "diri"), 2: (2, "neu"), 3: (3, "fluid")},
                 reverse=np.arange(m0.nele) % 3 == 0)
    m, bct = io.read_msh(p)
    assert (bct == 1).sum() == (bct0 == 1).sum() > 0
    u, f, gn = _exp_problem(m)
    g = DGDiffusion.from_mesh(m, None, bctype=bct, device=DEV, linear_solver="cg-mg")
    x = to_np(g.solve_for_diff(g.new_vector(), f, [u, gn], rtol=1e-12))
    sp = dg.make_space(m, bctype=bct)
    x_ref = np.linalg.solve(dg.assemble_dense(sp), dg.get_rhs(sp, f, u, u_neumann=gn))
    print("file solve vs oracle", rel(x, x_ref), "l2 err", dg.get_l2_error(sp, x, u)[0])
    assert rel(x, x_ref) < 1e-9
    assert dg.get_l2_error(sp, x, u)[0] < 1e-4


@pytest.mark.parametrize("D,n", [(2, 6), (3, 4)])
def test_double_double_residual_on_every_face_type(D, n):
    """dgmi_residual_dd evaluates the same operator as the FP64 apply on a mesh with
    interior, Dirichlet and Neumann faces (Neumann faces of either local index)."""
    print(f"[case] residual_dd D={D} n={n}", flush=True)
    m = meshgen.square_mesh(n) if D == 2 else meshgen.cube_mesh(n, jitter=0.0)
    kap = np.exp(np.random.default_rng(7).standard_normal(m.nele))
    g = DGDiffusion.from_mesh(m, kap, bctype=_top_neumann(m), device=DEV)
    rng = np.random.default_rng(8)
    x, b = to_dev(rng.standard_normal(g.ndof)), to_dev(rng.standard_normal(g.ndof))
    rdd, r64 = to_np(g.residual_dd(x, b)), to_np(g.apply(x, b))
    assert np.all(np.isfinite(rdd)) and rel(rdd, r64) < 1e-14
    g.close()


def _scrambled(m, seed):
    """The same mesh with every element's vertices listed in a random order (kept
    positively oriented), so that face pairs meet in many relative orientations."""
    rng = np.random.default_rng(seed)
    cells = np.stack([c[rng.permutation(m.D + 1)] for c in m.cells])
    cells = meshgen._orient(m.verts, cells)
    table = meshgen.TET_P3 if m.D == 3 else meshgen.TRI_P3
    return meshgen.Mesh(D=m.D, x_ref_in=meshgen._p3_nodes(m.verts, cells, table), cells=cells,
                        verts=m.verts)


def _frame_ranks(m):
    """Lexicographic ranks of the vertex relabellings that carry each interior face's
    neighbour into the face's frame (shared vertices keep this side's numbers)."""
    D = m.D
    nbf, _, _ = dg.topology(m.cells, D)
    fv = tables.TET_FACE_VERTS if D == 3 else tables.TRI_FACE_VERTS
    ranks = set()
    for z in np.nonzero(nbf >= 0)[0]:
        e, f = divmod(int(z), D + 1)
        nb = int(nbf[z]) // (D + 1)
        opp = [k for k in range(D + 1) if k not in fv[f]][0]
        own, other = list(m.cells[e]), list(m.cells[nb])
        sig = [other.index(own[k]) if k != opp else -1 for k in range(D + 1)]
        sig[opp] = (set(range(D + 1)) - set(sig)).pop()
        r = 0
        for i in range(D + 1):
            r = r * (D + 1 - i) + sum(sig[j] < sig[i] for j in range(i + 1, D + 1))
        ranks.add(r)
    return ranks


@pytest.mark.parametrize("D,n", [(2, 7), (3, 3)])
def test_single_pass_kernels_match_oracle_and_two_pass(D, n):
    """The one-kernel apply and FP32 sweeps form the neighbour's face traces from its row of
    x in the face's frame (its vertices relabelled so that the shared face is this side's
    face with this side's vertex order, whatever orientation the two elements meet in).  On
    a mesh whose elements list their vertices in random orders, with interior, Dirichlet
    and Neumann faces: the FP64 single-pass apply equals the oracle's operator to rounding,
    the FP32 one to FP32 rounding, both for r = b - A x and for A x; the FP32
    preconditioner with single-pass sweeps equals the one whose sweeps carry face traces;
    and an FP32-refinement solve with them reaches 1e-10 by the oracle's operator."""
    print(f"[case] single pass D={D} n={n}", flush=True)
    m = _scrambled(meshgen.square_mesh(n) if D == 2 else meshgen.cube_mesh(n), 11)
    # every relabelling two positively oriented elements can meet in (half of the (D+1)!)
    assert len(_frame_ranks(m)) == (3 if D == 2 else 12)
    kap = np.exp(np.random.default_rng(9).standard_normal(m.nele))
    bct = _top_neumann(m)
    g = DGDiffusion.from_mesh(m, kap, bctype=bct, device=DEV, face_kappa="harmonic",
                              blk_solver="direct")
    g.set_param("smoother_half", 0)  # every sweep form with the FP32 smoother
    sp = dg.make_space(m, kap, bctype=bct, face_kappa="harmonic")
    rng = np.random.default_rng(10)
    x, b = rng.standard_normal(sp.ndof), rng.standard_normal(sp.ndof)
    ax = dg.apply_A(sp, x)
    for mode in (0.0, 1.0):
        check(g.lib.dgmi_set_param(g.ctx, b"single_pass", mode), "dgmi_set_param")
        for dt, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
            r = to_np(g.apply(to_dev(x, dt), to_dev(b, dt)))
            y = -to_np(g.apply(to_dev(x, dt)))
            print(mode, dt, rel(r, b - ax), rel(y, ax), flush=True)
            assert rel(r, b - ax) < tol and rel(y, ax) < tol
    g._get_RAR_and_sfc_data_Um()
    z = {}
    for mode in (0.0, 2.0):
        check(g.lib.dgmi_set_param(g.ctx, b"single_pass", mode), "dgmi_set_param")
        z[mode] = to_np(g._um_left_precond(g.new_vector(torch.float32), to_dev(b, torch.float32)))
    # the single-pass sweeps in their k_residual_w form (rows read by every wave, each
    # smoother row read whole) against the block-staged kernel (the default)
    check(g.lib.dgmi_set_param(g.ctx, b"staged", 0.0), "dgmi_set_param")
    z["unstaged"] = to_np(g._um_left_precond(g.new_vector(torch.float32), to_dev(b, torch.float32)))
    check(g.lib.dgmi_set_param(g.ctx, b"staged", 1.0), "dgmi_set_param")
    print("preconditioner", rel(z[2.0], z[0.0]), rel(z[2.0], z["unstaged"]), flush=True)
    assert rel(z[2.0], z[0.0]) < 1e-5 and rel(z[2.0], z["unstaged"]) < 1e-5
    xs, its = g._gmres_mg_solver(g.new_vector(), to_dev(b), rtol=1e-10, inner=torch.float32)
    relres = np.linalg.norm(b - dg.apply_A(sp, to_np(xs))) / np.linalg.norm(b)
    print("solve", its, g.last_stats, relres, flush=True)
    assert relres <= 2e-10
    g.close()
