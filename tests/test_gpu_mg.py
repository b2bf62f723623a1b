"""HIP multigrid (RAR assembled and by coloured probing, transfers, SFC V-cycle,
preconditioner) against the FP64 oracle."""
import numpy as np
import pytest
import torch

from gpu_common import DEV, build, rel, to_dev, to_np
from dg_nn_public_amd import meshgen
from dg_nn_public_amd.diffusion import DGDiffusion
from oracle import dg, mg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[(2, 8, 0.5), (3, 4, 1.0)], ids=["2d", "3d"])
def prob(request):
    D, n, sig = request.param
    print(f"[case] D={D} n={n} sigma={sig}", flush=True)
    m, pb, g, sp = build(D, n, sigma=sig)
    g._get_RAR_and_sfc_data_Um()
    mgd = mg.build_mg(sp, m.verts)
    return m, pb, g, sp, mgd


def test_rar_and_sfc_match_oracle(prob):
    m, pb, g, sp, mgd = prob
    fina, cola, rar, sfc = g.mg_export()
    assert np.array_equal(fina, mgd.fina) and np.array_equal(cola, mgd.cola)
    assert np.array_equal(sfc, mgd.levels.sfc)
    assert rel(rar, mgd.rar) < 1e-12
    assert g.nodes_per_level == mgd.levels.nodes_per_level


def _mixed_faces(m, seed):
    """glb_bcface_type with every boundary kind: Dirichlet (0), Neumann (1) and no term (2)
    drawn per boundary face; interior faces -1."""
    _, _, bct = dg.topology(m.cells, m.D)
    rng = np.random.default_rng(seed)
    b = bct >= 0
    bct[b] = rng.integers(0, 3, int(b.sum()))
    return bct


@pytest.mark.parametrize("D,n", [(2, 6), (3, 3)])
def test_rar_assembled_and_probed_match_oracle(D, n):
    """RAR from the continuous P1 element matrices (the default) and by the reference's
    coloured probing both equal the oracle's coloured probing (z02/volume_mf_diff.py:
    1696-1754), on a mesh with every face kind and a 1e4 diffusivity contrast."""
    m = meshgen.square_mesh(n, seed=7) if D == 2 else meshgen.cube_mesh(n, seed=7)
    kap = 10.0 ** np.random.default_rng(8).uniform(-2, 2, m.nele)
    bct = _mixed_faces(m, 9)
    print(f"[case] D={D} n={n} faces {np.bincount(bct + 1)}", flush=True)
    sp = dg.make_space(m, kap, bctype=bct)
    mgd = mg.build_mg(sp, m.verts)
    out = {}
    for probing in (0, 1):
        g = DGDiffusion.from_mesh(m, kap, bctype=bct, device=DEV)
        g.set_param("rar_probing", probing)
        g._get_RAR_and_sfc_data_Um()
        fina, cola, rar, sfc = g.mg_export()
        assert np.array_equal(fina, mgd.fina) and np.array_equal(cola, mgd.cola)
        err = rel(rar, mgd.rar)
        print(f"rar_probing={probing} ncolor={g.ncolor} rel err {err:.2e}")
        assert err < 1e-12
        assert g.ncolor == (int(mg.color2(mgd.fina, mgd.cola, m.nverts)[1]) if probing else 0)
        out[probing] = rar
        g.close()
    # the two methods agree with each other to rounding
    assert rel(out[0], out[1]) < 1e-12


def test_restrict_and_prolong_match_oracle(prob):
    m, pb, g, sp, mgd = prob
    P = sp.el.p1_prolong
    r = np.random.default_rng(1).standard_normal(sp.ndof)
    ref = (mgd.I_cf @ mg.restrict_p3_to_p1(P, r))[np.argsort(mgd.levels.sfc)]
    assert rel(to_np(g.restrict_to_p1cg(to_dev(r))), ref) < 1e-13
    e = np.random.default_rng(2).standard_normal(m.nverts)  # natural order
    x0 = np.random.default_rng(3).standard_normal(sp.ndof)
    ref = x0 + mg.prolong_p1_to_p3(P, mgd.I_fc @ e)
    x = to_dev(x0)
    g.prolong_add_from_p1cg(to_dev(e[np.argsort(mgd.levels.sfc)]), x)
    assert rel(to_np(x), ref) < 1e-13


def _exact(g):
    """level the FP32 cycles solve exactly (None: every level is swept)"""
    return None if g.coarse_exact_level < 0 else g.coarse_exact_level


def test_sfc_vcycle_matches_oracle(prob):
    m, pb, g, sp, mgd = prob
    r = np.random.default_rng(4).standard_normal(m.nverts)
    ref = mg.mg_on_P1CG(r, mgd.levels)
    out = to_np(g.mg_on_P1CG(to_dev(r)))
    print("vcycle rel err", rel(out, ref))
    assert rel(out, ref) < 1e-11
    # FP32: exact coarse solve at the first level of at most 128 rows, and without it
    ref_x = mg.mg_on_P1CG(r, mgd.levels, exact_level=_exact(g))
    out32 = to_np(g.mg_on_P1CG(to_dev(r, torch.float32)))
    print("fp32 vcycle exact level", g.coarse_exact_level, "rel err", rel(out32, ref_x))
    assert rel(out32, ref_x) < 1e-4
    g.lib.dgmi_set_param(g.ctx, b"coarse_exact", 0.0)
    out32 = to_np(g.mg_on_P1CG(to_dev(r, torch.float32)))
    g.lib.dgmi_set_param(g.ctx, b"coarse_exact", 1.0)
    assert rel(out32, ref) < 1e-4


def test_preconditioner_matches_oracle(prob):
    m, pb, g, sp, mgd = prob
    b = np.random.default_rng(5).standard_normal(sp.ndof)
    ref = mg.um_left_precond(sp, mgd, b)
    out = to_np(g._um_left_precond(g.new_vector(), to_dev(b)))
    print("precond rel err", rel(out, ref))
    assert rel(out, ref) < 1e-10
    ref_x = mg.um_left_precond(sp, mgd, b, exact_level=_exact(g))
    out32 = to_np(g._um_left_precond(g.new_vector(torch.float32), to_dev(b, torch.float32)))
    assert rel(out32, ref_x) < 1e-4


@pytest.mark.parametrize("D,n", [(2, 8), (3, 4)], ids=["2d", "3d"])
def test_direct_smoother_fp32_neighbour_form(D, n):
    """With the exact block smoother the FP32 sweeps inside the preconditioner use
    x = (1 - omega) x + S (b - A_off x).  It matches the FP64 oracle (which sweeps
    x += S (b - A x)) to single precision and agrees with the FP32 residual form."""
    print(f"[case] D={D} n={n} direct smoother", flush=True)
    m, pb, g, sp = build(D, n, sigma=1.0, blk_solver="direct", face_kappa="harmonic")
    g.set_param("smoother_half", 0)  # the FP32 smoother (its FP16 copy: test_fp16_smoother_copy)
    g._get_RAR_and_sfc_data_Um()
    mgd = mg.build_mg(sp, m.verts)
    b = np.random.default_rng(6).standard_normal(sp.ndof)
    ref = mg.um_left_precond(sp, mgd, b, blk_solver="direct")
    assert rel(to_np(g._um_left_precond(g.new_vector(), to_dev(b))), ref) < 1e-10
    ref = mg.um_left_precond(sp, mgd, b, blk_solver="direct", exact_level=_exact(g))
    # "vector" with the neighbour form carries the face traces from sweep to sweep
    for kern in ("mfma", "vector"):
        g.set_f32_kernel(kern)
        out = {}
        for offd in (1, 0):
            g.lib.dgmi_set_param(g.ctx, b"smooth_offdiag", float(offd))
            out[offd] = to_np(g._um_left_precond(g.new_vector(torch.float32),
                                                 to_dev(b, torch.float32)))
            print(kern, "offdiag", offd, "rel err", rel(out[offd], ref), flush=True)
            assert rel(out[offd], ref) < 1e-4
        assert rel(out[1], out[0]) < 1e-4
    g.lib.dgmi_set_param(g.ctx, b"smooth_offdiag", 1.0)
    g.set_f32_kernel("mfma")


def test_grid_level_kernels_match_oracle(monkeypatch):
    """Force the grid-sized SFC-level kernels (k_sweep / k_sweep_restrict) on a small mesh
    by lowering the single-work-group threshold, and compare V-cycle and preconditioner."""
    monkeypatch.setenv("DGMI_SMALL_LEVEL_N", "8")
    m, pb, g, sp = build(3, 4, sigma=1.0)
    g._get_RAR_and_sfc_data_Um()
    mgd = mg.build_mg(sp, m.verts)
    r = np.random.default_rng(6).standard_normal(m.nverts)
    ref = mg.mg_on_P1CG(r, mgd.levels)
    out = to_np(g.mg_on_P1CG(to_dev(r)))
    print("grid-level vcycle rel err", rel(out, ref))
    assert rel(out, ref) < 1e-11
    b = np.random.default_rng(7).standard_normal(sp.ndof)
    ref = mg.um_left_precond(sp, mgd, b)
    assert rel(to_np(g._um_left_precond(g.new_vector(), to_dev(b))), ref) < 1e-10
    # FP32 on the same grid levels, no exact level: inside the preconditioner the
    # restriction also takes the cycle's first level-0 sweep (FP32 tolerance 1e-4)
    g.lib.dgmi_set_param(g.ctx, b"coarse_exact", 0.0)
    try:
        out32 = to_np(g._um_left_precond(g.new_vector(torch.float32), to_dev(b, torch.float32)))
    finally:
        g.lib.dgmi_set_param(g.ctx, b"coarse_exact", 1.0)
    print("grid-level fp32 precond rel err", rel(out32, ref))
    assert rel(out32, ref) < 1e-4


def test_graph_vcycle_replays_equal_eager(prob):
    """The preconditioner's V-cycle is replayed from a captured graph: several replays, with
    eager work and synchronisations between them, equal the eagerly launched cycle."""
    m, pb, g, sp, mgd = prob
    rng = np.random.default_rng(31)
    for dt in (torch.float64, torch.float32):
        outs = {}
        for mode in (0.0, 1.0):
            g.lib.dgmi_set_param(g.ctx, b"graph_vcycle", mode)
            res = []
            for k in range(3):
                b = to_dev(rng.standard_normal(sp.ndof), dt) if mode == 0.0 else outs[0.0][k][0]
                x = torch.zeros(sp.ndof, dtype=dt, device="cuda:0")
                g._um_left_precond(x, b)
                torch.cuda.synchronize()
                junk = torch.randn(1000, device="cuda:0").sum()  # eager work between replays
                res.append((b, x.clone(), float(junk)))
            outs[mode] = res
        for (b0, x0, _), (b1, x1, _) in zip(outs[0.0], outs[1.0]):
            assert torch.equal(x0, x1), dt
    g.lib.dgmi_set_param(g.ctx, b"graph_vcycle", 0.0)


@pytest.mark.parametrize("packed", [False, True], ids=["single", "batch"])
def test_fused_first_sweep_equals_unfused(monkeypatch, packed):
    """With grid-sized SFC levels the eagerly launched V-cycle lets the restriction take the
    first level-0 sweep from zero; the graph-replayed cycle launches that sweep on its own.
    The two orders of the same arithmetic agree to rounding in FP64 and FP32, on one
    context and on a packed batch of three components (padded level-0 layout)."""
    monkeypatch.setenv("DGMI_SMALL_LEVEL_N", "8")
    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.batch import DGBatch
    ms = [meshgen.cube_mesh(n, seed=n) for n in ((4,) if not packed else (3, 4, 2))]
    ks = [meshgen.sine_problem(m, sigma=1.0, kappa_seed=5 + i).kappa for i, m in enumerate(ms)]
    print(f"[case] fused vs unfused first sweep, {len(ms)} component(s)", flush=True)
    bt = DGBatch(ms, ks, graph_vcycle=False)
    g = bt.space
    rng = np.random.default_rng(40)
    try:
        for dt, coarse_exact, tol in ((torch.float64, 1.0, 1e-14), (torch.float32, 0.0, 2e-6)):
            g.lib.dgmi_set_param(g.ctx, b"coarse_exact", coarse_exact)
            b = to_dev(rng.standard_normal(g.ndof), dt)
            out = {}
            for mode in (0.0, 1.0):
                g.lib.dgmi_set_param(g.ctx, b"graph_vcycle", mode)
                out[mode] = to_np(g._um_left_precond(g.new_vector(dt), b.clone()))
            err = rel(out[0.0], out[1.0])
            print(dt, "fused vs unfused rel diff", err, flush=True)
            assert np.all(np.isfinite(out[0.0])) and err < tol
    finally:
        g.lib.dgmi_set_param(g.ctx, b"coarse_exact", 1.0)
        bt.close()


def test_vector_sweeps_with_carried_traces_equal_matrix_core_sweeps():
    """On a mesh of many work-groups, the FP32 preconditioner whose vector sweeps carry the
    face traces (zero sweep, fused prolongation, neighbour-form sweeps) agrees with the
    matrix-core family, which recomputes every trace: a trace written from a stale or
    doubly-updated row would show as a large difference."""
    print("[case] D=3 n=12 direct smoother, vector vs mfma preconditioner", flush=True)
    m, pb, g, sp = build(3, 12, sigma=1.0, blk_solver="direct", face_kappa="harmonic")
    g.set_param("smoother_half", 0)  # both families with the FP32 smoother
    g._get_RAR_and_sfc_data_Um()
    b = to_dev(np.random.default_rng(8).standard_normal(sp.ndof), torch.float32)
    out = {}
    for kern in ("vector", "mfma"):
        g.set_f32_kernel(kern)
        out[kern] = to_np(g._um_left_precond(g.new_vector(torch.float32), b))
    g.set_f32_kernel("mfma")
    err = rel(out["vector"], out["mfma"])
    print("vector vs mfma preconditioner rel diff", err, flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("D,n", [(2, 24), (3, 8)], ids=["2d", "3d"])
def test_fp32_dense_coarse_level_matches_oracle(D, n):
    """A single solve whose first SFC level has more than 512 rows solves it with the
    device Gauss-Jordan inverse (several 64-wide pivot blocks and identity padding),
    applied by the grid kernel: the FP32 cycle matches the oracle's cycle with the same
    exact level.  Tolerance 1e-3: an FP32 inverse of a level of ~700 rows whose
    condition number is ~1e3."""
    print(f"[case] dense coarse level D={D} n={n}", flush=True)
    m, pb, g, sp = build(D, n, sigma=0.5)
    g._get_RAR_and_sfc_data_Um()
    mgd = mg.build_mg(sp, m.verts)
    lvl = g.coarse_exact_level
    assert lvl >= 0 and mgd.levels.nodes_per_level[lvl] > 512
    r = np.random.default_rng(21).standard_normal(m.nverts)
    ref = mg.mg_on_P1CG(r, mgd.levels, exact_level=lvl)
    out = to_np(g.mg_on_P1CG(to_dev(r, torch.float32)))
    print("dense level", lvl, "rows", mgd.levels.nodes_per_level[lvl], "rel err", rel(out, ref))
    assert rel(out, ref) < 1e-3


def test_preconditioner_allocation_failure_is_an_error():
    """A work vector that cannot be allocated (here: the context's byte cap) makes the call
    fail with an error before any launch; with the cap lifted it runs and matches."""
    from dg_nn_public_amd._lib import DgmiError
    m, pb, g, sp = build(3, 2, sigma=0.5)
    g._get_RAR_and_sfc_data_Um()
    g.set_alloc_limit(g.device_bytes())
    b = np.random.default_rng(5).standard_normal(sp.ndof)
    for dt in (torch.float64, torch.float32):
        with pytest.raises(DgmiError, match="allocation failed"):
            g._um_left_precond(g.new_vector(dt), to_dev(b, dt))
    g.set_alloc_limit(0)
    mgd = mg.build_mg(sp, m.verts)
    out = to_np(g._um_left_precond(g.new_vector(), to_dev(b)))
    assert rel(out, mg.um_left_precond(sp, mgd, b)) < 1e-10


@pytest.mark.parametrize("D,n", [(2, 10), (3, 5)], ids=["2d", "3d"])
def test_fp16_smoother_copy(D, n):
    """The FP32 block-staged sweeps can read an FP16 copy of the packed exact block
    smoother (each element's entries divided by its largest magnitude; the default in 3D,
    switched on here for 2D before the set-up).  Against the FP32 smoother: smoothing steps
    and the preconditioner differ by no more than the FP16 rounding allows (a few 1e-4
    relative; bound 5e-3), but they do differ (the copy is used), and an FP32-refinement
    solve with it reaches 1e-10 by the oracle's operator."""
    print(f"[case] fp16 smoother D={D} n={n}", flush=True)
    m, pb, g, sp = build(D, n, sigma=1.0, blk_solver="direct", face_kappa="harmonic")
    g.set_param("smoother_half", 1)
    g._get_RAR_and_sfc_data_Um()
    rng = np.random.default_rng(12)
    x0, b = rng.standard_normal((2, sp.ndof))
    out = {}
    for half in (1, 0):
        g.set_param("smoother_half", half)
        xs = g.smooth_n(to_dev(x0, torch.float32), to_dev(b, torch.float32), 3)
        z = g._um_left_precond(g.new_vector(torch.float32), to_dev(b, torch.float32))
        out[half] = (to_np(xs), to_np(z))
    d_sweep, d_pc = rel(out[1][0], out[0][0]), rel(out[1][1], out[0][1])
    print("fp16 vs fp32 smoother: sweeps", d_sweep, "preconditioner", d_pc, flush=True)
    assert 1e-7 < d_sweep < 5e-3 and 1e-7 < d_pc < 5e-3
    g.set_param("smoother_half", 1)
    bb = dg.get_rhs(sp, pb.f, pb.u_bc)
    xs, its = g._gmres_mg_solver(g.new_vector(), to_dev(bb), rtol=1e-10, inner=torch.float32)
    relres = np.linalg.norm(bb - dg.apply_A(sp, to_np(xs))) / np.linalg.norm(bb)
    print("solve", its, relres, flush=True)
    assert relres <= 2e-10
    g.close()
