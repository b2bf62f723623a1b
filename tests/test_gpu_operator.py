"""HIP operator apply, RHS, block diagonal and smoother against the FP64 oracle."""
import numpy as np
import pytest
import torch

from gpu_common import build, rel, to_dev, to_np
from oracle import dg

pytestmark = pytest.mark.gpu

CASES = [(2, 6, 1.0), (3, 3, 1.0)]


@pytest.fixture(scope="module", params=CASES, ids=["2d", "3d"])
def prob(request):
    D, n, sig = request.param
    print(f"[case] D={D} n={n} sigma={sig}", flush=True)
    return build(D, n, sigma=sig)


def test_topology_matches_oracle(prob):
    m, pb, g, sp = prob
    nbf, aln, bct = g.topology()
    assert np.array_equal(nbf, sp.nbf) and np.array_equal(aln, sp.alnmt)
    assert np.array_equal(bct, sp.bctype)


def test_apply_f64_matches_oracle(prob):
    m, pb, g, sp = prob
    rng = np.random.default_rng(7)
    x = rng.standard_normal(sp.ndof)
    b = rng.standard_normal(sp.ndof)
    ref = b - dg.apply_A(sp, x)
    out = g.apply(to_dev(x), to_dev(b))
    err = rel(to_np(out), ref)
    print("f64 apply rel err", err)
    assert err < 1e-12
    out2 = g.apply(to_dev(x))  # b = None: r = -A x
    assert rel(to_np(out2), -dg.apply_A(sp, x)) < 1e-12


def test_apply_f32_within_tolerance(prob):
    m, pb, g, sp = prob
    x = np.random.default_rng(8).standard_normal(sp.ndof)
    ref = dg.apply_A(sp, x)
    out = -to_np(g.apply(to_dev(x, torch.float32)))
    err = rel(out, ref)
    print("f32 apply rel err", err)
    assert err < 1e-6


def test_get_residual_or_smooth_accumulates_like_the_reference(prob):
    m, pb, g, sp = prob
    rng = np.random.default_rng(9)
    r0_init, x, b = rng.standard_normal((3, sp.ndof))
    r0 = to_dev(r0_init)
    xi = to_dev(x)
    r0, xi = g.get_residual_or_smooth(r0, xi, to_dev(b), do_smooth=False)
    assert rel(to_np(r0), r0_init + b - dg.apply_A(sp, x)) < 1e-12
    r0 = to_dev(np.zeros(sp.ndof))
    r0, _ = g.get_residual_or_smooth(r0, to_dev(x), 0, do_smooth=False)
    assert rel(to_np(r0), -dg.apply_A(sp, x)) < 1e-12


def test_rhs_matches_oracle(prob):
    m, pb, g, sp = prob
    ref = dg.get_rhs(sp, pb.f, pb.u_bc)
    rng = np.random.default_rng(10)
    ub = rng.standard_normal(sp.ndof)  # non-zero Dirichlet data exercises the face terms
    ref2 = dg.get_rhs(sp, pb.f, ub)
    out = g._get_rhs(g.new_vector(), pb.f, pb.u_bc)
    out2 = g._get_rhs(g.new_vector(), pb.f, ub)
    assert rel(to_np(out), ref) < 1e-12 and rel(to_np(out2), ref2) < 1e-12


def test_bdiag_matches_oracle(prob):
    m, pb, g, sp = prob
    d_ref, B_ref = dg.get_bdiag_diag(sp)
    d, B = g.get_bdiag_diag()
    assert rel(to_np(B), B_ref) < 1e-12 and rel(to_np(d), d_ref) < 1e-12


@pytest.mark.parametrize("blk", ["jacobian", "direct"])
def test_smooth_step_matches_oracle(prob, blk):
    m, pb, g0, sp = prob
    from dg_nn_public_amd.diffusion import DGDiffusion
    g = DGDiffusion.from_mesh(m, pb.kappa, blk_solver=blk)
    _, B = dg.get_bdiag_diag(sp)
    rng = np.random.default_rng(11)
    x, b = rng.standard_normal((2, sp.ndof))
    r_ref = np.zeros(sp.ndof)
    x_ref = x.copy()
    dg.get_residual_or_smooth(sp, r_ref, x_ref, b, True, B, blk)
    r0 = g.new_vector()
    xi = to_dev(x)
    r0, xi = g.get_residual_or_smooth(r0, xi, to_dev(b), do_smooth=True)
    assert rel(to_np(r0), r_ref) < 1e-12
    assert rel(to_np(xi), x_ref) < 1e-11
    # FP32 smoothing step within single-precision tolerance
    r32 = g.new_vector(torch.float32)
    x32 = to_dev(x, torch.float32)
    g.get_residual_or_smooth(r32, x32, to_dev(b, torch.float32), do_smooth=True)
    assert rel(to_np(x32), x_ref) < 1e-5


def test_bad_arguments_raise(prob):
    m, pb, g, sp = prob
    with pytest.raises(ValueError):
        g.apply(torch.zeros(sp.ndof + 1, dtype=torch.float64, device="cuda:0"))
    with pytest.raises(ValueError):
        g.apply(torch.zeros(sp.ndof, dtype=torch.float16, device="cuda:0"))
    with pytest.raises(ValueError):
        g.apply(torch.zeros(sp.ndof, dtype=torch.float64))


@pytest.mark.parametrize("D,n", [(2, 6), (3, 3)])
def test_harmonic_face_kappa_apply_and_bdiag_match_oracle(D, n):
    m, pb, g, sp = build(D, n, sigma=2.0, face_kappa="harmonic")
    x = np.random.default_rng(13).standard_normal(sp.ndof)
    assert rel(-to_np(g.apply(to_dev(x))), dg.apply_A(sp, x)) < 1e-12
    d, B = g.get_bdiag_diag()
    assert rel(to_np(B), dg.get_bdiag_diag(sp)[1]) < 1e-12


def test_full_size_3d_apply_properties():
    """The 3D workload size (24^3 x 6 tets, 1.66M DOFs): FP64 apply against the oracle,
    exactness on linear fields off the boundary, FP32 against FP64."""
    m, pb, g, sp = build(3, 24, sigma=1.0)
    print("[case] full-size 3D n=24", flush=True)
    x = np.random.default_rng(12).standard_normal(sp.ndof)
    ref = dg.apply_A(sp, x)
    y64 = -to_np(g.apply(to_dev(x)))
    err64 = rel(y64, ref)
    y32 = -to_np(g.apply(to_dev(x, torch.float32)))
    err32 = rel(y32, ref)
    print("full-size rel err f64", err64, "f32", err32)
    assert err64 < 1e-12 and err32 < 1e-6
    from gpu_common import build as _b
    m1, pb1, g1, sp1 = _b(3, 24, sigma=0.0)
    c = np.array([0.3, -1.2, 0.7])
    u = (np.einsum("d,edn->en", c, m1.x_ref_in) + 0.5).reshape(-1)
    r = to_np(g1.apply(to_dev(u))).reshape(m1.nele, -1)
    hasb = (sp1.bctype.reshape(m1.nele, -1) == 0).any(1)
    assert np.abs(r[~hasb]).max() < 1e-10


# ---- matrix-core kernels (FP32 and bf16-input) ------------------------------------------
def test_f32_vector_and_mfma_kernels_agree_with_oracle(prob):
    m, pb, g, sp = prob
    rng = np.random.default_rng(21)
    x, b = rng.standard_normal((2, sp.ndof))
    ref = b - dg.apply_A(sp, x)
    errs = {}
    for k in ("vector", "mfma"):
        g.set_f32_kernel(k)
        errs[k] = rel(to_np(g.apply(to_dev(x, torch.float32), to_dev(b, torch.float32))), ref)
        pos = to_np(g.apply(to_dev(x, torch.float32)))
        assert rel(-pos, dg.apply_A(sp, x)) < 1e-6
    g.set_f32_kernel("mfma")
    print("f32 rel err", errs, flush=True)
    assert errs["vector"] < 1e-6 and errs["mfma"] < 1e-6


def test_bf16_input_apply_error_is_bf16_sized(prob):
    """bf16 inputs (8-bit significand) with fp32 accumulation: the relative error of A x
    for random x sits near the bf16 unit roundoff 2^-8, far above fp32's."""
    m, pb, g, sp = prob
    x = np.random.default_rng(22).standard_normal(sp.ndof)
    ref = dg.apply_A(sp, x)
    y = -to_np(g.apply(to_dev(x, torch.float32), bf16=True))
    err = rel(y, ref)
    print("bf16 apply rel err", err, flush=True)
    assert 1e-5 < err < 2e-2


@pytest.mark.parametrize("kernel", ["vector", "mfma", "mfma_bf16"])
def test_f32_smooth_kernels_match_oracle(prob, kernel):
    m, pb, g, sp = prob
    _, B = dg.get_bdiag_diag(sp)
    rng = np.random.default_rng(23)
    x, b = rng.standard_normal((2, sp.ndof))
    r_ref = np.zeros(sp.ndof)
    x_ref = x.copy()
    dg.get_residual_or_smooth(sp, r_ref, x_ref, b, True, B, "jacobian")
    g.set_f32_kernel(kernel)
    try:
        r32 = g.new_vector(torch.float32)
        x32 = to_dev(x, torch.float32)
        g.get_residual_or_smooth(r32, x32, to_dev(b, torch.float32), do_smooth=True)
        tol = 2e-2 if kernel == "mfma_bf16" else 1e-5
        ex, er = rel(to_np(x32), x_ref), rel(to_np(r32), r_ref)
        print(kernel, "smooth rel err x", ex, "r", er, flush=True)
        assert ex < tol and er < tol
    finally:
        g.set_f32_kernel("mfma")


def test_unknown_f32_kernel_raises(prob):
    m, pb, g, sp = prob
    with pytest.raises(ValueError):
        g.set_f32_kernel("tensor")
    with pytest.raises(ValueError):
        g.apply(to_dev(np.zeros(sp.ndof)), bf16=True)  # bf16 inputs need float32 vectors


@pytest.mark.parametrize("waves", [2, 4, 8])
def test_mfma_tile_shapes_match_oracle(prob, waves):
    m, pb, g, sp = prob
    x = np.random.default_rng(24).standard_normal(sp.ndof)
    ref = dg.apply_A(sp, x)
    g.lib.dgmi_set_param(g.ctx, b"mfma_waves", float(waves))
    g.lib.dgmi_set_param(g.ctx, b"bf16_waves", float(waves))
    try:
        e32 = rel(-to_np(g.apply(to_dev(x, torch.float32))), ref)
        e16 = rel(-to_np(g.apply(to_dev(x, torch.float32), bf16=True)), ref)
    finally:
        g.lib.dgmi_set_param(g.ctx, b"mfma_waves", 4.0)
        g.lib.dgmi_set_param(g.ctx, b"bf16_waves", 4.0)
    print("waves", waves, "f32", e32, "bf16", e16, flush=True)
    assert e32 < 1e-6 and e16 < 2e-2


def test_autotune_keeps_a_correct_kernel(prob):
    m, pb, g, sp = prob
    choice = g.autotune(reps=3)
    print("autotune", choice, flush=True)
    assert choice["f32_kernel"] in ("vector", "mfma") and len(choice["us"]) == 7
    assert all(v > 0 for v in choice["us"].values())
    x = np.random.default_rng(25).standard_normal(sp.ndof)
    assert rel(-to_np(g.apply(to_dev(x, torch.float32))), dg.apply_A(sp, x)) < 1e-6
    g.set_f32_kernel("mfma")


@pytest.mark.parametrize("kernel", ["vector", "mfma"])
def test_smooth_n_matches_repeated_oracle_steps(kernel):
    """n smoothing steps in one call (FP32 vector sweeps carry the traces from step to
    step) equal n steps of the oracle's get_residual_or_smooth with the exact block
    smoother: FP64 to 1e-11, FP32 to single precision."""
    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.diffusion import DGDiffusion
    from gpu_common import DEV
    m = meshgen.cube_mesh(4, seed=2)
    pb = meshgen.sine_problem(m, sigma=1.0)
    g = DGDiffusion.from_mesh(m, pb.kappa, blk_solver="direct", device=DEV,
                              face_kappa="harmonic")
    g.set_param("smoother_half", 0)  # FP32 arithmetic throughout (FP16 copy: test_gpu_mg)
    sp = dg.make_space(m, pb.kappa, face_kappa="harmonic")
    _, B = dg.get_bdiag_diag(sp)
    rng = np.random.default_rng(41)
    x0, b = rng.standard_normal((2, sp.ndof))
    x_ref = x0.copy()
    for _ in range(5):
        r_ref = np.zeros(sp.ndof)
        dg.get_residual_or_smooth(sp, r_ref, x_ref, b, True, B, "direct")
    x64 = g.smooth_n(to_dev(x0), to_dev(b), 5)
    assert rel(to_np(x64), x_ref) < 1e-11
    g.set_f32_kernel(kernel)
    x32 = g.smooth_n(to_dev(x0, torch.float32), to_dev(b, torch.float32), 5)
    print(kernel, "smooth_n fp32 rel err", rel(to_np(x32), x_ref), flush=True)
    assert rel(to_np(x32), x_ref) < 1e-5
    g.set_f32_kernel("mfma")
