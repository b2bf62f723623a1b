"""Oracle SIPG operator pinned by invariants and brute force (no golden vectors exist in
the reference for these): symmetry of the probed matrix, diagonal blocks, exactness on
linear fields, O(h^4) convergence to the manufactured solution, the L2-error definition."""
import numpy as np
import pytest

from dg_nn_public_amd import meshgen
from oracle import dg


def _mesh(D, n, jitter=None):
    if D == 2:
        return meshgen.square_mesh(n, **({} if jitter is None else {"jitter": jitter}))
    return meshgen.cube_mesh(n, **({} if jitter is None else {"jitter": jitter}))


@pytest.fixture(scope="module", params=[(2, 3), (3, 2)], ids=["2d", "3d"])
def small(request):
    D, n = request.param
    m = _mesh(D, n)
    kap = np.exp(0.5 * np.random.default_rng(3).standard_normal(m.nele))
    sp = dg.make_space(m, kap)
    return m, sp, dg.assemble_dense(sp)


def test_probed_matrix_is_symmetric(small):
    _, _, A = small
    assert np.abs(A - A.T).max() <= 1e-13 * np.abs(A).max()


def test_spd_for_moderate_contrast(small):
    _, _, A = small
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


def test_reference_mean_loses_definiteness_harmonic_keeps_it():
    """At high contrast the reference's arithmetic-mean face diffusivity gives an
    indefinite operator; the harmonic mean (weighted interior penalty) stays SPD and
    equals the mean where kappa is continuous."""
    m = _mesh(3, 3)
    kap = np.exp(2.0 * np.random.default_rng(1).standard_normal(m.nele))
    A_mean = dg.assemble_dense(dg.make_space(m, kap, face_kappa="mean"))
    sp_h = dg.make_space(m, kap, face_kappa="harmonic")
    A_h = dg.assemble_dense(sp_h)
    assert np.linalg.eigvalsh(0.5 * (A_mean + A_mean.T)).min() < 0
    assert np.linalg.eigvalsh(0.5 * (A_h + A_h.T)).min() > 0
    assert np.abs(A_h - A_h.T).max() <= 1e-13 * np.abs(A_h).max()
    d, B = dg.get_bdiag_diag(sp_h)
    nl = sp_h.nloc
    assert np.abs(np.stack([A_h[e * nl:(e + 1) * nl, e * nl:(e + 1) * nl]
                            for e in range(m.nele)]) - B).max() <= 1e-12 * np.abs(B).max()
    one = np.ones(m.nele)
    assert np.allclose(dg.assemble_dense(dg.make_space(m, one, face_kappa="harmonic")),
                       dg.assemble_dense(dg.make_space(m, one)), rtol=0, atol=1e-12)


def test_bdiag_is_diagonal_block_of_probed_matrix(small):
    _, sp, A = small
    d, B = dg.get_bdiag_diag(sp)
    nl = sp.nloc
    blocks = np.stack([A[e * nl:(e + 1) * nl, e * nl:(e + 1) * nl] for e in range(sp.nele)])
    assert np.abs(blocks - B).max() <= 1e-13 * np.abs(B).max()
    assert np.array_equal(d, np.einsum("enn->en", B))


@pytest.mark.parametrize("D,n", [(2, 4), (3, 3)])
def test_linear_fields_have_zero_residual_off_the_boundary(D, n):
    """Consistency: for linear u and constant kappa, (A u)_e = 0 on every element without
    a Dirichlet face.  Fails for a wrong +-1/2 factor, normal sign or face alignment."""
    m = _mesh(D, n)
    sp = dg.make_space(m)
    c = np.array([0.3, -1.2, 0.7][:D])
    u = (np.einsum("d,edn->en", c, m.x_ref_in) + 0.5).reshape(-1)
    r = dg.apply_A(sp, u).reshape(m.nele, sp.nloc)
    hasb = (sp.bctype.reshape(m.nele, -1) == 0).any(1)
    assert (~hasb).sum() > 0
    assert np.abs(r[~hasb]).max() < 1e-12


def test_constant_field_only_excites_dirichlet_elements():
    m = _mesh(3, 2)
    sp = dg.make_space(m)
    r = dg.apply_A(sp, np.ones(sp.ndof)).reshape(m.nele, sp.nloc)
    hasb = (sp.bctype.reshape(m.nele, -1) == 0).any(1)
    assert np.abs(r[~hasb]).max() < 1e-12 and np.abs(r[hasb]).max() > 1e-3


def test_element_relabelling_invariance():
    """Permuting the element order permutes the result (face/alignment bookkeeping)."""
    m = _mesh(2, 3)
    kap = np.exp(np.random.default_rng(1).standard_normal(m.nele) * 0.3)
    sp = dg.make_space(m, kap)
    x = np.random.default_rng(2).standard_normal(sp.ndof)
    perm = np.random.default_rng(4).permutation(m.nele)
    m2 = meshgen.Mesh(D=2, x_ref_in=m.x_ref_in[perm], cells=m.cells[perm], verts=m.verts)
    sp2 = dg.make_space(m2, kap[perm])
    y = dg.apply_A(sp, x).reshape(m.nele, -1)
    y2 = dg.apply_A(sp2, x.reshape(m.nele, -1)[perm].reshape(-1)).reshape(m.nele, -1)
    assert np.abs(y[perm] - y2).max() < 1e-12 * np.abs(y).max()


def test_manufactured_solution_converges_at_order_4():
    errs = []
    for n in (4, 8):
        m = _mesh(2, n, jitter=0.0)
        pb = meshgen.sine_problem(m)
        sp = dg.make_space(m)
        b = dg.get_rhs(sp, pb.f, pb.u_bc)
        u = np.linalg.solve(dg.assemble_dense(sp), b)
        errs.append(dg.get_l2_error(sp, u, pb.u_exact)[0])
    assert 13.0 < errs[0] / errs[1] < 19.0


def test_l2_error_definition_on_one_element():
    """get_l2_error squares each basis contribution separately (reference definition):
    for a single DOF difference d on one element it equals |d| sqrt(sum_g n_k^2 detwei)."""
    m = _mesh(2, 1)
    sp = dg.make_space(m)
    d = np.zeros(sp.ndof)
    d[3] = 0.5
    l2, linf = dg.get_l2_error(sp, d, np.zeros(sp.ndof))
    expect = 0.5 * np.sqrt(np.sum(sp.el.n[3] ** 2 * sp.detwei[0]))
    assert abs(l2 - expect) < 1e-15 and linf == 0.5


def test_rhs_mass_term_integrates_constants():
    """b = M f with f = 1 and zero Dirichlet data sums to the domain measure."""
    for D, n in ((2, 4), (3, 2)):
        m = _mesh(D, n)
        sp = dg.make_space(m)
        b = dg.get_rhs(sp, np.ones(sp.ndof), np.zeros(sp.ndof))
        assert abs(b.sum() - 1.0) < 1e-13
