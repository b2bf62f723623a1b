"""Oracle reference-element tables against the tables the reference itself holds
(tests/golden/reference_tables.json, read from the reference sources as text)."""
import json
import os

import numpy as np
import pytest

from oracle import tables

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tables.json")))


def _rows(a):
    return sorted(map(tuple, np.round(np.asarray(a), 14)))


def test_tet24_rule_matches_reference_points_and_weights():
    L, w = tables.tet_rule_24()
    Lr = np.array(GOLD["tet24_L"]).reshape(24, 3)
    wr = np.array(GOLD["tet24_weight_unscaled"]) / 6.0
    assert _rows(np.c_[L, w]) == _rows(np.c_[Lr, wr])


def test_tri12_rule_matches_reference_in_order():
    p, w = tables.tri_rule_12()
    assert np.array_equal(p[:, :2], np.array(GOLD["tri12_xy"]).reshape(12, 2))
    assert np.array_equal(w, np.array(GOLD["tri12_sweight_unscaled"]) * 0.5)


def test_line4_rule_matches_reference():
    p, w = tables.line_rule_4()
    a, b = GOLD["line4_a"], GOLD["line4_b"]
    assert np.allclose(sorted(p[:, 0]), sorted([0.5 - 0.5 * b, 0.5 - 0.5 * a, 0.5 + 0.5 * a, 0.5 + 0.5 * b]))
    assert np.allclose(sorted(w), sorted(np.array([GOLD["line4_w2"], GOLD["line4_w1"]] * 2) / 2))


@pytest.mark.parametrize("D", [2, 3])
def test_geometric_alignment_equals_reference_gi_align(D):
    ref = (np.array(GOLD["tri12_alignment_1based"]).reshape(3, 12) - 1 if D == 3
           else np.array(GOLD["line4_alignment"]))
    assert np.array_equal(tables.alignment_table(D), ref)


@pytest.mark.parametrize("D,key", [(3, "tet20_nodes"), (2, "tri10_nodes")])
def test_basis_is_kronecker_at_reference_nodes(D, key):
    el = tables.make_element(D)
    X = np.array(GOLD[key], dtype=float).reshape(-1, D)
    lam = np.c_[X, 1 - X.sum(1)]
    phi, _ = tables.lagrange(el.index, lam)
    assert np.abs(phi - np.eye(el.nloc)).max() < 1e-13


@pytest.mark.parametrize("D,key", [(3, "tet20_p1_prolong"), (2, "tri10_p1_prolong")])
def test_p1_prolongator_matches_reference(D, key):
    assert np.abs(tables.make_element(D).p1_prolong - np.array(GOLD[key], dtype=float)).max() < 1e-15


@pytest.mark.parametrize("D,key", [(3, "tet20_face_iloc"), (2, "tri10_face_iloc")])
def test_face_nodes_match_reference(D, key):
    el = tables.make_element(D)
    ref = np.array(GOLD[key]).reshape(D + 1, -1)
    assert all(set(a) == set(b) for a, b in zip(ref, el.face_nodes))


def test_drst_duv_and_constants_match_reference():
    assert np.array_equal(tables.TET_DRST_DUV, np.array(GOLD["tet_drst_duv"], dtype=float))
    assert np.array_equal(tables.TRI_DRST_DUV, np.array(GOLD["tri_drst_duv"], dtype=float))
    from oracle import dg
    assert dg.ETA_E == GOLD["eta_e"] and dg.JAC_WEI == GOLD["jac_wei"]
    assert GOLD["pre_smooth_its"] == 3 and GOLD["post_smooth_its"] == 3


@pytest.mark.parametrize("D", [2, 3])
def test_quadrature_exact_for_degree_6(D):
    """Closed-form simplex monomial integrals: int l1^a l2^b (l3^c) = a! b! c! D!/(a+b+c+D)! |T|."""
    from math import factorial
    if D == 3:
        L, w = tables.tet_rule_24()
        for a, b, c in [(6, 0, 0), (2, 2, 2), (3, 1, 2), (0, 4, 1), (1, 1, 1)]:
            exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
            assert abs(np.sum(w * L[:, 0] ** a * L[:, 1] ** b * L[:, 2] ** c) - exact) < 1e-14
    else:
        p, w = tables.tri_rule_12()
        for a, b in [(6, 0), (3, 3), (2, 4), (1, 2)]:
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert abs(np.sum(w * p[:, 0] ** a * p[:, 1] ** b) - exact) < 1e-13


@pytest.mark.parametrize("D", [2, 3])
def test_basis_derivatives_match_finite_differences(D):
    el = tables.make_element(D)
    rng = np.random.default_rng(0)
    xi = rng.uniform(0.05, 0.3, size=(5, D))
    h = 1e-6
    for k in range(D):
        lp = np.c_[xi, 1 - xi.sum(1)].copy()
        lm = lp.copy()
        lp[:, k] += h; lp[:, D] -= h
        lm[:, k] -= h; lm[:, D] += h
        fd = (tables.lagrange(el.index, lp)[0] - tables.lagrange(el.index, lm)[0]) / (2 * h)
        _, dxi = tables.lagrange(el.index, np.c_[xi, 1 - xi.sum(1)])
        assert np.abs(fd - dxi[k]).max() < 1e-7
