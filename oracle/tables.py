"""Reference-element tables for P3 triangles and tetrahedra (oracle).

Restates ``SHATRInew`` (z02/shape_function.py:459-1317) and the quadrature rules
``_gi_pnts_tetra`` / ``_gi_pnts_tri`` / ``_gi_pnts_line`` (z02/shape_function.py:19-455)
for the configuration the hot path uses: element order p = 3, quadrature degree
2p = 6 (z02/main.py:52), i.e. ngi/sngi = 24/12 in 3D and 12/4 in 2D
(z02/function_space.py:32-66).

The basis is written generically from the barycentric multi-index of every node,
phi_a(l) = prod_i prod_{j<a_i} (p l_i - j)/(j+1), which is the equispaced simplex
Lagrange basis; the node order is the reference's local order
(z02/shape_function.py:511-532 for tets, :1127-1140 for triangles).  Local
coordinates are xi_k = l_k for k < D and l_D = 1 - sum(xi), so
d/dxi_k = d/dl_k - d/dl_D.

Quadrature point data are the published rules the reference tabulates (Keast #7
for the 24-point tet rule, Strang-Fix #9 for the 12-point triangle rule, 4-point
Gauss-Legendre on [0, 1]); they are reproduced with the reference's digits so that
the oracle integrates with the same rule.  Face-point alignment ``gi_align`` is
computed here by matching face points geometrically, and checked against the
reference's own table in the tests.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np

P = 3  # element order (z02/config.py:65)

# --------------------------------------------------------------------------------------
# node multi-indices (barycentric, sum = P) in the reference's local order
# --------------------------------------------------------------------------------------
# 3D: vertices 0..3, edge pairs (4,5 on 0-2) (6,7 on 0-1) (8,9 on 1-2) (10,11 on 0-3)
# (12,13 on 1-3) (14,15 on 2-3), face centres 16 {1,2,3} 17 {0,1,2} 18 {0,2,3} 19 {0,1,3}
# (z02/shape_function.py:511-532; z02/mesh_init.py:610-631 for the same positions).
TET_P3_INDEX = np.array([
    (3, 0, 0, 0), (0, 3, 0, 0), (0, 0, 3, 0), (0, 0, 0, 3),
    (2, 0, 1, 0), (1, 0, 2, 0), (2, 1, 0, 0), (1, 2, 0, 0),
    (0, 2, 1, 0), (0, 1, 2, 0), (2, 0, 0, 1), (1, 0, 0, 2),
    (0, 2, 0, 1), (0, 1, 0, 2), (0, 0, 2, 1), (0, 0, 1, 2),
    (0, 1, 1, 1), (1, 1, 1, 0), (1, 0, 1, 1), (1, 1, 0, 1),
], dtype=np.int64)

# 2D: vertices 0..2 (CCW), 3,4 on edge 0-1, 5,6 on 1-2, 7,8 on 2-0, 9 centroid
# (z02/shape_function.py:1127-1140; z02/mesh_init.py:224-235).
TRI_P3_INDEX = np.array([
    (3, 0, 0), (0, 3, 0), (0, 0, 3),
    (2, 1, 0), (1, 2, 0), (0, 2, 1), (0, 1, 2),
    (1, 0, 2), (2, 0, 1), (1, 1, 1),
], dtype=np.int64)


# --------------------------------------------------------------------------------------
# quadrature rules
# --------------------------------------------------------------------------------------
def _orbit4(vals):
    """Distinct permutations of a barycentric 4-tuple, in lexicographic order."""
    seen = []
    for perm in itertools.permutations(vals):
        if perm not in seen:
            seen.append(perm)
    return seen


def tet_rule_24():
    """Keast #7, 24 points, degree 6 (z02/shape_function.py:25-64).

    Returns (L, w): L (24, 3) local coordinates (l1, l2, l3), w (24,) summing to 1/6.
    Points are generated from the rule's four symmetry orbits; the coordinates are
    the reference's tabulated digits.
    """
    a1, b1 = 0.3561913862225449, 0.2146028712591517
    a2, b2 = 0.8779781243961660, 0.0406739585346113
    a3, b3 = 0.0329863295731731, 0.3223378901422757
    a4, b4, c4 = 0.2696723314583159, 0.6030056647916491, 0.0636610018750175
    orbits = [
        ((a1, b1, b1, b1), 0.0399227502581679),
        ((a2, b2, b2, b2), 0.0100772110553207),
        ((a3, b3, b3, b3), 0.0553571815436544),
        ((a4, b4, c4, c4), 0.0482142857142857),
    ]
    pts, wts = [], []
    for gen, w in orbits:
        for perm in _orbit4(gen):
            pts.append(perm[:3])
            wts.append(w)
    return np.array(pts, dtype=np.float64), np.array(wts, dtype=np.float64) / 6.0


def tri_rule_12():
    """Strang-Fix #9, 12 points, degree 6 (z02/shape_function.py:271-308).

    Returns (pnts, w): pnts (12, 3) barycentric (x, y, 1-x-y) in the reference's
    point order, w (12,) summing to 1/2.
    """
    a, b = 0.873821971016996, 0.063089014491502
    c, d = 0.501426509658179, 0.249286745170910
    e, f, g = 0.636502499121399, 0.310352451033785, 0.053145049844816
    xy = []
    for big, small in ((a, b), (c, d)):
        xy += [(big, small), (small, big), (small, small)]
    for perm in itertools.permutations((e, f, g)):
        xy.append(perm[:2])
    xy = np.array(xy, dtype=np.float64)
    pnts = np.concatenate([xy, (1.0 - xy.sum(axis=1))[:, None]], axis=1)
    w = np.array([0.050844906370207] * 3 + [0.116786275726379] * 3
                 + [0.082851075618374] * 6, dtype=np.float64) * 0.5
    return pnts, w


def line_rule_4():
    """4-point Gauss-Legendre on [0, 1] (z02/shape_function.py:413-438).

    Returns (pnts, w): pnts (4, 2) as (s, 1-s), w (4,) summing to 1.
    """
    a, b = 0.339981043584856, 0.861136311594053
    w1, w2 = 0.652145154862546, 0.347854845137454
    s = np.array([0.5 - 0.5 * b, 0.5 - 0.5 * a, 0.5 + 0.5 * a, 0.5 + 0.5 * b])
    pnts = np.stack([s, 1.0 - s], axis=1)
    w = np.array([w2, w1, w1, w2]) / 2.0
    return pnts, w


# --------------------------------------------------------------------------------------
# Lagrange basis on barycentric coordinates
# --------------------------------------------------------------------------------------
def _factor(k, t):
    """f_k(t) = prod_{j<k} (P t - j)/(j+1) and its derivative, vectorised over t."""
    val = np.ones_like(t)
    der = np.zeros_like(t)
    for j in range(k):
        term = (P * t - j) / (j + 1)
        der = der * term + val * (P / (j + 1))
        val = val * term
    return val, der


def lagrange(index, lam):
    """Evaluate the basis of multi-index table ``index`` (nloc, D+1) at barycentric
    points ``lam`` (npt, D+1).

    Returns phi (nloc, npt) and dphi/dxi (D, nloc, npt)."""
    nloc, nb = index.shape
    D = nb - 1
    npt = lam.shape[0]
    phi = np.ones((nloc, npt))
    dlam = np.zeros((nb, nloc, npt))
    for n in range(nloc):
        vals, ders = [], []
        for i in range(nb):
            v, d = _factor(int(index[n, i]), lam[:, i])
            vals.append(v)
            ders.append(d)
        for i in range(nb):
            phi[n] *= vals[i]
            prod = ders[i].copy()
            for k in range(nb):
                if k != i:
                    prod = prod * vals[k]
            dlam[i, n] = prod
    dxi = dlam[:D] - dlam[D][None]
    return phi, dxi


# --------------------------------------------------------------------------------------
# face parameterisations
# --------------------------------------------------------------------------------------
def tet_face_points(pnts):
    """Map triangle-rule points (sngi, 3) to barycentric points on each tet face.

    Face 0 is l1 = 0 (nodes 2,1,3), face 1 is l2 = 0 (nodes 0,2,3), face 2 is l3 = 0
    (nodes 1,0,3), face 3 is l4 = 0 (nodes 0,1,2) -- z02/shape_function.py:763-783.
    Returns (nface, sngi, 4)."""
    p0, p1, p2 = pnts[:, 0], pnts[:, 1], pnts[:, 2]
    z = np.zeros_like(p0)
    return np.stack([
        np.stack([z, p1, p0, p2], axis=1),
        np.stack([p0, z, p1, p2], axis=1),
        np.stack([p1, p0, z, p2], axis=1),
        np.stack([p0, p1, p2, z], axis=1),
    ])


def tri_face_points(pnts):
    """Map line-rule points (sngi, 2) to barycentric points on each triangle edge.

    Face 0 is edge 0-1, face 1 is edge 1-2, face 2 is edge 2-0
    (z02/shape_function.py:1105-1116).  Returns (nface, sngi, 3)."""
    s0, s1 = pnts[:, 0], pnts[:, 1]
    z = np.zeros_like(s0)
    return np.stack([
        np.stack([s0, s1, z], axis=1),
        np.stack([z, s0, s1], axis=1),
        np.stack([s1, z, s0], axis=1),
    ])


# vertex triples (3D) / pairs (2D) of each face, oriented outward for detJ > 0
# (z02/mesh_init.py:305-320 for tets, :43-45 for triangles)
TET_FACE_VERTS = np.array([[2, 1, 3], [0, 2, 3], [1, 0, 3], [0, 1, 2]], dtype=np.int64)
TRI_FACE_VERTS = np.array([[0, 1], [1, 2], [2, 0]], dtype=np.int64)

# d(xi)/d(face parameters) per face: (nface, D, D-1) in 3D, (nface, D) in 2D
# (z02/function_space.py:200-213)
TET_DRST_DUV = np.array([
    [[0, 0], [0, 1], [1, 0]],
    [[1, 0], [0, 0], [0, 1]],
    [[0, 1], [1, 0], [0, 0]],
    [[-1, -1], [1, 0], [0, 1]],
], dtype=np.float64)
TRI_DRST_DUV = np.array([[-1.0, 1.0], [0.0, -1.0], [1.0, 0.0]])


def _face_vertex_weights(face_lam, face_verts):
    """Barycentric weights of the face points w.r.t. the face's own vertex order."""
    return face_lam[:, face_verts]


def alignment_table(D):
    """Neighbour face-point permutations, computed geometrically.

    For a shared face seen from both sides, the neighbour lists the same vertices in
    the opposite orientation.  ``alnmt`` = k means the neighbour's first face vertex is
    this side's k-th face vertex (z02/mesh_init.py:379-387 for tets, :87-97 for
    triangles).  Row k of the result maps this side's face point g to the matching
    neighbour face point.  Returns (D, sngi) int64.
    """
    if D == 3:
        pnts, _ = tri_rule_12()
        # use face 3 (nodes 0,1,2) parameterised by (l1,l2,l3) = pnts; its vertex
        # weights in face-vertex order are pnts itself.
        this = pnts  # (sngi, 3) weights on face vertices (A, B, C)
        rows = []
        for k in range(3):
            # this face vertices (A,B,C) = (0,1,2); neighbour lists them as a rotation
            # of (A,C,B) starting at this side's vertex k.
            rev = [0, 2, 1]
            start = rev.index(k)
            nb_order = [rev[(start + i) % 3] for i in range(3)]
            # neighbour's face point h has weight pnts[h, i] on vertex nb_order[i]
            nb_w = np.zeros_like(this)
            for i in range(3):
                nb_w[:, nb_order[i]] = pnts[:, i]
            perm = []
            for gpt in range(this.shape[0]):
                dist = np.abs(nb_w - this[gpt][None]).sum(axis=1)
                perm.append(int(np.argmin(dist)))
            rows.append(perm)
        return np.array(rows, dtype=np.int64)
    else:
        pnts, _ = line_rule_4()
        rows = []
        for k in range(2):
            nb_order = [0, 1] if k == 0 else [1, 0]
            nb_w = np.zeros_like(pnts)
            for i in range(2):
                nb_w[:, nb_order[i]] = pnts[:, i]
            perm = []
            for gpt in range(pnts.shape[0]):
                dist = np.abs(nb_w - pnts[gpt][None]).sum(axis=1)
                perm.append(int(np.argmin(dist)))
            rows.append(perm)
        return np.array(rows, dtype=np.int64)


# --------------------------------------------------------------------------------------
# bundled element
# --------------------------------------------------------------------------------------
@dataclass
class Element:
    """Reference P3 element, mirroring ``Element``/``SHATRInew`` outputs
    (z02/function_space.py:14-80, z02/shape_function.py:459-1317)."""
    D: int
    nloc: int
    ngi: int
    sngi: int
    nface: int
    n: np.ndarray        # (nloc, ngi)
    nlx: np.ndarray      # (D, nloc, ngi)
    weight: np.ndarray   # (ngi,)
    sn: np.ndarray       # (nface, nloc, sngi)
    snlx: np.ndarray     # (nface, D, nloc, sngi)
    sweight: np.ndarray  # (sngi,)
    gi_align: np.ndarray  # (D, sngi)
    drst_duv: np.ndarray
    index: np.ndarray    # node multi-indices (nloc, D+1)
    face_verts: np.ndarray
    face_nodes: np.ndarray  # (nface, snloc) local nodes on each face
    p1_prolong: np.ndarray  # (nloc, D+1) P1DG -> P3DG (z02/mesh_init.py:224-235, 610-631)


def make_element(D: int) -> Element:
    if D == 3:
        index = TET_P3_INDEX
        L, w = tet_rule_24()
        lam = np.concatenate([L, (1.0 - L.sum(axis=1))[:, None]], axis=1)
        fpnts, sw = tri_rule_12()
        flam = tet_face_points(fpnts)
        face_verts = TET_FACE_VERTS
        drst = TET_DRST_DUV
    elif D == 2:
        index = TRI_P3_INDEX
        pn, w = tri_rule_12()
        lam = pn
        fpnts, sw = line_rule_4()
        flam = tri_face_points(fpnts)
        face_verts = TRI_FACE_VERTS
        drst = TRI_DRST_DUV
    else:
        raise ValueError("D must be 2 or 3")
    nloc = index.shape[0]
    nface = D + 1
    n, nlx = lagrange(index, lam)
    sngi = flam.shape[1]
    sn = np.zeros((nface, nloc, sngi))
    snlx = np.zeros((nface, D, nloc, sngi))
    for f in range(nface):
        sn[f], snlx[f] = lagrange(index, flam[f])
    # local nodes lying on face f: zero barycentric weight on the opposite vertex
    face_nodes = []
    for f in range(nface):
        opp = [v for v in range(D + 1) if v not in face_verts[f]][0]
        face_nodes.append(np.nonzero(index[:, opp] == 0)[0])
    face_nodes = np.array(face_nodes, dtype=np.int64)
    return Element(D=D, nloc=nloc, ngi=lam.shape[0], sngi=sngi, nface=nface,
                   n=n, nlx=nlx, weight=w, sn=sn, snlx=snlx, sweight=sw,
                   gi_align=alignment_table(D), drst_duv=drst, index=index,
                   face_verts=face_verts, face_nodes=face_nodes,
                   p1_prolong=index.astype(np.float64) / P)
