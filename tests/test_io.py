"""On-disk formats: gmsh read/write round trips, a hand-written gmsh 4.1 file with tagged
boundary facets, mesh caches, checkpoints and VTU output (CPU only)."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest
import torch

from dg_nn_public_amd import host, io, meshgen


@pytest.mark.parametrize("D", [2, 3])
def test_msh_round_trip(tmp_path, D):
    m = meshgen.square_mesh(5, seed=3) if D == 2 else meshgen.cube_mesh(3, seed=4)
    p = tmp_path / "m.msh"
    io.write_msh(p, m)
    r, bct = io.read_msh(p)
    assert r.D == D and np.array_equal(r.cells, m.cells)
    assert np.allclose(r.verts, m.verts, rtol=0, atol=1e-15)
    assert np.allclose(r.x_ref_in, m.x_ref_in, rtol=0, atol=1e-14)
    _, _, ref_bct = host.topology(m.cells, D)
    assert np.array_equal(bct, ref_bct)  # boundary faces Dirichlet (0), interior -1


def test_msh_boundary_tags_map_to_face_types(tmp_path):
    m = meshgen.square_mesh(4)
    _, _, bct = host.topology(m.cells, 2)
    # mark the faces on x = 0 as Neumann (1) through a physical tag
    bct = bct.copy()
    for z in np.nonzero(bct == 0)[0]:
        e, f = divmod(int(z), 3)
        a, b = [(0, 1), (1, 2), (2, 0)][f]
        if abs(m.verts[m.cells[e, a], 0]) < 1e-12 and abs(m.verts[m.cells[e, b], 0]) < 1e-12:
            bct[z] = 1
    p = tmp_path / "tagged.msh"
    io.write_msh(p, m, bctype=bct, tag_of_bc={0: 11, 1: 12})
    _, got = io.read_msh(p, bc_of_tag={11: 0, 12: 1})
    assert np.array_equal(got, bct) and (got == 1).sum() == 4


GMSH41 = """$MeshFormat
4.1 0 8
$EndMeshFormat
$Entities
0 0 2 1
1 0 0 0 1 1 0 1 5 0
2 0 0 0 1 0 1 1 6 0
1 0 0 0 1 1 1 0 0
$EndEntities
$Nodes
1 4 1 4
3 1 0 4
1
2
3
4
0 0 0
1 0 0
0 1 0
0 0 1
$EndNodes
$Elements
3 5 1 5
2 1 2 2
1 1 2 3
2 1 2 4
2 2 2 2
3 1 3 4
4 2 3 4
3 1 4 1
5 1 2 3 4
$EndElements
"""


def test_gmsh41_parsing_with_physical_tags(tmp_path):
    p = tmp_path / "one.msh"
    p.write_text(GMSH41)
    m, bct = io.read_msh(p, bc_of_tag={5: 0, 6: 1})
    assert m.D == 3 and m.nele == 1 and m.nverts == 4
    # orientation: det(x_k - x_3) > 0 for the stored vertex order
    v = m.verts[m.cells[0]]
    assert np.linalg.det(v[:3] - v[3]) > 0
    faces = {tuple(sorted(m.cells[0, k] for k in fv)): bct[f]
             for f, fv in enumerate([(2, 1, 3), (0, 2, 3), (1, 0, 3), (0, 1, 2)])}
    # vertex ids 0..3 are the nodes 1..4: faces {1,2,3} and {1,2,4} tagged 5 -> 0,
    # {1,3,4} and {2,3,4} tagged 6 -> 1
    assert faces[(0, 1, 2)] == 0 and faces[(0, 1, 3)] == 0
    assert faces[(0, 2, 3)] == 1 and faces[(1, 2, 3)] == 1


def test_mesh_cache_and_checkpoint_round_trip(tmp_path):
    m = meshgen.cube_mesh(2)
    _, _, bct = host.topology(m.cells, 3)
    sfc = host.sfc_order(m.verts)
    io.save_mesh_cache(tmp_path / "c.npz", m, bct, sfc)
    r, b2, s2 = io.load_mesh_cache(tmp_path / "c.npz")
    assert np.array_equal(r.cells, m.cells) and np.array_equal(r.x_ref_in, m.x_ref_in)
    assert np.array_equal(b2, bct) and np.array_equal(s2, sfc)
    x = torch.randn(m.nele * m.nloc, dtype=torch.float64)
    io.save_checkpoint(tmp_path / "x.pt", x, its=42, relres=3.5e-11)
    y, meta = io.load_checkpoint(tmp_path / "x.pt")
    assert torch.equal(x, y) and meta == {"its": 42, "relres": 3.5e-11}


def test_vtu_output(tmp_path):
    m = meshgen.square_mesh(3)
    pb = meshgen.sine_problem(m, sigma=1.0)
    io.write_vtu(tmp_path / "u.vtu", m, point_fields={"u": pb.u_exact},
                 cell_fields={"kappa": pb.kappa})
    root = ET.parse(tmp_path / "u.vtu").getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfCells")) == m.nele
    assert int(piece.get("NumberOfPoints")) == 3 * m.nele
    u = np.array(piece.find("PointData/DataArray").text.split(), dtype=float)
    assert np.allclose(u, pb.u_exact.reshape(m.nele, m.nloc)[:, :3].ravel())
    types = np.array(piece.find("Cells/DataArray[@Name='types']").text.split(), dtype=int)
    assert (types == 5).all()


def _match_nodes(a, b):
    """per element, the permutation p with a[:, :, p] == b (node positions)"""
    out = np.empty(a.shape[::2], dtype=np.int64)
    for e in range(a.shape[0]):
        d = np.abs(a[e][:, :, None] - b[e][:, None, :]).max(axis=0)  # (nloc_a, nloc_b)
        out[e] = d.argmin(axis=0)
        assert d.min(axis=0).max() < 1e-12
    return out


@pytest.mark.parametrize("D", [2, 3])
def test_cubic_msh_is_read_in_the_reference_node_order(tmp_path, D):
    """A P3 gmsh file (tetra20 / triangle10 in gmsh's own node numbering, some cells
    reversed) is read into the reference's local order: every node sits where the
    reference's P3 table puts it for the element's vertices (z02/mesh_init.py:276-289),
    every element has det J > 0, and the operator equals the one of the original mesh up
    to the element-local renumbering."""
    from oracle import dg
    m = meshgen.square_mesh(4, seed=1) if D == 2 else meshgen.cube_mesh(2, seed=2)
    rev = np.random.default_rng(0).random(m.nele) < 0.4
    p = tmp_path / "p3.msh"
    io.write_msh(p, m, order=3, reverse=rev)
    r, bct = io.read_msh(p)
    assert r.D == D and r.nele == m.nele and r.x_ref_in.shape == m.x_ref_in.shape
    assert io.affine_deviation(r) < 1e-12
    v = r.verts[r.cells]
    J = v[:, :D] - v[:, D:]
    assert np.all(np.linalg.det(J) > 0)
    perm = _match_nodes(r.x_ref_in, m.x_ref_in)  # read node perm[e, n] = original node n
    kap = np.exp(np.random.default_rng(1).standard_normal(m.nele))
    u = np.random.default_rng(2).standard_normal((m.nele, m.nloc))
    ur = np.take_along_axis(np.zeros_like(u), perm, 1)
    ur[np.arange(m.nele)[:, None], perm] = u
    y = dg.apply_A(dg.make_space(m, kap), u.reshape(-1)).reshape(m.nele, -1)
    yr = dg.apply_A(dg.make_space(r, kap), ur.reshape(-1)).reshape(m.nele, -1)
    assert np.abs(yr[np.arange(m.nele)[:, None], perm] - y).max() < 1e-11 * np.abs(y).max()
    _, _, ref_bct = host.topology(r.cells, D)
    assert np.array_equal(bct, ref_bct)


def test_tetra20_right_handed_cells_use_the_reference_renumbering(tmp_path):
    """For a gmsh cell with det(x1-x0, x2-x0, x3-x0) > 0 the reader's local node i is
    gmsh node TET20_ORDER[i] (the reference's table): its vertices are gmsh 3, 1, 2, 0."""
    m = meshgen.cube_mesh(1, jitter=0.0)
    one = meshgen.Mesh(D=3, x_ref_in=m.x_ref_in[:1], cells=m.cells[:1], verts=m.verts)
    gm = one.verts[one.cells[0]]
    if np.linalg.det(gm[1:] - gm[0]) < 0:  # make the written cell right-handed in gmsh terms
        c = one.cells.copy()
        c[0, [1, 2]] = c[0, [2, 1]]
        one = meshgen.Mesh(D=3, x_ref_in=one.x_ref_in, cells=c, verts=one.verts)
    p = tmp_path / "one.msh"
    io.write_msh(p, one, order=3)
    r, _ = io.read_msh(p)
    g = one.cells[0]
    assert np.allclose(r.verts[r.cells[0]], one.verts[[g[3], g[1], g[2], g[0]]], atol=1e-15)
    assert io.affine_deviation(r) < 1e-12


def test_physical_names_number_boundary_groups_in_file_order(tmp_path):
    """The reference's cube (z21-cube-mesh/cube_ho_poi.geo): groups "diri" (five faces)
    and "neu" (z = 1) become glb_bcface_type 0 and 1, as its cell-set order numbers them
    (z02/mesh_init.py:697-741)."""
    m = meshgen.cube_mesh(2, seed=5)
    _, _, bct = host.topology(m.cells, 3)
    fv = [(2, 1, 3), (0, 2, 3), (1, 0, 3), (0, 1, 2)]
    want = bct.copy()
    for z in np.nonzero(bct == 0)[0]:
        e, f = divmod(int(z), 4)
        if np.all(np.abs(m.verts[m.cells[e, list(fv[f])], 2] - 1.0) < 1e-12):
            want[z] = 1
    p = tmp_path / "cube.msh"
    io.write_msh(p, m, bctype=want, tag_of_bc={0: 1, 1: 2}, order=3,
                 physical_names={1: (2, "diri"), 2: (2, "neu"), 3: (3, "fluid")})
    r, got = io.read_msh(p)

    def by_face(mesh, t):  # face centroid -> type (the reader renumbers local faces)
        out = {}
        for z in np.nonzero(t >= 0)[0]:
            e, f = divmod(int(z), 4)
            c = mesh.verts[mesh.cells[e, list(fv[f])]].mean(axis=0)
            out[tuple(np.round(c, 12))] = int(t[z])
        return out

    assert by_face(r, got) == by_face(m, want) and (got == 1).sum() == 8
