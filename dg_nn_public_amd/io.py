"""On-disk formats around the solve: gmsh meshes in, mesh caches, solution checkpoints and
VTU output.

The reference reads gmsh ``.msh`` files through meshio (z02/config.py:53-56), caches its
pre-processing as ``.npy`` files next to the mesh, saves the solution with ``torch.save``
(z02/main.py:216) and writes VTU for visualisation (z02/output.py).  meshio is not a
dependency here; the reader below parses the ASCII gmsh formats 2.2 and 4.1 itself.

Cells: the reference reads P3 meshes (``tetra20`` / ``triangle10``, gmsh types 29 / 21,
z02/config.py:69-80); linear ones (types 4 / 2) are read too and get straight-sided P3
nodes.  A P3 cell's nodes are put in the reference's local order: a gmsh tetrahedron with
det(x1-x0, x2-x0, x3-x0) > 0 is renumbered with the reference's table
(z02/mesh_init.py:276-289); one with det < 0, which the reference leaves in gmsh order,
is first reflected (vertices 1 and 2 exchanged, with their edge and face nodes) and then
renumbered the same way; a clockwise triangle is reversed with the reference's table
(z02/mesh_init.py:19-31, whose orientation test reads the wrong points and never fires).
Boundary faces take ``glb_bcface_type`` from the physical group of the matching boundary
facet: ``bc_of_tag`` maps physical tags to types; without it, the boundary groups named in
``$PhysicalNames`` are numbered 0, 1, ... in file order, as the reference numbers its
cell sets (z02/mesh_init.py:697-741: e.g. "diri" 0, "neu" 1); untagged facets are 0
(Dirichlet); interior faces -1.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np
import torch

from . import meshgen

_GMSH_TRI, _GMSH_TET, _GMSH_LINE = 2, 4, 1
_GMSH_TRI10, _GMSH_TET20, _GMSH_LINE4 = 21, 29, 26
# the reference's renumbering of a gmsh tetra20 with det(x1-x0, x2-x0, x3-x0) > 0 into its
# local order (z02/mesh_init.py:276-281): local node i is gmsh node TET20_ORDER[i]
TET20_ORDER = [3, 1, 2, 0, 12, 13, 14, 15, 6, 7, 10, 11, 5, 4, 8, 9, 16, 19, 18, 17]
# gmsh tetra20 with vertices 1 and 2 exchanged (its edge and face nodes follow): a left-
# handed cell becomes right-handed
TET20_REFLECT = [0, 2, 1, 3, 9, 8, 7, 6, 5, 4, 10, 11, 14, 15, 12, 13, 16, 18, 17, 19]
# triangle10: gmsh order is the reference's for counter-clockwise cells; clockwise ones
# are reversed with the reference's table (z02/mesh_init.py:20)
TRI10_REVERSE = [0, 2, 1, 8, 7, 6, 5, 4, 3, 9]


# ---------------------------------------------------------------------------- gmsh
def _tokens(path):
    with open(path, "r") as fh:
        return [ln.strip() for ln in fh]


def _section(lines, name):
    try:
        a = lines.index(f"${name}")
        b = lines.index(f"$End{name}", a)
    except ValueError:
        return None
    return lines[a + 1:b]


def _parse_v2(nodes_sec, elems_sec):
    nn = int(nodes_sec[0])
    tags = np.empty(nn, dtype=np.int64)
    xyz = np.empty((nn, 3))
    for i in range(nn):
        p = nodes_sec[1 + i].split()
        tags[i] = int(p[0])
        xyz[i] = [float(v) for v in p[1:4]]
    ne = int(elems_sec[0])
    elems = []  # (type, physical tag, node tags)
    for i in range(ne):
        p = [int(v) for v in elems_sec[1 + i].split()]
        etype, ntag = p[1], p[2]
        phys = p[3] if ntag > 0 else 0
        elems.append((etype, phys, p[3 + ntag:]))
    return tags, xyz, elems


def _parse_v4(nodes_sec, elems_sec, ent_sec):
    # physical tags of entities (dim, tag) from $Entities
    phys = {}
    if ent_sec is not None:
        counts = [int(v) for v in ent_sec[0].split()]
        row = 1
        for dim, cnt in enumerate(counts):
            for _ in range(cnt):
                p = ent_sec[row].split()
                row += 1
                tag = int(p[0])
                k = 4 if dim == 0 else 7  # point: tag x y z; others: tag + bounding box
                nph = int(p[k])
                phys[(dim, tag)] = int(p[k + 1]) if nph > 0 else 0
    nb, nn = (int(v) for v in nodes_sec[0].split()[:2])
    tags = np.empty(nn, dtype=np.int64)
    xyz = np.empty((nn, 3))
    row, k = 1, 0
    for _ in range(nb):
        _, _, _, cnt = (int(v) for v in nodes_sec[row].split())
        row += 1
        for i in range(cnt):
            tags[k + i] = int(nodes_sec[row + i])
        for i in range(cnt):
            xyz[k + i] = [float(v) for v in nodes_sec[row + cnt + i].split()[:3]]
        row += 2 * cnt
        k += cnt
    nb = int(elems_sec[0].split()[0])
    elems = []
    row = 1
    for _ in range(nb):
        dim, etag, etype, cnt = (int(v) for v in elems_sec[row].split())
        row += 1
        ph = phys.get((dim, etag), 0)
        for i in range(cnt):
            p = [int(v) for v in elems_sec[row + i].split()]
            elems.append((etype, ph, p[1:]))
        row += cnt
    return tags, xyz, elems


def _physical_names(lines):
    """[(dim, tag, name)] of $PhysicalNames in file order."""
    sec = _section(lines, "PhysicalNames")
    out = []
    if sec:
        for ln in sec[1:1 + int(sec[0])]:
            p = ln.split(maxsplit=2)
            out.append((int(p[0]), int(p[1]), p[2].strip().strip('"')))
    return out


def _p3_cells(vol, index, xyz, D):
    """Node tags of P3 cells in the reference's local order (see the module docstring)."""
    tags = np.array([e[2][:10 if D == 2 else 20] for e in vol], dtype=np.int64)
    pts = xyz[np.vectorize(index.__getitem__, otypes=[np.int64])(tags)][..., :D]
    if D == 3:
        v = pts[:, :4]
        det = np.linalg.det(v[:, 1:4] - v[:, 0:1])
        refl = det < 0
        tags[refl] = tags[refl][:, TET20_REFLECT]
        return tags[:, TET20_ORDER]
    v = pts[:, :3]
    det = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - \
        (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0])
    cw = det < 0
    tags[cw] = tags[cw][:, TRI10_REVERSE]
    return tags


def read_msh(path, bc_of_tag=None):
    """Read an ASCII gmsh 2.2 / 4.1 mesh of P3 (tetra20 / triangle10) or linear
    tetrahedra / triangles.

    Returns (mesh, bctype): ``meshgen.Mesh`` with the P3 nodes in the reference's local
    order (the file's node positions for P3 cells, straight-sided ones for linear cells)
    and ``glb_bcface_type`` per (element, face) as described in the module docstring.
    """
    lines = _tokens(path)
    fmt = _section(lines, "MeshFormat")
    if fmt is None:
        raise ValueError(f"{path}: no $MeshFormat section")
    version, filetype = fmt[0].split()[:2]
    if int(filetype) != 0:
        raise ValueError(f"{path}: binary gmsh files are not supported")
    nodes_sec, elems_sec = _section(lines, "Nodes"), _section(lines, "Elements")
    if nodes_sec is None or elems_sec is None:
        raise ValueError(f"{path}: missing $Nodes or $Elements")
    if version.startswith("2"):
        tags, xyz, elems = _parse_v2(nodes_sec, elems_sec)
    elif version.startswith("4"):
        tags, xyz, elems = _parse_v4(nodes_sec, elems_sec, _section(lines, "Entities"))
    else:
        raise ValueError(f"{path}: gmsh format {version} is not supported")
    types = {t for t, _, _ in elems}
    if _GMSH_TET20 in types or _GMSH_TET in types:
        D = 3
    elif _GMSH_TRI10 in types or _GMSH_TRI in types:
        D = 2
    else:
        raise ValueError(f"{path}: no triangles or tetrahedra (gmsh types 2, 4, 21, 29)")
    p3_type, lin_type = (_GMSH_TET20, _GMSH_TET) if D == 3 else (_GMSH_TRI10, _GMSH_TRI)
    facet_types = (_GMSH_TRI10, _GMSH_TRI) if D == 3 else (_GMSH_LINE4, _GMSH_LINE)
    index = {int(t): i for i, t in enumerate(tags)}
    p3 = [e for e in elems if e[0] == p3_type]
    lin = [e for e in elems if e[0] == lin_type]
    if p3 and lin:
        raise ValueError(f"{path}: mixes P3 and linear cells")
    if p3:
        local = _p3_cells(p3, index, xyz, D)
        cells_tag = local[:, :D + 1]
    else:
        cells_tag = np.array([e[2][:D + 1] for e in lin], dtype=np.int64)
    used = np.unique(cells_tag)
    new_id = {int(t): i for i, t in enumerate(used)}
    verts = np.ascontiguousarray(xyz[[index[int(t)] for t in used], :D])
    cells = np.vectorize(new_id.__getitem__, otypes=[np.int64])(cells_tag)
    if p3:
        rows = np.vectorize(index.__getitem__, otypes=[np.int64])(local)
        x_ref_in = np.ascontiguousarray(np.moveaxis(xyz[rows][..., :D], 2, 1))
    else:
        cells = meshgen._orient(verts, cells)
        table = meshgen.TET_P3 if D == 3 else meshgen.TRI_P3
        x_ref_in = meshgen._p3_nodes(verts, cells, table)
    mesh = meshgen.Mesh(D=D, x_ref_in=x_ref_in, cells=cells, verts=verts)
    # physical group of every boundary facet -> glb_bcface_type
    if bc_of_tag is None:
        bnd = [t for dim, t, _ in _physical_names(lines) if dim == D - 1]
        bc_of_tag = {t: i for i, t in enumerate(bnd)}
    facet_bc = {}
    for etype, ph, nodes in elems:
        if etype in facet_types and all(int(t) in new_id for t in nodes[:D]):
            key = tuple(sorted(new_id[int(t)] for t in nodes[:D]))
            facet_bc[key] = bc_of_tag.get(ph, 0)
    from .host import topology
    nbf, _, _ = topology(cells, D)
    face_verts = _FACE_VERTS[D]
    bctype = np.full(cells.shape[0] * (D + 1), -1, dtype=np.int64)
    for z in np.nonzero(nbf < 0)[0]:
        e, f = divmod(int(z), D + 1)
        key = tuple(sorted(int(cells[e, k]) for k in face_verts[f]))
        bctype[z] = facet_bc.get(key, 0)
    return mesh, bctype


def affine_deviation(mesh):
    """Largest distance of a P3 node from its straight-sided position, relative to the
    element size (0 for an affine mesh; the solver accepts only affine elements)."""
    table = meshgen.TET_P3 if mesh.D == 3 else meshgen.TRI_P3
    ref = meshgen._p3_nodes(mesh.verts, mesh.cells, table)
    v = mesh.verts[mesh.cells]
    h = np.abs(v[:, :mesh.D] - v[:, mesh.D:]).max(axis=(1, 2))
    return float((np.abs(mesh.x_ref_in - ref).max(axis=(1, 2)) / h).max())


# local vertex lists of the faces (z02/mesh_init.py:305-320 and z02/shape_function.py:1105-1116)
_FACE_VERTS = {3: [(2, 1, 3), (0, 2, 3), (1, 0, 3), (0, 1, 2)], 2: [(0, 1), (1, 2), (2, 0)]}


# gmsh's own node numbering of the cubic cells (gmsh reference manual, "node ordering"):
# edges and their direction, then face-interior nodes by face vertex triples
_GMSH_TET_EDGES = [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)]
_GMSH_TET_FACES = [(0, 2, 1), (0, 1, 3), (0, 3, 2), (3, 1, 2)]
_GMSH_TRI_EDGES = [(0, 1), (1, 2), (2, 0)]


def write_msh(path, mesh, bctype=None, tag_of_bc=None, order=1, physical_names=None,
              reverse=None):
    """Write ``mesh`` as ASCII gmsh 2.2: nodes, boundary facets (physical tag
    ``tag_of_bc[bctype]``, default 1) and the cells (physical tag 0, or the tag named by
    ``physical_names`` for dimension D).

    order 1 writes linear cells (types 4 / 2, facets 2 / 1); order 3 writes cubic cells
    (tetra20 / triangle10, facets triangle10 / line4) with nodes shared between cells and
    numbered as gmsh numbers them, at straight-sided positions.  ``physical_names``
    ({tag: (dim, name)}) adds a $PhysicalNames section; ``reverse`` (bool per cell)
    writes those cells with their vertices 1 and 2 exchanged (opposite orientation)."""
    D = mesh.D
    if order not in (1, 3):
        raise ValueError("order must be 1 or 3")
    from .host import topology
    nbf, _, bct = topology(mesh.cells, D)
    if bctype is None:
        bctype = bct
    pts = [np.asarray(v, dtype=np.float64) for v in mesh.verts]
    edge_nodes, face_nodes = {}, {}

    def node_at(x):
        pts.append(np.asarray(x, dtype=np.float64))
        return len(pts) - 1

    def edge(a, b):
        """the two nodes on edge a -> b, in that direction"""
        key = (min(a, b), max(a, b))
        if key not in edge_nodes:
            xa, xb = mesh.verts[key[0]], mesh.verts[key[1]]
            edge_nodes[key] = (node_at(xa + (xb - xa) / 3.0), node_at(xa + 2.0 * (xb - xa) / 3.0))
        n0, n1 = edge_nodes[key]
        return [n0, n1] if key == (a, b) else [n1, n0]

    def face(a, b, c):
        key = tuple(sorted((a, b, c)))
        if key not in face_nodes:
            face_nodes[key] = node_at(mesh.verts[list(key)].mean(axis=0))
        return face_nodes[key]

    def cubic(vs):
        if D == 3:
            out = list(vs)
            for a, b in _GMSH_TET_EDGES:
                out += edge(vs[a], vs[b])
            out += [face(vs[a], vs[b], vs[c]) for a, b, c in _GMSH_TET_FACES]
            return out
        out = list(vs)
        for a, b in _GMSH_TRI_EDGES:
            out += edge(vs[a], vs[b])
        return out + [face(*vs)]

    vol_type = (_GMSH_TET20 if D == 3 else _GMSH_TRI10) if order == 3 else \
        (_GMSH_TET if D == 3 else _GMSH_TRI)
    facet_type = (_GMSH_TRI10 if D == 3 else _GMSH_LINE4) if order == 3 else \
        (_GMSH_TRI if D == 3 else _GMSH_LINE)
    vol_tag = 0
    for t, (dim, _) in (physical_names or {}).items():
        if dim == D:
            vol_tag = t
    elems = []
    for z in np.nonzero(nbf < 0)[0]:
        e, f = divmod(int(z), D + 1)
        tag = (tag_of_bc or {}).get(int(bctype[z]), 1)
        fv = [int(mesh.cells[e, k]) for k in _FACE_VERTS[D][f]]
        if order == 3:
            if D == 3:  # triangle10 facet: vertices, edges (0,1), (1,2), (2,0), centre
                nodes = fv + edge(fv[0], fv[1]) + edge(fv[1], fv[2]) + edge(fv[2], fv[0]) + \
                    [face(*fv)]
            else:  # line4: end points, then the two interior nodes
                nodes = fv + edge(fv[0], fv[1])
        else:
            nodes = fv
        elems.append(f"{facet_type} 2 {tag} {tag} " + " ".join(str(k + 1) for k in nodes))
    for i, c in enumerate(mesh.cells):
        vs = [int(k) for k in c]
        if reverse is not None and reverse[i]:
            vs[1], vs[2] = vs[2], vs[1]
        nodes = cubic(vs) if order == 3 else vs
        elems.append(f"{vol_type} 2 {vol_tag} {vol_tag} " + " ".join(str(k + 1) for k in nodes))
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat"]
    if physical_names:
        lines += ["$PhysicalNames", str(len(physical_names))]
        lines += [f'{dim} {t} "{name}"' for t, (dim, name) in physical_names.items()]
        lines.append("$EndPhysicalNames")
    lines += ["$Nodes", str(len(pts))]
    for i, v in enumerate(pts):
        z = float(v[2]) if D == 3 else 0.0
        lines.append(f"{i + 1} {float(v[0])!r} {float(v[1])!r} {z!r}")
    lines.append("$EndNodes")
    lines += ["$Elements", str(len(elems))]
    lines += [f"{i + 1} {s}" for i, s in enumerate(elems)]
    lines.append("$EndElements")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------------------- caches
def save_mesh_cache(path, mesh, bctype=None, sfc=None):
    """Pre-processed mesh as one ``.npz`` (the reference keeps such arrays as ``.npy``
    files next to the mesh): cells, vertices, P3 nodes, face types and optionally the
    SFC order of the P1 vertices."""
    arrays = dict(D=np.int64(mesh.D), cells=mesh.cells, verts=mesh.verts, x_ref_in=mesh.x_ref_in)
    if bctype is not None:
        arrays["bctype"] = np.asarray(bctype, dtype=np.int64)
    if sfc is not None:
        arrays["sfc"] = np.asarray(sfc, dtype=np.int64)
    np.savez(path, **arrays)


def load_mesh_cache(path):
    """Inverse of ``save_mesh_cache`` (no pickled objects are read)."""
    with np.load(path, allow_pickle=False) as z:
        mesh = meshgen.Mesh(D=int(z["D"]), x_ref_in=z["x_ref_in"].copy(), cells=z["cells"].copy(),
                            verts=z["verts"].copy())
        bct = z["bctype"].copy() if "bctype" in z else None
        sfc = z["sfc"].copy() if "sfc" in z else None
    return mesh, bct, sfc


def save_checkpoint(path, x, **meta):
    """Solution checkpoint, as z02/main.py:216 saves ``x_i`` with torch.save; ``meta`` holds
    plain numbers or strings (iterations, residual, time)."""
    torch.save({"x": x.detach().cpu(), "meta": dict(meta)}, path)


def load_checkpoint(path, device=None):
    """Load a checkpoint written by ``save_checkpoint`` (tensors and plain values only)."""
    ck = torch.load(path, map_location="cpu", weights_only=True)
    x = ck["x"] if device is None else ck["x"].to(device)
    return x, ck.get("meta", {})


# ---------------------------------------------------------------------------- VTU
def write_vtu(path, mesh, point_fields=None, cell_fields=None):
    """ASCII VTK unstructured grid of the DG solution: every element carries its own
    copies of its vertices (the field is discontinuous); ``point_fields`` are DG nodal
    vectors (nele*nloc, the reference layout) sampled at the vertex nodes (local nodes
    0..D), ``cell_fields`` one value per element."""
    D, ne, nb = mesh.D, mesh.nele, mesh.D + 1
    pts = np.zeros((ne * nb, 3))
    pts[:, :D] = mesh.verts[mesh.cells].reshape(-1, D)
    vtk_type = 10 if D == 3 else 5  # VTK_TETRA / VTK_TRIANGLE
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="0.1", byte_order="LittleEndian")
    grid = ET.SubElement(root, "UnstructuredGrid")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(ne * nb), NumberOfCells=str(ne))

    def arr(parent, name, data, ncomp=1, dtype="Float64"):
        a = ET.SubElement(parent, "DataArray", type=dtype, Name=name,
                          NumberOfComponents=str(ncomp), format="ascii")
        a.text = " ".join(repr(float(v)) if dtype.startswith("Float") else str(int(v))
                          for v in np.asarray(data).ravel())

    arr(ET.SubElement(piece, "Points"), "Points", pts, 3)
    cells = ET.SubElement(piece, "Cells")
    arr(cells, "connectivity", np.arange(ne * nb), dtype="Int64")
    arr(cells, "offsets", nb * np.arange(1, ne + 1), dtype="Int64")
    arr(cells, "types", np.full(ne, vtk_type), dtype="UInt8")
    if point_fields:
        pd = ET.SubElement(piece, "PointData")
        for name, v in point_fields.items():
            v = v.detach().double().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
            arr(pd, name, v.reshape(ne, mesh.nloc)[:, :nb])
    if cell_fields:
        cd = ET.SubElement(piece, "CellData")
        for name, v in cell_fields.items():
            arr(cd, name, np.asarray(v).reshape(ne))
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")
