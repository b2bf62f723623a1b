"""On-disk formats around the solve: gmsh meshes in, mesh caches, solution checkpoints and
VTU output.

The reference reads gmsh ``.msh`` files through meshio (z02/config.py:53-56), caches its
pre-processing as ``.npy`` files next to the mesh, saves the solution with ``torch.save``
(z02/main.py:216) and writes VTU for visualisation (z02/output.py).  meshio is not a
dependency here; the reader below parses the ASCII gmsh formats 2.2 and 4.1 itself and
builds the P3 DG mesh (straight-sided P3 nodes in the reference's local order, oriented
as z02/mesh_init.py:276-289).  Boundary faces take ``glb_bcface_type`` from the physical
tag of the matching boundary facet (default 0, Dirichlet), interior faces -1.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np
import torch

from . import meshgen

_GMSH_TRI, _GMSH_TET, _GMSH_LINE = 2, 4, 1


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


def read_msh(path, bc_of_tag=None):
    """Read an ASCII gmsh 2.2 / 4.1 mesh of triangles (2D) or tetrahedra (3D).

    Returns (mesh, bctype): ``meshgen.Mesh`` with P3 nodes at the affine positions, and
    ``glb_bcface_type`` per (element, face): -1 interior, else the value of
    ``bc_of_tag[physical tag]`` of the matching boundary facet (default 0, Dirichlet).
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
    has_tet = any(t == _GMSH_TET for t, _, _ in elems)
    D = 3 if has_tet else 2
    vol_type, facet_type = (_GMSH_TET, _GMSH_TRI) if D == 3 else (_GMSH_TRI, _GMSH_LINE)
    vol = [e for e in elems if e[0] == vol_type]
    if not vol:
        raise ValueError(f"{path}: no triangles or tetrahedra")
    index = {int(t): i for i, t in enumerate(tags)}
    cells_tag = np.array([e[2][:D + 1] for e in vol], dtype=np.int64)
    used = np.unique(cells_tag)
    new_id = {int(t): i for i, t in enumerate(used)}
    verts = np.ascontiguousarray(xyz[[index[int(t)] for t in used], :D])
    cells = np.vectorize(new_id.__getitem__, otypes=[np.int64])(cells_tag)
    cells = meshgen._orient(verts, cells)
    table = meshgen.TET_P3 if D == 3 else meshgen.TRI_P3
    mesh = meshgen.Mesh(D=D, x_ref_in=meshgen._p3_nodes(verts, cells, table), cells=cells,
                        verts=verts)
    # boundary condition type of every face from the tagged facets
    facet_bc = {}
    for etype, ph, nodes in elems:
        if etype == facet_type and all(int(t) in new_id for t in nodes[:D]):
            key = tuple(sorted(new_id[int(t)] for t in nodes[:D]))
            facet_bc[key] = (bc_of_tag or {}).get(ph, 0)
    from .host import topology
    nbf, _, _ = topology(cells, D)
    face_verts = _FACE_VERTS[D]
    bctype = np.full(cells.shape[0] * (D + 1), -1, dtype=np.int64)
    for z in np.nonzero(nbf < 0)[0]:
        e, f = divmod(int(z), D + 1)
        key = tuple(sorted(int(cells[e, k]) for k in face_verts[f]))
        bctype[z] = facet_bc.get(key, 0)
    return mesh, bctype


# local vertex lists of the faces (z02/mesh_init.py:305-320 and z02/shape_function.py:1105-1116)
_FACE_VERTS = {3: [(2, 1, 3), (0, 2, 3), (1, 0, 3), (0, 1, 2)], 2: [(0, 1), (1, 2), (2, 0)]}


def write_msh(path, mesh, bctype=None, tag_of_bc=None):
    """Write ``mesh`` as ASCII gmsh 2.2: vertices, boundary facets (physical tag
    ``tag_of_bc[bctype]``, default 1) and the triangles / tetrahedra (physical tag 0)."""
    D = mesh.D
    vol_type, facet_type = (_GMSH_TET, _GMSH_TRI) if D == 3 else (_GMSH_TRI, _GMSH_LINE)
    from .host import topology
    nbf, _, bct = topology(mesh.cells, D)
    if bctype is None:
        bctype = bct
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(mesh.nverts)]
    for i, v in enumerate(mesh.verts):
        z = float(v[2]) if D == 3 else 0.0
        lines.append(f"{i + 1} {float(v[0])!r} {float(v[1])!r} {z!r}")
    lines.append("$EndNodes")
    elems = []
    for z in np.nonzero(nbf < 0)[0]:
        e, f = divmod(int(z), D + 1)
        tag = (tag_of_bc or {}).get(int(bctype[z]), 1)
        nodes = " ".join(str(int(mesh.cells[e, k]) + 1) for k in _FACE_VERTS[D][f])
        elems.append(f"{facet_type} 2 {tag} {tag} {nodes}")
    for c in mesh.cells:
        elems.append(f"{vol_type} 2 0 0 " + " ".join(str(int(k) + 1) for k in c))
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
