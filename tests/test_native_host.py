"""libdgmi.so loads, exports every entry point include/dgmi.h declares, and its host-only
set-up (element tables, topology, P1CG pattern, colouring, SFC, levels) equals the
oracle's.  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

from dg_nn_public_amd import _lib, host, meshgen
from oracle import dg, mg, tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dgmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dgmi_\w+)\s*\(", src)))


def test_header_declares_the_bound_entry_points():
    assert _declared() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


@pytest.mark.parametrize("D", [2, 3])
def test_native_element_tables_equal_oracle(D):
    t = host.element_tables(D)
    el = tables.make_element(D)
    for k in ("n", "nlx", "weight", "sn", "snlx", "sweight"):
        assert np.abs(t[k] - getattr(el, k)).max() < 1e-14, k
    assert np.array_equal(t["gi_align"], el.gi_align)


@pytest.fixture(scope="module", params=[2, 3], ids=["2d", "3d"])
def mesh(request):
    return meshgen.square_mesh(9) if request.param == 2 else meshgen.cube_mesh(4)


def test_native_topology_equals_oracle(mesh):
    got = host.topology(mesh.cells, mesh.D)
    ref = dg.topology(mesh.cells, mesh.D)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_native_sparsity_coloring_sfc_levels_equal_oracle(mesh):
    fina, cola = host.p1cg_sparsity(mesh.cells, mesh.D, mesh.nverts)
    fo, co = mg.p1cg_sparsity(mesh.cells, mesh.nverts)
    assert np.array_equal(fina, fo) and np.array_equal(cola, co)
    wc, nc = host.color2(fina, cola)
    wo, no = mg.color2(fo, co, mesh.nverts)
    assert nc == no and np.array_equal(wc, wo)
    sfc = host.sfc_order(mesh.verts)
    assert np.array_equal(sfc, mg.hilbert_order(mesh.verts))
    vals = np.random.default_rng(0).standard_normal(cola.size)
    lv = host.sfc_levels(fina, cola, vals, sfc)
    lo = mg.mg_on_P1CG_prep(fo, co, vals, sfc)
    assert len(lv) == lo.nlevel
    for (r, c, v), A in zip(lv, lo.A):
        assert np.array_equal(r, A.indptr) and np.array_equal(c, A.indices)
        assert np.allclose(v, A.data, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [1089, 15625, 117649, 1, 7])
def test_native_nlevel(n):
    assert host.nlevel_sfc(n) == mg.nlevel_sfc(n)


def test_load_fails_loudly_without_library(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "missing.so"))
        finally:
            _lib._lib = None
    _lib.load()
