"""NumPy wrappers of the host-only set-up entry points of libdgmi.so.

These replace the reference's f2py Fortran set-up modules (``get_nb``, ``sfc``,
``map_to_sfc_matrix``) and its Python set-up loops; they need no GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def _i64(a):
    return a.ctypes.data_as(_lib._I64P)


def _f64(a):
    return a.ctypes.data_as(_lib._DP)


def element_tables(D: int) -> dict:
    nl, ng, sg = (20, 24, 12) if D == 3 else (10, 12, 4)
    nf = D + 1
    t = dict(n=np.empty((nl, ng)), nlx=np.empty((D, nl, ng)), weight=np.empty(ng),
             sn=np.empty((nf, nl, sg)), snlx=np.empty((nf, D, nl, sg)), sweight=np.empty(sg),
             gi_align=np.empty((D, sg), np.int64))
    lib = _lib.load()
    check(lib.dgmi_host_element_tables(D, *[_f64(t[k]) for k in
                                            ("n", "nlx", "weight", "sn", "snlx", "sweight")],
                                       _i64(t["gi_align"])), "dgmi_host_element_tables")
    return t


def topology(cells: np.ndarray, D: int):
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    n = cells.shape[0] * (D + 1)
    nbf, aln, bct = (np.empty(n, np.int64) for _ in range(3))
    check(_lib.load().dgmi_host_topology(_i64(cells), cells.shape[0], D, _i64(nbf), _i64(aln),
                                         _i64(bct)), "dgmi_host_topology")
    return nbf, aln, bct


def p1cg_sparsity(cells: np.ndarray, D: int, nverts: int):
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    cap = cells.shape[0] * (D + 1) ** 2
    fina = np.empty(nverts + 1, np.int64)
    cola = np.empty(cap, np.int64)
    ncola = C.c_int64()
    check(_lib.load().dgmi_host_p1cg_sparsity(_i64(cells), cells.shape[0], D, nverts, cap,
                                              _i64(fina), _i64(cola), C.byref(ncola)),
          "dgmi_host_p1cg_sparsity")
    return fina, cola[:ncola.value].copy()


def color2(fina, cola):
    fina = np.ascontiguousarray(fina, dtype=np.int64)
    cola = np.ascontiguousarray(cola, dtype=np.int64)
    n = fina.size - 1
    wc = np.empty(n, np.int32)
    nc = C.c_int32()
    check(_lib.load().dgmi_host_color2(_i64(fina), _i64(cola), n,
                                       wc.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nc)),
          "dgmi_host_color2")
    return wc.astype(np.int64), nc.value


def sfc_order(verts: np.ndarray):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    sfc = np.empty(verts.shape[0], np.int64)
    check(_lib.load().dgmi_host_sfc_order(_f64(verts), verts.shape[0], verts.shape[1], _i64(sfc)),
          "dgmi_host_sfc_order")
    return sfc


def nlevel_sfc(n: int) -> int:
    return int(_lib.load().dgmi_host_nlevel_sfc(n))


def sfc_levels(fina, cola, values, sfc):
    """Per-level CSR (rowptr, col, val) lists of the SFC level operators."""
    fina = np.ascontiguousarray(fina, dtype=np.int64)
    cola = np.ascontiguousarray(cola, dtype=np.int64)
    values = np.ascontiguousarray(values, dtype=np.float64)
    sfc = np.ascontiguousarray(sfc, dtype=np.int64)
    n = fina.size - 1
    L = nlevel_sfc(n)
    cap_rows = 2 * n + 2 * L + 2
    cap_nnz = cola.size * L
    nlev = C.c_int64()
    level_n = np.empty(L, np.int64)
    rp = np.empty(cap_rows, np.int64)
    col = np.empty(cap_nnz, np.int64)
    val = np.empty(cap_nnz)
    check(_lib.load().dgmi_host_sfc_levels(_i64(fina), _i64(cola), _f64(values), _i64(sfc), n,
                                           cap_rows, cap_nnz, C.byref(nlev), _i64(level_n),
                                           _i64(rp), _i64(col), _f64(val)), "dgmi_host_sfc_levels")
    out = []
    ro = no = 0
    for l in range(nlev.value):
        nl = int(level_n[l])
        r = rp[ro:ro + nl + 1].copy()
        nnz = int(r[-1])
        out.append((r, col[no:no + nnz].copy(), val[no:no + nnz].copy()))
        ro += nl + 1
        no += nnz
    return out
