"""ctypes binding of libdgmi.so (include/dgmi.h).

The library is built in-tree by ``python -m dg_nn_public_amd.build``.  Loading fails
loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgmi.so")

DGMI_F64, DGMI_F32, DGMI_BF16 = 0, 1, 2
DGMI_SMOOTH_JACOBI3, DGMI_SMOOTH_DIRECT = 0, 1

# every entry point declared in include/dgmi.h, with its ctypes signature
_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)


class MeshDesc(C.Structure):
    _fields_ = [
        ("D", C.c_int32),
        ("nele", C.c_int64),
        ("nverts", C.c_int64),
        ("x_ref_in", _DP),
        ("cells", _I64P),
        ("verts", _DP),
        ("kappa", _DP),
        ("bctype", _I64P),
        ("eta_e", C.c_double),
        ("face_kappa", C.c_int32),
        ("stream", C.c_void_p),
    ]


SIGNATURES = {
    "dgmi_ctx_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(_P)]),
    "dgmi_ctx_destroy": (C.c_int, [_P]),
    "dgmi_ctx_create_batch": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, _I64P, _I64P,
                                        C.POINTER(_P)]),
    "dgmi_ctx_create_batch_parts": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, _I64P, _I64P,
                                              C.POINTER(_DP), C.POINTER(_P)]),
    "dgmi_ctx_info": (C.c_int, [_P, _I64P]),
    "dgmi_setup_timing": (C.c_int, [_P, _DP]),
    "dgmi_topology": (C.c_int, [_P, _I64P, _I64P, _I64P]),
    "dgmi_element_tables": (C.c_int, [_P, _DP, _DP, _DP, _DP, _DP, _DP, _I64P]),
    "dgmi_apply": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "dgmi_residual_dd": (C.c_int, [_P, _P, _P, _P, _P]),
    "dgmi_release_cached_memory": (C.c_int, []),
    "dgmi_rhs": (C.c_int, [_P, _P, _P, _P, _P, C.c_double, _P, _P]),
    "dgmi_l2_error": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "dgmi_bdiag": (C.c_int, [_P, _P, _P, _P]),
    "dgmi_smoother_setup": (C.c_int, [_P, C.c_int, C.c_double, _P]),
    "dgmi_smooth": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "dgmi_smooth_n": (C.c_int, [_P, C.c_int, _P, _P, C.c_int, _P]),
    "dgmi_mg_setup": (C.c_int, [_P, _P]),
    "dgmi_mg_info": (C.c_int, [_P, _I64P, C.c_int]),
    "dgmi_mg_export": (C.c_int, [_P, _I64P, _I64P, _DP, _I64P]),
    "dgmi_restrict": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "dgmi_prolong_add": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "dgmi_sfc_vcycle": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "dgmi_precond": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "dgmi_gmres_solve": (C.c_int, [_P, C.c_int, _P, _P, C.c_double, C.c_int, C.c_int, _DP, _P]),
    "dgmi_cg_solve": (C.c_int, [_P, _P, _P, C.c_double, C.c_double, C.c_int, _DP, _P]),
    "dgmi_level_smooth": (C.c_int, [C.c_int, C.c_int64, _P, _P, _P, _P, C.c_double, _P, _P, _P,
                                    _P]),
    "dgmi_level_restrict": (C.c_int, [C.c_int, C.c_int64, _P, _P, _P]),
    "dgmi_level_prolong_add": (C.c_int, [C.c_int, C.c_int64, _P, _P, _P]),
    "dgmi_pn_p1dg": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int64, _P, _P, _P]),
    "dgmi_set_param": (C.c_int, [_P, C.c_char_p, C.c_double]),
    "dgmi_autotune": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int32), _DP, _P]),
    "dgmi_dot": (C.c_int, [C.c_int, C.c_int64, _P, _P, _P, _P]),
    "dgmi_axpy": (C.c_int, [C.c_int, C.c_int64, C.c_double, _P, _P, _P]),
    "dgmi_upload": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_void_p), _I64P, _P]),
    "dgmi_dot_noalloc": (C.c_int, [C.c_int, C.c_int64, _P, _P, _P, _P]),
    "dgmi_dot_noalloc_max_n": (C.c_int64, []),
    "dgmi_last_error": (C.c_char_p, []),
    # host-only set-up
    "dgmi_host_element_tables": (C.c_int, [C.c_int, _DP, _DP, _DP, _DP, _DP, _DP, _I64P]),
    "dgmi_host_topology": (C.c_int, [_I64P, C.c_int64, C.c_int, _I64P, _I64P, _I64P]),
    "dgmi_host_p1cg_sparsity": (C.c_int, [_I64P, C.c_int64, C.c_int, C.c_int64, C.c_int64,
                                          _I64P, _I64P, _I64P]),
    "dgmi_host_color2": (C.c_int, [_I64P, _I64P, C.c_int64, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]),
    "dgmi_host_sfc_order": (C.c_int, [_DP, C.c_int64, C.c_int, _I64P]),
    "dgmi_host_nlevel_sfc": (C.c_int, [C.c_int64]),
    "dgmi_host_sfc_levels": (C.c_int, [_I64P, _I64P, _DP, _I64P, C.c_int64, C.c_int64,
                                       C.c_int64, _I64P, _I64P, _I64P, _I64P, _DP]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libdgmi.so and declare its signatures.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -m dg_nn_public_amd.build` (no fallback exists)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class DgmiError(RuntimeError):
    pass


def check(rc: int, what: str):
    if rc != 0:
        msg = load().dgmi_last_error().decode(errors="replace")
        raise DgmiError(f"{what} failed ({rc}): {msg}")
