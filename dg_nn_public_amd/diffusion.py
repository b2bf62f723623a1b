"""Reference-compatible host interface of the DG diffusion solve on MI355X.

The reference exposes module-level functions over torch tensors that read a global
state bag (z02/cmmn_data.py:19-64, ``config.sf_nd_nb``).  ``DGDiffusion`` holds that state
explicitly -- one library context per mesh -- and its methods keep the reference's
names, argument meaning, in-place mutation and return values:

  get_residual_or_smooth(r0, x_i, x_rhs, do_smooth=False) -> (r0, x_i)
                                          z02/volume_mf_diff.py:403-483
  get_bdiag_diag() -> (diagK, bdiagK)      z02/get_bdiag_diag.py:30-62
  _get_RAR_and_sfc_data_Um()               z02/volume_mf_diff.py:1642-1693
  _um_left_precond(x_i, x_rhs) -> x_i      z02/volume_mf_diff.py:1934-2029
  mg_on_P1CG(r0, variables_sfc, nlevel, nodes_per_level, cg_nonods) -> e
                                          z02/multigrid_linearelastic.py:112-172
  mg_smooth_one_level(level, e_i, b, variables_sfc) -> (e_i, rr_i)   :25-39
  mg_on_P1CG_prep(fina, cola, RARvalues, sparse_in)
      -> (sfc, variables_sfc, nlevel, nodes_per_level)               :175-283
  vel_pndg_to_p1dg_restrictor(x), vel_p1dg_to_pndg_prolongator(x)    :286-297
  _gmres_mg_solver(x_i, x_rhs, tol) -> (x_i, its)   z02/volume_mf_diff.py:1757-1841
  _mg_solver / _cg_mg_solver               z02/volume_mf_diff.py:1844-1931
  _get_rhs(x_rhs, f, u_bc, alpha_x_i) -> x_rhs      z02/volume_mf_diff.py:153-400
  solve_for_diff(x_i, f, u_bc, alpha_x_i, t=None) -> x_i   z02/volume_mf_diff.py:33-148
  get_l2_error(x_num, x_ana) -> (l2, linf) z02/volume_mf_diff.py:2032-2056

Module-level functions of the same names act on the space installed with
``set_space`` (the analogue of ``config.sf_nd_nb``).  All arithmetic runs in libdgmi.so;
torch tensors are device buffers and streams only.  Errors raise ``ValueError`` for bad
arguments and ``DgmiError`` for failures inside the library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check

ETA_E = 36.0        # z02/config.py:245
JAC_WEI = 2.0 / 3.0  # z02/config.py:108
GMRES_M = 30        # restart length of the workloads (the reference config uses 10, :149)
GMRES_ITS = 400     # z02/config.py:150
PRE_SMOOTH_ITS = 3   # z02/config.py:112
POST_SMOOTH_ITS = 3  # z02/config.py:113
LINEAR_SOLVERS = ("gmres-mg", "cg-mg", "mg")  # z02/config.py:95

_DTYPES = {torch.float64: _lib.DGMI_F64, torch.float32: _lib.DGMI_F32}
# inner precision of the iterative-refinement solve; bfloat16 means FP32 Krylov vectors and
# operator with bfloat16 inside the preconditioner: a bf16 copy of the smoother (vector
# kernels, exact block smoother) or bf16-input matrix-core applies otherwise
_INNER = {torch.float64: _lib.DGMI_F64, torch.float32: _lib.DGMI_F32,
          torch.bfloat16: _lib.DGMI_BF16}
F32_KERNELS = {"vector": 0, "mfma": 1, "mfma_bf16": 2}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class LevelCSR:
    """One SFC level operator on the device: CSR with int32 ``crow``/``col`` and values in
    the level's dtype (the reference keeps a torch sparse CSR tensor here,
    z02/multigrid_linearelastic.py:93-99)."""

    def __init__(self, crow, col, values, n):
        self.crow, self.col, self.values, self.n = crow, col, values, int(n)

    def to(self, dtype):
        return LevelCSR(self.crow, self.col, self.values.to(dtype), self.n)


class SfcData:
    """The reference's ``sf_nd_nb.sfc_data_Um``: SFC numbering (1-based, shape (n, 1) as
    the Fortran curve returns it), level operators, level count and sizes."""

    def __init__(self, space_filling_curve_numbering, variables_sfc, nlevel, nodes_per_level):
        self.space_filling_curve_numbering = space_filling_curve_numbering
        self.variables_sfc = variables_sfc
        self.nlevel = nlevel
        self.nodes_per_level = nodes_per_level


def _levels_to_device(levels, device):
    """[(rowptr, col, val)] numpy per level -> variables_sfc list of
    (LevelCSR, diag (1, n), n), the reference's (a_sfc_sparse, diagA, nonods)."""
    out = []
    for rp, col, val in levels:
        n = rp.size - 1
        rows = np.repeat(np.arange(n), np.diff(rp))
        diag = np.zeros(n)
        on = col == rows
        diag[rows[on]] = val[on]
        csr = LevelCSR(torch.as_tensor(rp.astype(np.int32), device=device),
                       torch.as_tensor(col.astype(np.int32), device=device),
                       torch.as_tensor(val, dtype=torch.float64, device=device), n)
        out.append((csr, torch.as_tensor(diag, device=device).view(1, n), n))
    return out


class _ContextLevels:
    """variables_sfc of a context's own hierarchy: built on first access from the exported
    RAR and SFC numbering (the same host builder the context used); ``mg_on_P1CG`` given
    this object runs the context's fused V-cycle."""

    def __init__(self, space):
        self._space = space
        self._levels = None

    def _get(self):
        if self._levels is None:
            sp = self._space
            if sp.ncomp != 1:
                raise ValueError("the level operators of a packed context are not exported")
            from . import host
            fina, cola, rar, sfc = sp.mg_export()
            self._levels = _levels_to_device(host.sfc_levels(fina, cola, rar, sfc), sp.device)
        return self._levels

    def __len__(self):
        return self._space.nlevel

    def __getitem__(self, i):
        return self._get()[i]

    def __iter__(self):
        return iter(self._get())


class DGDiffusion:
    """One P3 DG mesh, its SIPG operator and SFC multigrid hierarchy on one GPU."""

    def __init__(self, x_ref_in, cells, verts, kappa=None, bctype=None, eta_e: float = ETA_E,
                 blk_solver: str = "jacobian", jac_wei: float = JAC_WEI, device="cuda:0",
                 face_kappa: str = "mean", components=None, is_transient: bool = False,
                 dt: float | None = None, linear_solver: str = "gmres-mg",
                 alloc_limit: int = 0):
        """face_kappa: "mean" is the reference's interior-face diffusivity (kE + kN)/2;
        "harmonic" uses 2 kE kN/(kE + kN) (symmetric weighted interior penalty), which
        keeps the operator positive definite at high contrast.  ``components`` =
        (elem_off, vert_off) packs independent solves into this one mesh (see
        ``dg_nn_public_amd.batch``): each keeps its own multigrid hierarchy, Krylov space and
        convergence test, while every kernel launch covers all of them.
        ``is_transient`` / ``dt``: the reference's ``sf_nd_nb.isTransient`` and ``dt``
        (z02/config.py:33-36, 62-63): when transient, ``_get_rhs`` adds M alpha_x_i / dt.
        ``linear_solver``: the solver ``solve_for_diff`` runs (z02/config.py:95, :137-148).
        ``alloc_limit``: device bytes the context may hold (0: no cap), set right after
        creation (``set_alloc_limit``)."""
        if linear_solver not in LINEAR_SOLVERS:
            raise ValueError(f"linear_solver must be one of {LINEAR_SOLVERS}")
        if is_transient and not (dt is not None and dt > 0 and np.isfinite(dt)):
            raise ValueError("a transient space needs a positive, finite dt")
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("DGDiffusion runs on a GPU device")
        torch.cuda.set_device(self.device)
        # x_ref_in: one (nele, D, nloc) array, or with components one array per component
        parts = None
        if isinstance(x_ref_in, (list, tuple)):
            if components is None:
                raise ValueError("x_ref_in given in pieces needs components")
            parts = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in x_ref_in]
            if not parts or any(a.ndim != 3 or a.shape[1:] != parts[0].shape[1:] for a in parts):
                raise ValueError("every piece of x_ref_in must be (nele_c, D, nloc)")
            x = parts[0][:0]
            D, nele = parts[0].shape[1], sum(a.shape[0] for a in parts)
        else:
            x = np.ascontiguousarray(np.asarray(x_ref_in, dtype=np.float64))
            if x.ndim != 3:
                raise ValueError("x_ref_in must be (nele, D, nloc) with D in (2, 3)")
            D, nele = x.shape[1], x.shape[0]
        cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int64))
        verts = np.ascontiguousarray(np.asarray(verts, dtype=np.float64))
        if D not in (2, 3) or x.shape[2] != (10 if D == 2 else 20):
            raise ValueError("x_ref_in must be (nele, D, nloc) with D in (2, 3)")
        if cells.shape != (nele, D + 1) or verts.ndim != 2 or verts.shape[1] != D:
            raise ValueError("cells must be (nele, D+1) and verts (nverts, D)")
        kappa = np.ones(nele) if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        if kappa.shape != (nele,):
            raise ValueError("kappa must have one value per element")
        self._keep = [x, cells, verts, kappa] + (parts or [])
        bc = None
        if bctype is not None:
            bc = np.ascontiguousarray(np.asarray(bctype, dtype=np.int64).reshape(-1))
            if bc.size != nele * (D + 1):
                raise ValueError("bctype must have nele*nface entries")
            self._keep.append(bc)
        desc = _lib.MeshDesc(
            D=D, nele=nele, nverts=verts.shape[0],
            x_ref_in=x.ctypes.data_as(_lib._DP), cells=cells.ctypes.data_as(_lib._I64P),
            verts=verts.ctypes.data_as(_lib._DP), kappa=kappa.ctypes.data_as(_lib._DP),
            bctype=bc.ctypes.data_as(_lib._I64P) if bc is not None else None, eta_e=eta_e,
            face_kappa={"mean": 0, "harmonic": 1}[face_kappa], stream=_stream().value)
        ctx = C.c_void_p()
        self.ncomp = 1
        if components is None:
            check(self.lib.dgmi_ctx_create(C.byref(desc), C.byref(ctx)), "dgmi_ctx_create")
            self.elem_off = np.array([0, nele], dtype=np.int64)
            self.vert_off = np.array([0, verts.shape[0]], dtype=np.int64)
        else:
            eo = np.ascontiguousarray(components[0], dtype=np.int64)
            vo = np.ascontiguousarray(components[1], dtype=np.int64)
            if eo.ndim != 1 or eo.shape != vo.shape or eo.size < 2:
                raise ValueError("components must be two offset arrays of equal length >= 2")
            self.ncomp = eo.size - 1
            self.elem_off, self.vert_off = eo, vo
            if parts is None:
                check(self.lib.dgmi_ctx_create_batch(C.byref(desc), self.ncomp,
                                                     eo.ctypes.data_as(_lib._I64P),
                                                     vo.ctypes.data_as(_lib._I64P), C.byref(ctx)),
                      "dgmi_ctx_create_batch")
            else:
                if len(parts) != self.ncomp or any(a.shape[0] != eo[q + 1] - eo[q]
                                                   for q, a in enumerate(parts)):
                    raise ValueError("one piece of x_ref_in per component, of its element count")
                ptrs = (_lib._DP * len(parts))(*[a.ctypes.data_as(_lib._DP) for a in parts])
                check(self.lib.dgmi_ctx_create_batch_parts(C.byref(desc), self.ncomp,
                                                           eo.ctypes.data_as(_lib._I64P),
                                                           vo.ctypes.data_as(_lib._I64P), ptrs,
                                                           C.byref(ctx)),
                      "dgmi_ctx_create_batch_parts")
        self.ctx = ctx
        self._params = {}
        if alloc_limit:
            self.set_alloc_limit(alloc_limit)
        info = (C.c_int64 * 8)()
        check(self.lib.dgmi_ctx_info(ctx, info), "dgmi_ctx_info")
        self.D = D
        self.nele, self.nloc, self.ndof, self.nface = info[0], info[1], info[2], info[3]
        self.ngi, self.sngi, self.nverts = info[4], info[5], info[6]
        self.x_ref_in = x if parts is None else parts
        self.cells = cells
        self.verts = verts
        self.kappa = kappa
        self.blk_solver = blk_solver
        self.face_kappa = face_kappa
        self.jac_wei = jac_wei
        self.isTransient = bool(is_transient)
        self.dt = dt
        self.linear_solver = linear_solver
        self.bctype = bc if bc is not None else self.topology()[2]
        self.has_neumann = bool(np.any(self.bctype == 1))
        self._smoother_ready = False
        self.diagK = None
        self.bdiagK = None

    @classmethod
    def from_mesh(cls, mesh, kappa=None, **kw):
        """Build from a ``meshgen.Mesh``."""
        return cls(mesh.x_ref_in, mesh.cells, mesh.verts, kappa, **kw)

    def close(self, synced=False):
        """Give the context's device memory back to the library's block cache.  With
        ``synced`` the caller has finished all work it queued for this context (e.g. it
        synchronised its stream), and the device-wide synchronisation is skipped."""
        if getattr(self, "ctx", None):
            if synced:
                self.lib.dgmi_set_param(self.ctx, b"destroy_sync", 0.0)
            self.lib.dgmi_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- helpers
    def _vec(self, t, name, dtypes=(torch.float64, torch.float32)):
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{name} must be a torch tensor")
        if t.device != self.device and not (t.device.type == "cuda" and self.device.index is None):
            raise ValueError(f"{name} must live on {self.device}")
        if t.dtype not in dtypes:
            raise ValueError(f"{name} has dtype {t.dtype}; expected one of {dtypes}")
        if t.numel() != self.ndof or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous tensor of {self.ndof} values")
        return t

    def new_vector(self, dtype=torch.float64):
        return torch.zeros(self.ndof, dtype=dtype, device=self.device)

    def topology(self):
        n = self.nele * self.nface
        nbf, aln, bct = (np.empty(n, np.int64) for _ in range(3))
        check(self.lib.dgmi_topology(self.ctx, nbf.ctypes.data_as(_lib._I64P),
                                     aln.ctypes.data_as(_lib._I64P),
                                     bct.ctypes.data_as(_lib._I64P)), "dgmi_topology")
        return nbf, aln, bct

    def element_tables(self):
        D, nl, ng, sg, nf = self.D, self.nloc, self.ngi, self.sngi, self.nface
        arrs = dict(n=np.empty((nl, ng)), nlx=np.empty((D, nl, ng)), weight=np.empty(ng),
                    sn=np.empty((nf, nl, sg)), snlx=np.empty((nf, D, nl, sg)),
                    sweight=np.empty(sg), gi_align=np.empty((D, sg), np.int64))
        check(self.lib.dgmi_element_tables(
            self.ctx, *[arrs[k].ctypes.data_as(_lib._DP) for k in
                        ("n", "nlx", "weight", "sn", "snlx", "sweight")],
            arrs["gi_align"].ctypes.data_as(_lib._I64P)), "dgmi_element_tables")
        return arrs

    # ---------------------------------------------------------------- operator
    def set_f32_kernel(self, name):
        """Kernel family of FP32 applies and smooths: "vector" (FMA, the default), "mfma"
        (fp32 matrix cores) or "mfma_bf16" (bf16 inputs, fp32 accumulation)."""
        if name not in F32_KERNELS:
            raise ValueError(f"f32 kernel must be one of {sorted(F32_KERNELS)}")
        check(self.lib.dgmi_set_param(self.ctx, b"f32_kernel", float(F32_KERNELS[name])),
              "dgmi_set_param")

    def set_param(self, name, value):
        """Set a solver parameter of the context (``dgmi_set_param``: e.g. "coarse_exact",
        "smooth_offdiag", "single_pass", "staged", "smoother_half", "f32_smooth_its"; see
        include/dgmi.h)."""
        check(self.lib.dgmi_set_param(self.ctx, str(name).encode(), float(value)),
              "dgmi_set_param")
        self._params[str(name)] = float(value)

    @property
    def smoother_half(self):
        """Whether the FP32 block-staged sweeps read the FP16 copy of the smoother (the
        library's default: in 3D)."""
        return bool(self._params.get("smoother_half", 1.0 if self.D == 3 else 0.0))

    def set_alloc_limit(self, nbytes):
        """Cap the device bytes this context may hold (0: no cap).  An allocation that would
        pass the cap fails and the call that needed it raises ``DgmiError`` without
        launching work."""
        check(self.lib.dgmi_set_param(self.ctx, b"alloc_limit_bytes", float(nbytes)),
              "dgmi_set_param")

    SETUP_PHASES = ("topology", "geometry", "upload", "smoother", "graph", "probe", "levels",
                    "layout")

    def setup_timing(self):
        """Host wall-clock seconds of this context's set-up phases (dgmi_setup_timing)."""
        out = (C.c_double * 8)()
        check(self.lib.dgmi_setup_timing(self.ctx, out), "dgmi_setup_timing")
        return dict(zip(self.SETUP_PHASES, list(out)))

    def device_bytes(self):
        """Device bytes this context holds."""
        info = (C.c_int64 * 8)()
        check(self.lib.dgmi_ctx_info(self.ctx, info), "dgmi_ctx_info")
        return int(info[7])

    def autotune(self, reps=5):
        """Time the FP32 apply kernels (vector; matrix cores with 32/64/128 elements per
        work-group, fp32 and bf16 inputs) on this mesh and keep the fastest.  Returns the
        choice and the timings in microseconds."""
        ch = (C.c_int32 * 3)()
        tm = (C.c_double * 7)()
        check(self.lib.dgmi_autotune(self.ctx, reps, ch, tm, _stream()), "dgmi_autotune")
        names = ["vector", "mfma_w2", "mfma_w4", "mfma_w8", "bf16_w2", "bf16_w4", "bf16_w8"]
        return {"f32_kernel": ["vector", "mfma"][ch[0]], "mfma_waves": ch[1],
                "bf16_waves": ch[2], "us": dict(zip(names, list(tm)))}

    def apply(self, x, b=None, out=None, bf16=False):
        """out = b - A x (b None: out = -A x) in the precision of x.  ``bf16`` (FP32 x only)
        rounds the contraction inputs to bf16 and accumulates in fp32."""
        self._vec(x, "x")
        dt = _DTYPES[x.dtype]
        if bf16:
            if x.dtype != torch.float32:
                raise ValueError("the bf16-input apply takes float32 vectors")
            dt = _lib.DGMI_BF16
        if b is not None:
            self._vec(b, "b", (x.dtype,))
        out = torch.empty_like(x) if out is None else self._vec(out, "out", (x.dtype,))
        check(self.lib.dgmi_apply(self.ctx, dt, _ptr(x), _ptr(b), _ptr(out), _stream()),
              "dgmi_apply")
        return out

    def residual_dd(self, x, b, out=None):
        """out = b - A x (FP64 vectors) evaluated in double-double arithmetic and rounded
        once to FP64: the residual without the rounding noise of the FP64 evaluation,
        which can reach 1e-10 ||b|| at strong diffusivity contrast on fine meshes."""
        self._vec(x, "x", (torch.float64,))
        self._vec(b, "b", (torch.float64,))
        out = torch.empty_like(x) if out is None else self._vec(out, "out", (torch.float64,))
        check(self.lib.dgmi_residual_dd(self.ctx, _ptr(x), _ptr(b), _ptr(out), _stream()),
              "dgmi_residual_dd")
        return out

    def set_residual_dd(self, mode):
        """True residuals of the solvers: 0 FP64, 1 double-double near convergence or on
        stagnation (default), 2 always double-double."""
        check(self.lib.dgmi_set_param(self.ctx, b"residual_dd", float(mode)), "dgmi_set_param")

    def get_residual_or_smooth(self, r0, x_i, x_rhs, do_smooth=False):
        """r0 += x_rhs - A x_i; with do_smooth also x_i += omega c, c from the block
        smoother applied to r0.  Both tensors are mutated and returned flattened
        (z02/volume_mf_diff.py:403-483)."""
        r0 = r0.view(-1)
        x_i = x_i.view(-1)
        self._vec(r0, "r0")
        self._vec(x_i, "x_i", (r0.dtype,))
        dt = _DTYPES[r0.dtype]
        if not (isinstance(x_rhs, (int, float)) and x_rhs == 0):
            x_rhs = self._vec(x_rhs.view(-1), "x_rhs", (r0.dtype,))
            check(self.lib.dgmi_axpy(dt, self.ndof, 1.0, _ptr(x_rhs), _ptr(r0), _stream()),
                  "dgmi_axpy")
        if do_smooth:
            self._ensure_smoother()
            check(self.lib.dgmi_smooth(self.ctx, dt, _ptr(x_i), _ptr(r0), _ptr(r0), _stream()),
                  "dgmi_smooth")
        else:
            check(self.lib.dgmi_apply(self.ctx, dt, _ptr(x_i), _ptr(r0), _ptr(r0), _stream()),
                  "dgmi_apply")
        return r0, x_i

    def smooth_n(self, x, b, n):
        """n smoothing steps x += omega S (b - A x) in place (the reference's smoothing loop,
        z02/volume_mf_diff.py:1947-1951, without the residual output)."""
        x = self._vec(x.view(-1), "x")
        b = self._vec(b.view(-1), "b", (x.dtype,))
        self._ensure_smoother()
        check(self.lib.dgmi_smooth_n(self.ctx, _DTYPES[x.dtype], _ptr(x), _ptr(b), int(n),
                                     _stream()), "dgmi_smooth_n")
        return x

    def get_bdiag_diag(self):
        """Diagonal blocks and diagonal of A in FP64 (z02/get_bdiag_diag.py:30-62)."""
        bd = torch.empty((self.nele, self.nloc, self.nloc), dtype=torch.float64, device=self.device)
        dg = torch.empty((self.nele, self.nloc), dtype=torch.float64, device=self.device)
        check(self.lib.dgmi_bdiag(self.ctx, _ptr(bd), _ptr(dg), _stream()), "dgmi_bdiag")
        self.bdiagK, self.diagK = bd, dg
        return dg, bd

    def _ensure_smoother(self):
        if not self._smoother_ready:
            v = {"jacobian": _lib.DGMI_SMOOTH_JACOBI3, "direct": _lib.DGMI_SMOOTH_DIRECT}[self.blk_solver]
            check(self.lib.dgmi_smoother_setup(self.ctx, v, self.jac_wei, _stream()),
                  "dgmi_smoother_setup")
            self._smoother_ready = True

    # ---------------------------------------------------------------- multigrid
    def _get_RAR_and_sfc_data_Um(self):
        """P1CG operator RAR + SFC level operators (z02/volume_mf_diff.py:1642-1693); sets
        ``sfc_data_Um`` as the reference does.  RAR is assembled from the continuous P1
        element matrices unless ``set_param("rar_probing", 1)`` asks for the reference's
        coloured probing (z02/volume_mf_diff.py:1696-1754); the two agree to rounding."""
        self._ensure_smoother()
        check(self.lib.dgmi_mg_setup(self.ctx, _stream()), "dgmi_mg_setup")
        info = (C.c_int64 * 128)()
        check(self.lib.dgmi_mg_info(self.ctx, info, 128), "dgmi_mg_info")
        self.cg_nonods, self.ncola, self.ncolor, self.nlevel = info[0], info[1], info[2], info[3]
        self.nodes_per_level = [info[4 + i] for i in range(self.nlevel)]
        # FP32 cycles: level solved exactly (-1: none)
        self.coarse_exact_level = int(info[4 + self.nlevel])
        self._mg_ready = True
        self._sfc = None
        self.sfc_data_Um = SfcData(None, _ContextLevels(self), self.nlevel,
                                   list(self.nodes_per_level))
        if self.ncomp == 1:
            self.sfc_data_Um.space_filling_curve_numbering = self.mg_export()[3].reshape(-1, 1)
        return 0

    def mg_export(self):
        """Host copies of (fina, cola, RAR values, sfc numbering)."""
        fina = np.empty(self.cg_nonods + 1, np.int64)
        cola = np.empty(self.ncola, np.int64)
        rar = np.empty(self.ncola)
        sfc = np.empty(self.cg_nonods, np.int64)
        check(self.lib.dgmi_mg_export(self.ctx, fina.ctypes.data_as(_lib._I64P),
                                      cola.ctypes.data_as(_lib._I64P), rar.ctypes.data_as(_lib._DP),
                                      sfc.ctypes.data_as(_lib._I64P)), "dgmi_mg_export")
        return fina, cola, rar, sfc

    def _need_mg(self):
        if not getattr(self, "_mg_ready", False):
            self._get_RAR_and_sfc_data_Um()

    def _cg_vec(self, t, name, dtypes=(torch.float64, torch.float32), n=None):
        """A P1CG / SFC-level vector: device tensor on this space's device, contiguous, n
        values (cg_nonods by default), one of ``dtypes``."""
        n = self.cg_nonods if n is None else n
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{name} must be a torch tensor")
        if t.device.type != "cuda" or (self.device.index is not None and t.device != self.device):
            raise ValueError(f"{name} must live on {self.device}")
        if t.dtype not in dtypes:
            raise ValueError(f"{name} has dtype {t.dtype}; expected one of {dtypes}")
        if t.numel() != n or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous tensor of {n} values")
        return t

    def restrict_to_p1cg(self, r):
        """P3DG -> P1CG in SFC order (I_cf P^T, permuted)."""
        self._need_mg()
        self._vec(r, "r")
        out = torch.empty(self.cg_nonods, dtype=r.dtype, device=self.device)
        check(self.lib.dgmi_restrict(self.ctx, _DTYPES[r.dtype], _ptr(r), _ptr(out), _stream()),
              "dgmi_restrict")
        return out

    def prolong_add_from_p1cg(self, e_sfc, x):
        """x += P I_fc e (e in SFC order)."""
        self._need_mg()
        self._vec(x, "x")
        self._cg_vec(e_sfc, "e_sfc", (x.dtype,))
        check(self.lib.dgmi_prolong_add(self.ctx, _DTYPES[x.dtype], _ptr(e_sfc), _ptr(x),
                                        _stream()), "dgmi_prolong_add")
        return x

    def vel_pndg_to_p1dg_restrictor(self, x):
        """y[e, k] = sum_n P[n, k] x[e, n]: P3DG -> P1DG per element
        (z02/multigrid_linearelastic.py:286-290)."""
        x = self._vec(x.reshape(-1), "x")
        y = torch.empty(self.nele * (self.D + 1), dtype=x.dtype, device=self.device)
        check(self.lib.dgmi_pn_p1dg(self.D, _DTYPES[x.dtype], 0, self.nele, _ptr(x), _ptr(y),
                                    _stream()), "dgmi_pn_p1dg")
        return y

    def vel_p1dg_to_pndg_prolongator(self, x):
        """y[e, n] = sum_k P[n, k] x[e, k]: P1DG -> P3DG per element (:293-297)."""
        x = self._cg_vec(x.reshape(-1), "x", n=self.nele * (self.D + 1))
        y = torch.empty(self.ndof, dtype=x.dtype, device=self.device)
        check(self.lib.dgmi_pn_p1dg(self.D, _DTYPES[x.dtype], 1, self.nele, _ptr(x), _ptr(y),
                                    _stream()), "dgmi_pn_p1dg")
        return y

    def mg_smooth_one_level(self, level, e_i, b, variables_sfc=None):
        """rr = b - A_l e; e += jac_wei rr / diag(A_l) on SFC level ``level``
        (z02/multigrid_linearelastic.py:25-39).  ``variables_sfc`` as returned by
        ``mg_on_P1CG_prep`` (default: this space's own levels).  Returns (e_i, rr_i), e_i
        updated in place."""
        if variables_sfc is None:
            self._need_mg()
            variables_sfc = self.sfc_data_Um.variables_sfc
        a, diag, n = variables_sfc[level]
        e = e_i.reshape(-1)
        bb = b.reshape(-1)
        self._cg_vec(e, "e_i", n=n)
        self._cg_vec(bb, "b", (e.dtype,), n=n)
        if a.values.dtype != e.dtype:
            a = a.to(e.dtype)
            diag = diag.to(e.dtype)
        rr = torch.empty_like(e)
        check(self.lib.dgmi_level_smooth(_DTYPES[e.dtype], n, _ptr(a.crow), _ptr(a.col),
                                         _ptr(a.values), _ptr(diag), self.jac_wei, _ptr(e),
                                         _ptr(bb), _ptr(rr), _stream()), "dgmi_level_smooth")
        return e_i, rr

    def mg_on_P1CG_prep(self, fina, cola, RARvalues, sparse_in=None):
        """SFC numbering and level operators of a P1CG matrix given by (fina, cola,
        RARvalues) (z02/multigrid_linearelastic.py:175-283).  The numbering is this build's
        Hilbert order of the P1 vertices (``sparse_in`` may supply ``sfc``, 1-based, or
        ``verts``; default: this space's vertices).  Returns (sfc (n, 1), variables_sfc,
        nlevel, nodes_per_level) with variables_sfc[l] = (LevelCSR, diag (1, n_l), n_l)."""
        from . import host
        to_np = lambda a: a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)  # noqa: E731
        fina = np.ascontiguousarray(to_np(fina), dtype=np.int64)
        cola = np.ascontiguousarray(to_np(cola), dtype=np.int64)
        vals = np.ascontiguousarray(to_np(RARvalues), dtype=np.float64).reshape(-1)
        if vals.size != cola.size or fina[-1] != cola.size:
            raise ValueError("fina, cola and RARvalues do not describe one CSR matrix")
        n = fina.size - 1
        src = self if sparse_in is None else sparse_in
        sfc = getattr(src, "sfc", None)
        if sfc is None:
            verts = getattr(src, "verts", None)
            if verts is None:
                raise ValueError("sparse_in must provide sfc or verts")
            sfc = host.sfc_order(np.asarray(verts, dtype=np.float64))
        sfc = np.ascontiguousarray(np.asarray(sfc, dtype=np.int64).reshape(-1))
        if sfc.size != n:
            raise ValueError("the SFC numbering must have one entry per P1CG node")
        levels = host.sfc_levels(fina, cola, vals, sfc)
        variables_sfc = _levels_to_device(levels, self.device)
        nodes = [v[2] for v in variables_sfc]
        return sfc.reshape(-1, 1), variables_sfc, len(nodes), nodes

    def mg_on_P1CG(self, r0, variables_sfc=None, nlevel=None, nodes_per_level=None,
                   cg_nonods=None):
        """One SFC V-cycle, r and e in SFC order (z02/multigrid_linearelastic.py:112-172).
        With this space's own levels (the default, or ``sfc_data_Um.variables_sfc``) the
        cycle runs fused in the library; with levels from ``mg_on_P1CG_prep`` it runs the
        reference's steps one call each: pre-smoothing, the residual sweep (which also
        updates e, as the reference's call does), pairwise restriction of r and e, the
        one-node direct solve, nearest upsampling and post-smoothing."""
        own = variables_sfc is None or isinstance(variables_sfc, _ContextLevels)
        if own:
            self._need_mg()
            r0 = self._cg_vec(r0.reshape(-1), "r0")
            e = torch.empty_like(r0)
            check(self.lib.dgmi_sfc_vcycle(self.ctx, _DTYPES[r0.dtype], _ptr(r0), _ptr(e),
                                           _stream()), "dgmi_sfc_vcycle")
            return e
        nlevel = len(variables_sfc) if nlevel is None else int(nlevel)
        nodes = [v[2] for v in variables_sfc] if nodes_per_level is None else list(nodes_per_level)
        n0 = nodes[0] if cg_nonods is None else int(cg_nonods)
        r0 = self._cg_vec(r0.reshape(-1), "r0", n=n0)
        dt = _DTYPES[r0.dtype]
        start = nlevel - 1
        r_s = [r0]
        e_s = [torch.zeros_like(r0)]
        for level in range(start):
            for _ in range(PRE_SMOOTH_ITS):
                self.mg_smooth_one_level(level, e_s[level], r_s[level], variables_sfc)
            _, rr = self.mg_smooth_one_level(level, e_s[level], r_s[level], variables_sfc)
            rc = torch.empty(nodes[level + 1], dtype=r0.dtype, device=self.device)
            ec = torch.empty_like(rc)
            check(self.lib.dgmi_level_restrict(dt, nodes[level], _ptr(rr), _ptr(rc), _stream()),
                  "dgmi_level_restrict")
            check(self.lib.dgmi_level_restrict(dt, nodes[level], _ptr(e_s[level]), _ptr(ec),
                                               _stream()), "dgmi_level_restrict")
            r_s.append(rc)
            e_s.append(ec)
        # direct solve on the last level: one node (calculate_nlevel_sfc goes down to 1)
        if nodes[start] != 1:
            raise ValueError("the last SFC level must hold one node")
        a, diag, _ = variables_sfc[start]
        if a.values.dtype != r0.dtype:
            a, diag = a.to(r0.dtype), diag.to(r0.dtype)
        e_s[start] = torch.zeros(1, dtype=r0.dtype, device=self.device)
        scratch = torch.empty_like(e_s[start])
        check(self.lib.dgmi_level_smooth(dt, 1, _ptr(a.crow), _ptr(a.col), _ptr(a.values),
                                         _ptr(diag), 1.0, _ptr(e_s[start]), _ptr(r_s[start]),
                                         _ptr(scratch), _stream()), "dgmi_level_smooth")
        for level in range(start - 1, -1, -1):
            check(self.lib.dgmi_level_prolong_add(dt, nodes[level], _ptr(e_s[level + 1]),
                                                  _ptr(e_s[level]), _stream()),
                  "dgmi_level_prolong_add")
            for _ in range(POST_SMOOTH_ITS):
                self.mg_smooth_one_level(level, e_s[level], r_s[level], variables_sfc)
        return e_s[0]

    def _um_left_precond(self, x_i, x_rhs):
        """x_i = M^-1 x_rhs, one V-cycle from x_i = 0 (every caller in the reference passes a
        zeroed x_i, z02/volume_mf_diff.py:1785-1786, 1807-1808)."""
        self._need_mg()
        x_i = self._vec(x_i.view(-1), "x_i")
        x_rhs = self._vec(x_rhs.view(-1), "x_rhs", (x_i.dtype,))
        check(self.lib.dgmi_precond(self.ctx, _DTYPES[x_i.dtype], _ptr(x_rhs), _ptr(x_i),
                                    _stream()), "dgmi_precond")
        return x_i

    # ---------------------------------------------------------------- solvers
    def _gmres_mg_solver(self, x_i, x_rhs, tol=None, rtol=None, m=GMRES_M, max_its=GMRES_ITS,
                         inner=torch.float64, orth="cgs2"):
        """Left-preconditioned GMRES(m).  The reference stops on the absolute residual
        ``tol``; pass ``rtol`` for ||b - A x|| <= rtol ||b||.  ``inner`` float32 runs the
        Krylov loop in FP32 inside FP64 iterative refinement; bfloat16 does the same with
        bfloat16 inside the preconditioner (a bf16 copy of the smoother with the vector
        kernels and the exact block smoother, bf16-input matrix-core applies otherwise).  ``orth`` "mgs" is the
        reference's modified Gram-Schmidt, "cgs2" fused classical Gram-Schmidt with
        re-orthogonalisation.  Returns (x_i, its) and keeps the statistics in
        ``self.last_stats`` (for a packed batch a list of per-component dicts, and ``its``
        is the largest count)."""
        self._need_mg()
        if orth not in ("mgs", "cgs2"):
            raise ValueError("orth must be 'mgs' or 'cgs2'")
        check(self.lib.dgmi_set_param(self.ctx, b"gmres_orth", 0.0 if orth == "mgs" else 1.0),
              "dgmi_set_param")
        x_i = self._vec(x_i.view(-1), "x_i", (torch.float64,))
        b = self._vec(x_rhs.view(-1), "x_rhs", (torch.float64,))
        if rtol is None:
            if tol is None:
                raise ValueError("give tol (absolute) or rtol (relative)")
            bn = float(torch.linalg.vector_norm(b))
            rtol = tol / bn if bn > 0 else 0.0
        stats = (C.c_double * (4 * self.ncomp))()
        check(self.lib.dgmi_gmres_solve(self.ctx, _INNER[inner], _ptr(b), _ptr(x_i), rtol, m,
                                        max_its, stats, _stream()), "dgmi_gmres_solve")
        per = [dict(its=int(stats[4 * q]), relres=stats[4 * q + 1], outer=int(stats[4 * q + 2]),
                    cycles=int(stats[4 * q + 2]), f64_cycles=int(stats[4 * q + 3]))
               for q in range(self.ncomp)]
        self.last_stats = per[0] if self.ncomp == 1 else per
        return x_i, max(p["its"] for p in per)

    def _mg_solver(self, x_i, x_rhs, tol=None, rtol=None, max_its=GMRES_ITS):
        """Stand-alone V-cycle iteration (z02/volume_mf_diff.py:1844-1873): x += M^-1 r
        until ||r|| <= tol (absolute, the reference's test) or rtol ||b||."""
        self._need_mg()
        x_i = self._vec(x_i.view(-1), "x_i", (torch.float64,))
        b = self._vec(x_rhs.view(-1), "x_rhs", (torch.float64,))
        if tol is None and rtol is None:
            raise ValueError("give tol (absolute) or rtol (relative)")
        bn = float(torch.linalg.vector_norm(b))
        thr = tol if rtol is None else rtol * bn
        r = self.apply(x_i, b)
        e = torch.zeros_like(x_i)
        its = 0
        while float(torch.linalg.vector_norm(r)) > thr and its < max_its:
            self._um_left_precond(e, r)
            check(self.lib.dgmi_axpy(_lib.DGMI_F64, self.ndof, 1.0, _ptr(e), _ptr(x_i), _stream()),
                  "dgmi_axpy")
            self.apply(x_i, b, out=r)
            its += 1
        rn = float(torch.linalg.vector_norm(r))
        self.last_stats = dict(its=its, relres=rn / bn if bn > 0 else 0.0, outer=its, cycles=its)
        return x_i, its

    def _cg_mg_solver(self, x_i, x_rhs, tol=None, rtol=None, max_its=GMRES_ITS):
        """MG-preconditioned conjugate gradients (z02/volume_mf_diff.py:1875-1931), on the
        device in FP64.  The reference stops on the absolute norm of the updated residual
        ``tol``; pass ``rtol`` for ||r|| <= rtol ||b||.  Returns (x_i, its); the final true
        relative residual is in ``last_stats`` (a list per component for a packed space)."""
        self._need_mg()
        x_i = self._vec(x_i.view(-1), "x_i", (torch.float64,))
        b = self._vec(x_rhs.view(-1), "x_rhs", (torch.float64,))
        if tol is None and rtol is None:
            raise ValueError("give tol (absolute) or rtol (relative)")
        abs_tol = -1.0 if rtol is not None else float(tol)
        stats = (C.c_double * (4 * self.ncomp))()
        check(self.lib.dgmi_cg_solve(self.ctx, _ptr(b), _ptr(x_i),
                                     0.0 if rtol is None else float(rtol), abs_tol, int(max_its),
                                     stats, _stream()), "dgmi_cg_solve")
        per = [dict(its=int(stats[4 * q]), relres=stats[4 * q + 1], outer=1, cycles=1)
               for q in range(self.ncomp)]
        self.last_stats = per[0] if self.ncomp == 1 else per
        return x_i, max(p["its"] for p in per)

    def _dof_data(self, a, name):
        if isinstance(a, torch.Tensor) and a.device.type == "cuda":
            return self._vec(a.to(torch.float64).reshape(-1).contiguous(), name,
                             (torch.float64,))
        return self._upload([a], name)

    def _upload(self, pieces, name):
        """A new device vector of host pieces (arrays or CPU tensors, taken as float64)
        laid end to end, copied through the library's pinned staging buffers."""
        arrs = [np.ascontiguousarray(np.asarray(p, dtype=np.float64)).reshape(-1)
                for p in pieces]
        if sum(a.size for a in arrs) != self.ndof:
            raise ValueError(f"{name} must hold {self.ndof} values")
        out = torch.empty(self.ndof, dtype=torch.float64, device=self.device)
        src = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        nb = (C.c_int64 * len(arrs))(*[a.nbytes for a in arrs])
        check(self.lib.dgmi_upload(self.ctx, _ptr(out), len(arrs), src, nb, _stream()),
              "dgmi_upload")
        return out

    def _get_rhs(self, x_rhs, f, u_bc, alpha_x_i=None):
        """x_rhs += M f + Dirichlet terms + Neumann terms (+ M alpha_x_i / dt when the
        space is transient), as z02/volume_mf_diff.py:153-400 accumulates into x_rhs.

        ``u_bc`` is the reference's list of per-boundary-label nodal data (z02/bc_f.py:892):
        u_bc[0] Dirichlet values, u_bc[1] the Neumann flux on faces of glb_bcface_type 1;
        a single array is taken as the Dirichlet data of a mesh without Neumann faces.
        ``alpha_x_i`` is used only when the space is transient (``is_transient``), as the
        reference tests ``sf_nd_nb.isTransient`` (:230)."""
        x_rhs = self._vec(x_rhs.view(-1), "x_rhs", (torch.float64,))
        f = self._dof_data(f, "f")
        if isinstance(u_bc, (list, tuple)):
            if len(u_bc) == 0:
                raise ValueError("u_bc holds no boundary data")
            ud = self._dof_data(u_bc[0], "u_bc[0]")
            un = self._dof_data(u_bc[1], "u_bc[1]") if len(u_bc) > 1 else None
        else:
            ud, un = self._dof_data(u_bc, "u_bc"), None
        if self.has_neumann and un is None:
            raise ValueError("the mesh has Neumann faces (glb_bcface_type 1): pass u_bc as "
                             "[dirichlet, neumann]")
        al = None
        if self.isTransient and alpha_x_i is not None:
            al = self._dof_data(alpha_x_i, "alpha_x_i")
        tmp = torch.empty_like(x_rhs)
        check(self.lib.dgmi_rhs(self.ctx, _ptr(f), _ptr(ud), _ptr(un), _ptr(al),
                                float(self.dt) if al is not None else 0.0, _ptr(tmp), _stream()),
              "dgmi_rhs")
        check(self.lib.dgmi_axpy(_lib.DGMI_F64, self.ndof, 1.0, _ptr(tmp), _ptr(x_rhs),
                                 _stream()), "dgmi_axpy")
        return x_rhs

    def solve_for_diff(self, x_i, f, u_bc, alpha_x_i=None, t=None, rtol=1e-10, tol=None,
                       inner=torch.float64, m=GMRES_M, max_its=GMRES_ITS):
        """RHS, multigrid set-up and the solve with ``linear_solver`` (z02/volume_mf_diff.py:
        33-148).  ``tol`` (absolute) replaces ``rtol`` when given; ``inner`` applies to
        "gmres-mg"."""
        self._need_mg()
        shape = x_i.shape
        b = self._get_rhs(self.new_vector(), f, u_bc, alpha_x_i)
        kw = dict(tol=tol) if tol is not None else dict(rtol=rtol)
        if self.linear_solver == "gmres-mg":
            x_i, _ = self._gmres_mg_solver(x_i.view(-1), b, inner=inner, m=m, max_its=max_its, **kw)
        elif self.linear_solver == "cg-mg":
            x_i, _ = self._cg_mg_solver(x_i.view(-1), b, max_its=max_its, **kw)
        else:
            x_i, _ = self._mg_solver(x_i.view(-1), b, max_its=max_its, **kw)
        return x_i.view(shape)

    def get_l2_error(self, x_num, x_ana):
        """(l2, linf) with the reference's definition (z02/volume_mf_diff.py:2032-2056):
        l2^2 = sum_e sum_n sum_g (n[n,g] (x_num - x_ana)[e,n])^2 detwei[e,g], computed on
        the device (FP64 sums)."""
        x_num = self._vec(x_num.view(-1), "x_num")
        x_ana = self._vec(x_ana.view(-1), "x_ana", (x_num.dtype,))
        out = (C.c_double * 2)()
        check(self.lib.dgmi_l2_error(self.ctx, _DTYPES[x_num.dtype], _ptr(x_num), _ptr(x_ana),
                                     out, _stream()), "dgmi_l2_error")
        return float(out[0]), float(out[1])


# ---------------------------------------------------------------- module-level API
_space: DGDiffusion | None = None


def set_space(space: DGDiffusion):
    """Install the active space (the analogue of ``config.sf_nd_nb``)."""
    global _space
    _space = space


def _active():
    if _space is None:
        raise ValueError("no active space: call set_space(DGDiffusion(...)) first")
    return _space


def get_residual_or_smooth(r0, x_i, x_rhs, do_smooth=False):
    return _active().get_residual_or_smooth(r0, x_i, x_rhs, do_smooth)


def get_bdiag_diag():
    return _active().get_bdiag_diag()


def _get_RAR_and_sfc_data_Um():
    return _active()._get_RAR_and_sfc_data_Um()


def _um_left_precond(x_i, x_rhs):
    return _active()._um_left_precond(x_i, x_rhs)


def release_cached_memory():
    """Free the device blocks the library keeps for reuse after contexts are closed."""
    check(_lib.load().dgmi_release_cached_memory(), "dgmi_release_cached_memory")


def _gmres_mg_solver(x_i, x_rhs, tol):
    return _active()._gmres_mg_solver(x_i, x_rhs, tol)


def _mg_solver(x_i, x_rhs, tol):
    return _active()._mg_solver(x_i, x_rhs, tol)


def _cg_mg_solver(x_i, x_rhs, tol):
    return _active()._cg_mg_solver(x_i, x_rhs, tol)


def _get_rhs(x_rhs, f, u_bc, alpha_x_i):
    return _active()._get_rhs(x_rhs, f, u_bc, alpha_x_i)


def mg_on_P1CG(r0, variables_sfc, nlevel, nodes_per_level, cg_nonods):
    return _active().mg_on_P1CG(r0, variables_sfc, nlevel, nodes_per_level, cg_nonods)


def mg_smooth_one_level(level, e_i, b, variables_sfc):
    return _active().mg_smooth_one_level(level, e_i, b, variables_sfc)


def mg_on_P1CG_prep(fina, cola, RARvalues, sparse_in):
    return _active().mg_on_P1CG_prep(fina, cola, RARvalues, sparse_in)


def vel_pndg_to_p1dg_restrictor(x):
    return _active().vel_pndg_to_p1dg_restrictor(x)


def vel_p1dg_to_pndg_prolongator(x):
    return _active().vel_p1dg_to_pndg_prolongator(x)


def solve_for_diff(x_i, f, u_bc, alpha_x_i=None, t=None):
    return _active().solve_for_diff(x_i, f, u_bc, alpha_x_i, t)


def get_l2_error(x_num, x_ana):
    return _active().get_l2_error(x_num, x_ana)
