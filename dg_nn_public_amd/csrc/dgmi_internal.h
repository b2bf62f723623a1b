// Internal declarations shared by the libdgmi translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "setup_host.h"

namespace dgmi {

void set_error(const std::string& msg);

#define DGMI_HIP(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::dgmi::set_error(std::string(#call) + ": " + hipGetErrorString(e_));     \
      return -3;                                                                \
    }                                                                           \
  } while (0)

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

// Per-precision device copies of the element tables and geometry.
// Table buffer layout (kernel-friendly, contiguous per quadrature point):
//   vtab[g][d*nloc + n]          = dphi_n/dxi_d at volume point g       (ngi*D*nloc)
//   wgt[g]                                                              (ngi)
//   ftab[f][g][c*nloc + n]       c = 0: phi_n, c = 1+d: dphi_n/dxi_d    (nface*sngi*(1+D)*nloc)
//   sw[g]                                                               (sngi)
template <typename T> struct DevGeom {
  T* tab = nullptr;
  size_t off_vtab = 0, off_wgt = 0, off_ftab = 0, off_sw = 0;
  T* G = nullptr;      // (ng, nele)
  T* cf = nullptr;     // (nface*D, nele)
  T* wf = nullptr;     // (nface, nele)
  T* pen = nullptr;    // (nface, nele)
  T* S = nullptr;      // smoother (nloc*nloc, nele), row n col m at (n*nloc+m)*nele + e
  typename Vec2<T>::type* trace = nullptr;  // (nface*sngi, nele)
};

struct SfcLevelsDev {
  int nlevel = 0;
  std::vector<int64_t> n, nnz;
  std::vector<int32_t*> rowptr, col;
  std::vector<double*> val64, diag64;
  std::vector<float*> val32, diag32;
  // per-level work vectors (both precisions share the byte buffers, sized for FP64)
  std::vector<void*> r, e, e2;
  // device arrays of LevelDesc<double> / LevelDesc<float> for the single-work-group kernel
  void* desc64 = nullptr;
  void* desc32 = nullptr;
};

struct MGDev {
  bool ready = false;
  int64_t cg_nonods = 0, ncola = 0;
  int ncolor = 0;
  std::vector<int64_t> fina, cola, sfc;
  std::vector<double> rar;
  // restriction gather lists in SFC order: for SFC position p, entries
  // [rptr[p], rptr[p+1]) hold p1dg indices e*(D+1)+k; weight 1/mult
  int32_t* rptr = nullptr;
  int32_t* rent = nullptr;
  double* rw64 = nullptr;
  float* rw32 = nullptr;
  // prolongation: per p1dg index e*(D+1)+k, the SFC position of its vertex
  int32_t* cellpos = nullptr;
  SfcLevelsDev lv;
  int small_level = 0;  // first level handled by the single-workgroup kernel
  bool small_lds = true;  // stage the small levels in LDS
  void* r_sfc = nullptr;  // (cg_nonods) fp64-sized scratch
  void* e_sfc = nullptr;
};

namespace kry {
template <typename T>
int axpy_dev(int64_t n, const double* a, double sgn, const T* x, T* y, hipStream_t s);
template <typename T>
int scale_dev(int64_t n, const double* a, const T* x, T* y, hipStream_t s);
template <typename Tin, typename Tout>
int cast(int64_t n, const Tin* x, Tout* y, double scale, hipStream_t s);
template <typename Tin>
int add_cast(int64_t n, const Tin* x, double* y, const double* scale, hipStream_t s);
template <typename T>
int fill(int64_t n, T* y, T v, hipStream_t s);
template <typename T>
int combine(int64_t n, int k, const T* V, int64_t ldv, const double* y, T* x, hipStream_t s);
int givens(int j, int ldh, double* H, double* cs, double* sn, double* g, double* res,
           hipStream_t s);
int backsolve(int k, int ldh, const double* H, const double* g, double* y, hipStream_t s);
int store_h(const double* src, double* dst, int mode, hipStream_t s);
int set_g(double* g, int len, const double* beta, hipStream_t s);
template <typename T>
int cgs_pass(int64_t n, int k, const T* V, int64_t ldv, T* w, double* h, int accumulate,
             hipStream_t s);
}  // namespace kry

}  // namespace dgmi

struct dgmi_ctx {
  int D = 0, nloc = 0, ngi = 0, sngi = 0, nface = 0, ng = 0;
  int64_t nele = 0, nverts = 0, ndof = 0;
  double eta_e = 36.0;
  dgmi::ElementTables el;
  dgmi::Topology topo;
  std::vector<int64_t> cells;
  std::vector<double> verts, kappa;
  int32_t* nbface = nullptr;
  int32_t* fcode = nullptr;
  int32_t* aln = nullptr;
  double* detJ = nullptr;  // (nele) FP64
  dgmi::DevGeom<double> g64;
  dgmi::DevGeom<float> g32;
  int smoother_variant = -1;
  int gmres_orth = 1;  // 0: modified Gram-Schmidt, 1: CGS2 (fused passes)
  double omega = 2.0 / 3.0;
  dgmi::MGDev mg;
  // solver scratch (FP64 and FP32 vectors of ndof), created on demand
  std::vector<void*> allocs;
  size_t dev_bytes = 0;
  void* alloc(size_t bytes);
};

namespace dgmi {
// kernel launchers (apply.hip)
template <typename T>
int launch_apply(dgmi_ctx* c, DevGeom<T>& g, const T* x, const T* b, T* r, hipStream_t s);
template <typename T>
int launch_apply_pos(dgmi_ctx* c, DevGeom<T>& g, const T* x, T* r, hipStream_t s);
template <typename T>
int launch_smooth(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, T* r_out, hipStream_t s);
int launch_extract_diag(dgmi_ctx* c, const double* B, double* d, hipStream_t s);
int launch_rhs(dgmi_ctx* c, const double* refm, const double* detJ, const double* f,
               const double* ubc, double* b, hipStream_t s);
template <typename T>
int launch_smooth_zero(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, hipStream_t s);
int launch_bdiag(dgmi_ctx* c, const double* refmats, double* bdiag, hipStream_t s);
template <typename T>
int launch_build_S_jacobi(dgmi_ctx* c, const double* bdiag, double omega, T* S, hipStream_t s);
template <typename T>
int launch_convert(const double* in, T* out, int64_t n, hipStream_t s);
// transfers + SFC levels (mg.hip)
template <typename T>
int launch_restrict(dgmi_ctx* c, const T* r_dg, T* r_sfc, hipStream_t s);
template <typename T>
int launch_prolong_add(dgmi_ctx* c, const T* e_sfc, T* x_dg, hipStream_t s);
template <typename T>
int launch_sfc_vcycle(dgmi_ctx* c, const T* r_sfc, T* e_sfc, hipStream_t s);
int upload_level_descs(dgmi_ctx* c);
int choose_small_level(const SfcLevelsDev& lv, const std::vector<int64_t>& nnz, size_t budget);
template <typename T>
size_t lds_bytes_from(const SfcLevelsDev& lv, const std::vector<int64_t>& nnz, int l0);
template <typename T>
int set_small_lds_attr(size_t bytes);
int launch_probe_mask(dgmi_ctx* c, const int32_t* whichc_dev, int color, double* x_dg,
                      hipStream_t s);
int launch_probe_collect(dgmi_ctx* c, const double* r_sfc, const int32_t* spidx,
                         const int32_t* rowpos, int64_t nent, double* values, hipStream_t s);
// vector kernels (krylov.hip)
// mode 0: out[0] = x.y ; mode 1: out[0] = x.y, out[1] = sqrt(x.y)
template <typename T>
int launch_dot(int64_t n, const T* x, const T* y, double* out, int mode, hipStream_t s);
template <typename T>
int launch_axpy(int64_t n, double a, const T* x, T* y, hipStream_t s);
}  // namespace dgmi
