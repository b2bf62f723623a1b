// Internal declarations shared by the libdgmi translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "setup_host.h"

namespace dgmi {

void set_error(const std::string& msg);
// message for a failed device allocation, with the HIP error of the last failure on this
// thread
std::string alloc_failure(const char* what = nullptr);
void note_alloc_error(const std::string& msg);

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
  size_t off_fc = 0;  // face-interpolation block (FaceDims in faces.h)
  T* G = nullptr;      // (ng, nele)
  T* cf = nullptr;     // (nface*D, nele)
  T* cnb = nullptr;    // (nface*D, nele) the neighbour's J^-T n in the face's frame
  T* wf = nullptr;     // (nface, nele)
  T* pen = nullptr;    // (nface, nele)
  T* S = nullptr;      // smoother, packed symmetric: (a, b), a <= b, at sym_index(a, b)*nele + e
  // face traces: (nface*sngi, nele) for the vector kernels, (nele, nface, sngi) for the
  // matrix-core kernels
  typename Vec2<T>::type* trace = nullptr;
  typename Vec2<T>::type* trace2 = nullptr;  // second buffer (FP32 vector sweeps)
  // bf16 inner cycles (FP32 only, exact block smoother): the packed symmetric smoother in
  // bfloat16
  uint16_t* S16 = nullptr;
  // FP32 sweeps with "smoother_half": the packed smoother of each element scaled by its
  // largest magnitude and stored in FP16, two entries per 32-bit word (packed entries 2p
  // and 2p+1 at Sh[p*nele + e]), and the scale Shs[e]
  uint32_t* Sh = nullptr;
  float* Shs = nullptr;
};

// SFC levels on the device.  "Grid" levels 0..small_level hold every component whose
// single-work-group start level is at least l, concatenated in component order (largest
// start level first); a component's segment at level l starts at its level-0 offset / 2^l,
// so the pairwise restriction (2I, 2I+1) -> I and the nearest upsample i >> 1 act within
// the component.  Segments are padded with empty rows (diagonal 1) where needed.  Levels
// past a component's start level live in a packed per-component pool.
struct SfcLevelsDev {
  int nlevel = 0;                 // grid-layout levels held (small_level + 1)
  std::vector<int64_t> n, nnz;    // rows and entries per grid level
  std::vector<int64_t> grid_rows;    // rows swept by the FP64 grid kernels at level l
  std::vector<int64_t> grid_rows32;  // the same for FP32 (its LDS kernel may start lower)
  std::vector<int32_t*> rowptr, col;
  std::vector<double*> val64, diag64;
  std::vector<float*> val32, diag32;
  // per-level work vectors (both precisions share the byte buffers, sized for FP64)
  std::vector<void*> r, e, e2;
  // packed deeper levels of every component
  int32_t* pool_rp = nullptr;
  int32_t* pool_col = nullptr;
  double* pool_val64 = nullptr;
  double* pool_diag64 = nullptr;
  float* pool_val32 = nullptr;
  float* pool_diag32 = nullptr;
  void* pool_vec = nullptr;  // r, A, B of every pooled level (FP64-sized)
  // device tables for the single-work-group kernel: LevelDesc<T>[ncomp][MAX_LEVELS] and
  // (start level, level count) per component
  void* desc64 = nullptr;
  void* desc32 = nullptr;
  int32_t* rng64 = nullptr;  // FP64 (start, count); FP32 starts one level lower where its
  int32_t* rng32 = nullptr;  // half-size hierarchy fits in LDS
  // LDS images of every component's levels (16-byte aligned pieces) and per component
  // (first 16-byte word, words to copy)
  void* img64 = nullptr;
  void* img32 = nullptr;
  int64_t* imginfo64 = nullptr;
  int64_t* imginfo32 = nullptr;
  // FP32 exact coarse solves: inverses (row-major) and per component (offset, level or
  // -1), components in layout order
  float* cinv32 = nullptr;
  int64_t* cinfo32 = nullptr;
  // a context of one solve: the inverse of its exact level (row-major), applied by a grid
  // kernel in place of the single-work-group kernel and of the grid levels below
  float* dense_rm32 = nullptr;
  int dense_n32 = 0;
  int dense_level32 = -1;
  bool any_start0 = false;    // some component's FP64 kernel starts at level 0
  bool any_start0_32 = false;
  size_t lds64 = 0, lds32 = 0;  // dynamic LDS of the single-work-group kernel
};

// Blocks of DOF vectors that never cross a component (an independent solve packed into
// the context).  Reductions sum a component's blocks in a fixed order.
struct SegBlk {
  int64_t beg, end;
  int32_t comp, pad;
};
constexpr int MDOT_K = 32;     // basis vectors per fused Gram-Schmidt pass
constexpr int SEG_CHUNK = 1024;  // DOFs per segmented block
constexpr int GS_CHUNK = 4096;   // DOFs per wide block of the Gram-Schmidt kernels
struct SegDev {
  int ncomp = 1;
  int nblk = 0;
  SegBlk* blk = nullptr;        // (nblk)
  int32_t* comp_blk = nullptr;  // (ncomp + 1) first block of each component
  int64_t* eoff = nullptr;      // (ncomp + 1) element offsets of the components
  double* partial = nullptr;    // (nblk * MDOT_K)
  double* coef = nullptr;       // (ncomp * MDOT_K)
  // wide blocks (GS_CHUNK DOFs) of the fused Gram-Schmidt kernels
  int nwblk = 0;
  SegBlk* wblk = nullptr;        // (nwblk)
  int32_t* comp_wblk = nullptr;  // (ncomp + 1)
  bool wvec4 = false;            // every wide block starts at a multiple of 4 DOFs
  // contexts with a component of many wide blocks sum the Gram-Schmidt partials in two
  // stages: super-blocks of GS_SUPER consecutive wide blocks of one component first
  bool two_stage = false;
  int nsblk = 0;
  int32_t* sblk = nullptr;       // (nsblk + 1) first wide block of each super-block
  int32_t* comp_sblk = nullptr;  // (ncomp + 1) first super-block of each component
  double* spartial = nullptr;    // (nsblk * MDOT_K)
};
constexpr int GS_SUPER = 32;         // wide blocks per super-block
constexpr int GS_TWO_STAGE = 256;    // a component with more wide blocks uses two stages

// Operand tables of the matrix-core apply (apply_mfma.hip), in MFMA lane order.
struct MfTabs {
  void* b1 = nullptr;   // bf16 A operands of Y = T Xe, lane order
  void* b2 = nullptr;   // bf16 A operands of r = T^T Z, lane order
  float* wv = nullptr;  // quadrature weights: volume points, then face points
  float* tT = nullptr;  // fp32 T, row-major (rows x nloc) plus 16 zero words
};

// a captured V-cycle and the arguments it was captured with
struct VGraph {
  hipGraphExec_t exec = nullptr;
  const void* r = nullptr;
  void* e = nullptr;
  double omega = 0.0;
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
  int small_level = 0;    // FP64 grid kernels sweep levels below this one
  int small_level32 = 0;  // FP32 grid kernels sweep levels below this one
  // FP32 cycles solve the first level of at most 128 rows exactly (default); 0 runs every
  // level's sweeps as the FP64 cycle does
  int coarse_exact = 1;
  std::vector<int> exact32;  // per component (context order): exact level or -1
  // level sizes of every component (host), components in context order
  std::vector<std::vector<int64_t>> comp_levels;
  // V-cycle replayed from a captured graph (set_param "graph_vcycle"; off by default: the
  // switch between the graph and the stream's own work costs more than the launches)
  bool use_graph = false;
  VGraph graph64, graph32;
  hipStream_t capture_stream = nullptr;
};

namespace kry {
// active (may be null): components with a zero flag are skipped (their result is 0)
template <typename T>
int sdot(const SegDev& sg, const T* x, const T* y, double* out, double* sq, hipStream_t s,
         const int32_t* active = nullptr);
template <typename T>
int sorth_pass(const SegDev& sg, int k, const T* V, int64_t ldv, T* w, double* h, int64_t hs,
               int accumulate, const int32_t* active, hipStream_t s);
// CGS2 of w against V_0..V_{k-1}, 1 <= k <= MDOT_K, per component: h[c*hs + j] = the sum
// of both passes' coefficients, hn[c] = ||w|| afterwards (0 for inactive components)
template <typename T>
int sorth_cgs2(const SegDev& sg, int k, const T* V, int64_t ldv, T* w, double* h, int64_t hs,
               double* hn, const int32_t* active, hipStream_t s);
template <typename Tin, typename Tout>
int sscale(const SegDev& sg, const double* a, int mode, const int32_t* active, const Tin* x,
           Tout* y, hipStream_t s);
template <typename Tin>
int sadd(const SegDev& sg, const double* a, const Tin* x, double* y, hipStream_t s);
template <typename T>
int scombine(const SegDev& sg, const T* V, int64_t ldv, const double* yv, int64_t ys,
             const int32_t* kk, T* x, hipStream_t s);
int scycle_start(int ncomp, const double* beta, const double* eps, const int32_t* want,
                 int32_t* active, double* thr, double* g, int64_t gs, int glen, int32_t* kk,
                 hipStream_t s);
// status[0] = active components after the step, status[1 + c] = res[c] / thr[c] for the
// still-active components (0 for the others)
int sgivens(int ncomp, int j, int ldh, int64_t hs, double* H, const double* hn, double* cs,
            double* sn, int64_t rs, double* g, int64_t gs, double* res, const double* thr,
            int32_t* active, int32_t* kk, int32_t* nact, double* status, hipStream_t s);
int sbacksolve(int ncomp, int ldh, int64_t hs, const double* H, const double* g, int64_t gs,
               const int32_t* kk, double* yv, int64_t ys, hipStream_t s);
// MG-preconditioned CG per component (z02/volume_mf_diff.py:1875-1931): start flags,
// x += alpha p and r -= alpha Ap, p = z + beta p, and the end-of-iteration step (its,
// rz = rznew, convergence; status[0] = components still active)
int cg_start(int ncomp, const double* rr, const double* tol, const int32_t* want, double* thr,
             int32_t* its, int32_t* active, hipStream_t s);
template <typename T>
int cg_xr(const SegDev& sg, const double* rz, const double* pap, const int32_t* active,
          const T* p, const T* ap, T* x, T* r, hipStream_t s);
template <typename T>
int cg_p(const SegDev& sg, const double* rznew, const double* rz, const int32_t* active,
         const T* z, T* p, hipStream_t s);
int cg_step(int ncomp, const double* rr, double* rz, const double* rznew, const double* pap,
            const double* thr, int max_its, int32_t* its, int32_t* active, double* status,
            hipStream_t s);
// act64[b] = any component overlapping elements [64b, 64b+64) has flag[c] != 0
int expand_active(const SegDev& sg, int64_t nele, const int32_t* flag, int32_t* act64,
                  hipStream_t s);
}  // namespace kry

}  // namespace dgmi

namespace dgmi {
// The P1CG pattern and every component's Hilbert order need only the cells and the
// vertices: a host thread started at context creation forms them while the rest of the
// set-up (topology, geometry, smoother) runs, and dgmi_mg_setup takes them.
struct GraphJob {
  std::thread th;
  std::vector<int64_t> voff;  // the component vertex ranges the orders are for
  std::vector<int64_t> fina, cola;
  std::vector<std::vector<int64_t>> lsfc;  // local 1-based SFC positions per component
  std::vector<SfcLevelPattern> pat;        // each component's SFC level patterns
  bool ok = false;
};
}  // namespace dgmi

struct dgmi_ctx {
  int D = 0, nloc = 0, ngi = 0, sngi = 0, nface = 0, ng = 0;
  int64_t nele = 0, nverts = 0, ndof = 0;
  double eta_e = 36.0;
  dgmi::ElementTables el;
  dgmi::Topology topo;
  dgmi::GraphJob* graph = nullptr;  // joined by dgmi_mg_setup or dgmi_ctx_destroy
  std::vector<int64_t> cells;
  std::vector<double> verts, kappa;
  int32_t* nbface = nullptr;
  int32_t* fcode = nullptr;
  int32_t* aln = nullptr;
  int32_t* fnodes = nullptr;  // (nface, snloc) the nodes on each face
  // node relabellings of the face frames ((D+1)! x nloc; geometry.hip), and whether every
  // interior face has its frame (the single-pass kernels need them)
  uint8_t* nperm = nullptr;
  bool frames_ok = false;
  double* detJ = nullptr;  // (nele) FP64
  dgmi::DevGeom<double> g64;
  dgmi::DevGeom<float> g32;
  int smoother_variant = -1;
  int gmres_orth = 1;  // 0: modified Gram-Schmidt, 1: CGS2 (fused passes)
  // kernel family of every FP32-vector apply/smooth: 0 vector FMA (default: faster on
  // every mesh measured, and its sweeps carry the face traces), 1 MFMA fp32, 2 MFMA with
  // bf16 inputs
  int f32_kernel = 0;
  // vector kernels in one pass (face traces formed from the rows of x, own and
  // neighbours', the neighbour's row relabelled into the face's frame) instead of a trace
  // pass before them: 0 never, 1 the applies, 2 the applies and the FP32 preconditioner's
  // sweeps (at 48^3x6 tets: FP32 apply 292 -> 234 us, FP64 609 -> 466 us, sweep 321 ->
  // ~240 us)
  int single_pass = 2;
  // FP32 single-pass applies and sweeps and the FP32 sweep from zero: 1 = k_staged (the
  // block's rows read and written through LDS with coalesced vector accesses, each packed
  // smoother entry read once per block), 0 = k_residual_w / k_smooth_zero
  int staged = 1;
  // FP32 block-staged sweeps read the FP16 copy of the packed smoother (g32.Sh, formed by
  // dgmi_smoother_setup when set): 3D by default (48^3x6: sweep 188 -> 158 us, 102 -> 107
  // iterations, solve 9% faster); 2D off (its smoother is a smaller share of the sweep's
  // bytes, and at high contrast 219 -> 254 iterations)
  int smoother_half = -1;  // -1: by dimension
  // P1CG operator RAR (read by dgmi_mg_setup): 1 = coloured probing with FP64 operator
  // applies as the reference does; 0 = assembled from the element matrices of the continuous
  // P1 hat functions, where the interior-face terms vanish (every one carries a jump), so
  // RAR[i][j] = sum over the elements holding i and j of the P1-reduced volume and
  // Dirichlet-face terms, over mult_i mult_j.  Equal to rounding; 0 needs no colouring and
  // no operator applies.
  int rar_probing = 0;
  // pre- and post-smoothing sweeps of the FP32 cycles with carried traces (the reference's
  // pre_smooth_its, z02/config.py:112)
  int f32_smooth_its = 3;
  // FP32 cycles of a context of one solve: the exact (dense inverse) level is the first SFC
  // level of at most this many rows (set before dgmi_mg_setup)
  int dense_max_n = 4096;
  // FP32 smoothing sweeps with the exact block smoother in the neighbour-only form
  // x = (1 - omega) x + S (b - A_off x) (no volume term, half the face terms)
  int smooth_offdiag = 1;
  // bf16 inner cycles with the exact block smoother: 1 = the sweeps read a bfloat16 copy of
  // the smoother (vector kernels); 0 = bf16-input matrix-core applies
  int bf16_smoother = 0;
  // waves (16 elements each) per work-group of the fp32 / bf16 matrix-core applies
  int mf_wpb = 4, bf_wpb = 4;
  bool pc_bf16 = false;  // FP32 preconditioner applies with bf16 inputs
  // FP32 refinement cycles: 1 flexible (right-preconditioned) GMRES, 0 left-preconditioned
  int inner_flexible = 1;
  // the solvers' true residual b - A x: 0 FP64 always; 1 double-double once a component
  // nears its threshold or stalls (FP64 rounding of A x can reach the threshold); 2 always
  int residual_dd = 1;
  // dgmi_ctx_destroy: 1 = synchronise the device before giving the memory back; 0 = the
  // caller has finished all work it queued for this context
  int destroy_sync = 1;
  // FP32/BF16 refinement: 1 = once a cycle leaves a component's residual nearly unchanged,
  // the remaining cycles run in FP64; 0 = every cycle in the inner precision
  int fp64_tail = 1;
  // independent solves packed in this context: component c owns elements
  // [comp_eoff[c], comp_eoff[c+1]) and vertices [comp_voff[c], comp_voff[c+1])
  int ncomp = 1;
  std::vector<int64_t> comp_eoff, comp_voff;
  dgmi::SegDev seg;
  // per 64 elements: 1 while the element's solve is unfinished; element kernels skip
  // finished solves while mask_on (set only inside the GMRES driver)
  int32_t* act64 = nullptr;
  bool mask_on = false;
  void* scratch = nullptr;  // solver scratch vectors (dgmi_api.hip)
  void* krylov = nullptr;   // Krylov workspace (dgmi_api.hip)
  void* cgwork = nullptr;   // CG workspace (dgmi_api.hip)
  dgmi::MfTabs mf;
  double omega = 2.0 / 3.0;
  dgmi::MGDev mg;
  // solver scratch (FP64 and FP32 vectors of ndof), created on demand
  std::vector<void*> allocs;
  size_t dev_bytes = 0;
  // allocations that would take dev_bytes above this many bytes fail (0: no limit)
  size_t alloc_limit = 0;
  // host wall-clock seconds of the set-up phases (dgmi_setup_timing)
  double t_setup[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // stream of the set-up work of context creation (dgmi_mesh_desc::stream)
  hipStream_t ustream = nullptr;
  void* stager = nullptr;  // pinned staging buffers of h2d (dgmi_api.hip)
  void* alloc(size_t bytes);
};

namespace dgmi {
// Device memory of contexts and set-up work buffers.  A block given back with dev_free is
// kept in a per-device cache and handed out again by a later dev_alloc of about its size
// (within 1/4): hipFree waits for the whole device to go idle, which in a serving process
// means for every other job's queued work.  A block must not be given back while queued
// work still uses it.  When hipMalloc fails, the cached blocks of the device are freed and
// the allocation is tried once more; dgmi_release_cached_memory frees them on request.
void* dev_alloc(size_t bytes);
void dev_free(void* p);
// Set-up transfers between host memory and the device.  A copy from pageable host memory
// is staged by the HIP runtime in pieces, each of which waits in a hardware queue: where
// streams share one (more streams than GPU_MAX_HW_QUEUES), behind every kernel another
// stream queued there before it, which in a serving process can be a whole solve.  From
// pinned memory the copy engine moves the data without waiting there.  h2d copies the
// host data into the context's two pinned staging buffers (filled in turn, so that the
// host fills one while the other's copies run) and queues the copies on s; the host data
// may be reused at once, the device data are ready in the order of s.  h2d_release waits
// for the staged copies and gives the staging buffers back.  d2h copies device data into
// host memory through a pinned buffer and returns when they are there.
int h2d(dgmi_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s);
int h2d_release(dgmi_ctx* c);
int d2h(void* dst, const void* src, size_t bytes, hipStream_t s);
// set-up phases timed into dgmi_ctx::t_setup
enum SetupPhase {
  PH_TOPOLOGY = 0, PH_GEOMETRY, PH_UPLOAD, PH_SMOOTHER, PH_GRAPH, PH_PROBE, PH_LEVELS, PH_LAYOUT
};
double wall_seconds();
// FP64 element and face geometry (c->g64.G, cf, wf, pen; c->nbface, fcode, detJ) from the
// host P3 node coordinates, given in nparts pieces (piece q holds elements
// [part_off[q], part_off[q+1])), formed on the device; *max_dev = the largest node
// deviation from the affine position relative to the element size, *bad_elem = the first
// degenerate element or -1 (geometry.hip)
int build_geometry_device(dgmi_ctx* c, const double* const* x_parts, const int64_t* part_off,
                          int nparts, int face_kappa, double* max_dev, int64_t* bad_elem);
// kernel launchers (apply.hip)
template <typename T>
int launch_apply(dgmi_ctx* c, DevGeom<T>& g, const T* x, const T* b, T* r, hipStream_t s);
template <typename T>
int launch_apply_pos(dgmi_ctx* c, DevGeom<T>& g, const T* x, T* r, hipStream_t s);
// r = b - A x of the FP64 operator evaluated in double-double arithmetic, rounded to FP64
int launch_residual_dd(dgmi_ctx* c, const double* x, const double* b, double* r, hipStream_t s);
template <typename T>
int launch_smooth(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, T* r_out, hipStream_t s);
int launch_extract_diag(dgmi_ctx* c, const double* B, double* d, hipStream_t s);
// matrix-core apply (apply_mfma.hip); mode 0 r=-Ax, 1 r=b-Ax, 2 x+=S(b-Ax), 3 as 2 + r, 4 r=Ax
int build_mfma_tables(dgmi_ctx* c);
int launch_apply_mf(dgmi_ctx* c, bool bf16, float* x, const float* b, float* r, int mode,
                    hipStream_t s);
// gneu (Neumann flux data) and alpha (transient term, scaled by inv_dt) may be null
int launch_rhs(dgmi_ctx* c, const double* refm, const double* detJ, const double* f,
               const double* ubc, const double* gneu, const double* alpha, double inv_dt,
               double* b, hipStream_t s);
template <typename T>
int launch_smooth_zero(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, hipStream_t s);
// bdiag from the reference tables in the form of k_bdiag (bdiag_tables on the host)
int launch_bdiag(dgmi_ctx* c, const double* tab, double* bdiag, hipStream_t s);
// partial (2 nb doubles): per block of 256 elements the l2^2 part and the max |x_num - x_ana|
template <typename T>
int launch_l2_error(dgmi_ctx* c, const T* xn, const T* xa, const double* cn, double* partial,
                    int nb, hipStream_t s);
template <typename T>
int launch_build_S_jacobi(dgmi_ctx* c, const double* bdiag, double omega, T* S, hipStream_t s);
// omega * inverse of every diagonal block, written as the packed smoother copies (g64.S,
// g32.S, g32.S16 when allocated, and with `half` the FP16 copy g32.Sh/Shs); *fail = 1 +
// element index of a singular block
int launch_build_S_direct(dgmi_ctx* c, const double* bdiag, double omega, bool half, int* fail,
                          hipStream_t s);
// FP32 vector-family sweeps carrying traces between them (see apply.hip)
bool trace_io(const dgmi_ctx* c);
// face-trace buffers of the two-pass kernels, allocated at first use (false: no memory)
template <typename T> bool ensure_traces(dgmi_ctx* c, DevGeom<T>& g);
// single-pass applies / FP32 neighbour-form sweeps (face traces from the rows of x)
bool single_pass_apply(const dgmi_ctx* c);
bool single_pass_sweeps(const dgmi_ctx* c);
int launch_sweep_nbx(dgmi_ctx* c, const float* x, const float* b, float* xnew, hipStream_t s);
// (s16: read the bfloat16 copy of the smoother)
int launch_smooth_io(dgmi_ctx* c, float* x, const float* b, int tin, int tout, hipStream_t s,
                     bool s16 = false);
// x = S b (the first sweep from x = 0), traces of x into buffer tout
int launch_smooth_zero_io(dgmi_ctx* c, float* x, const float* b, int tout, hipStream_t s,
                          bool s16 = false);
int launch_apply_io(dgmi_ctx* c, const float* x, const float* b, float* r, int mode, int tin,
                    hipStream_t s);
// the exact block smoother has a bfloat16 copy for the bf16 inner cycles
bool bf16_smoother(const dgmi_ctx* c);
// g64.S / g32.S = packed symmetric FP64 / FP32 copies of the full FP64 smoother S (and
// g32.S16 its bfloat16 rounding, when allocated)
int launch_pack_S_sym(dgmi_ctx* c, const double* S, hipStream_t s);
// g32.Sh / g32.Shs from the packed FP64 smoother g64.S
int launch_pack_S_half(dgmi_ctx* c, hipStream_t s);
template <typename T>
int launch_convert(const double* in, T* out, int64_t n, hipStream_t s);
// transfers + SFC levels (mg.hip)
// with e_first set (only when vcycle_fuses_first), the restriction also writes the
// V-cycle's first level-0 sweep from zero into e_first, the cycle's e_sfc
template <typename T>
int launch_restrict(dgmi_ctx* c, const T* r_dg, T* r_sfc, hipStream_t s, T* e_first = nullptr);
template <typename T> bool vcycle_fuses_first(const dgmi_ctx* c);
template <typename T>
int launch_prolong_add(dgmi_ctx* c, const T* e_sfc, T* x_dg, hipStream_t s);
// FP32 vector family: x += P e with the traces of the new x written into buffer tout
int launch_prolong_trace(dgmi_ctx* c, const float* e_sfc, float* x_dg, int tout, hipStream_t s);
template <typename T>
int launch_sfc_vcycle(dgmi_ctx* c, const T* r_sfc, T* e_sfc, hipStream_t s,
                      bool first_done = false);
// single-work-group part of the V-cycle: LDS bytes of levels l0.. of one component, the
// first level from which a component fits in `budget`, and the device level tables
template <typename T>
size_t lds_bytes_from(const std::vector<int64_t>& n, const std::vector<int64_t>& nnz, int l0);
template <typename T>
int lds_start_level(const std::vector<int64_t>& n, const std::vector<int64_t>& nnz, size_t budget);
// Host description of the SFC hierarchy of every component and of its device layout.
struct SfcLayout {
  std::vector<SfcLevelsHost> lh;  // per component (context order), local indices
  std::vector<int> start;         // single-work-group start level per component (FP64)
  std::vector<int> start32;       // the same for FP32: start or start - 1
  std::vector<int> order;         // components in layout order (start level descending)
  std::vector<int64_t> off0;      // level-0 offset of each component (context order)
  std::vector<int64_t> len0;      // level-0 segment length (padded except the last)
  int smax = 0;
};
int upload_sfc_levels(dgmi_ctx* c, const SfcLayout& lay, hipStream_t s);
template <typename T>
int set_small_lds_attr(size_t bytes);
// stand-alone level operations on caller-owned CSR levels (mg.hip): one Jacobi step with
// its residual rr = b - A e, e += omega rr / diag; pairwise restriction (zero pad); nearest
// upsampling added into the fine vector; P3DG <-> P1DG element transfers (dir 0: to P1DG)
template <typename T>
int launch_level_smooth(int n, const int32_t* rp, const int32_t* col, const T* val,
                        const T* diag, T omega, T* e, const T* b, T* rr, hipStream_t s);
template <typename T>
int launch_level_pairsum(int n, const T* fine, T* coarse, hipStream_t s);
template <typename T>
int launch_level_upsample_add(int n, const T* coarse, T* fine, hipStream_t s);
template <typename T>
int launch_pn_p1dg(int D, int dir, int nele, const T* x, T* y, hipStream_t s);
int launch_probe_mask(dgmi_ctx* c, const int32_t* whichc_dev, int color, double* x_dg,
                      hipStream_t s);
int launch_probe_collect(dgmi_ctx* c, const double* r_sfc, const int32_t* spidx,
                         const int32_t* rowpos, int64_t nent, double* values, hipStream_t s);
// RAR assembled (rar_probing 0): ke = the P1-reduced element matrices (rarm = the reduced
// reference tables), then vals[k] for every CSR slot of the rows by position (slots
// [rlo[p], rhi[p]) of position p, column positions colpos); vals[nslot] = 1 when an element
// pair has no slot
int launch_rar_assemble(dgmi_ctx* c, const double* rarm, double* ke, const int32_t* rlo,
                        const int32_t* rhi, const int32_t* colpos, int64_t nslot,
                        double* vals, hipStream_t s);
// vector kernels (krylov.hip)
// mode 0: out[0] = x.y ; mode 1: out[0] = x.y, out[1] = sqrt(x.y)
template <typename T>
int launch_dot(int64_t n, const T* x, const T* y, double* out, int mode, hipStream_t s);
template <typename T>
int launch_axpy(int64_t n, double a, const T* x, T* y, hipStream_t s);
// the dot product with its partial sums in a fixed device array (no allocation); n up to
// dot_noalloc_max_n()
template <typename T>
int launch_dot_noalloc(int64_t n, const T* x, const T* y, double* out, hipStream_t s);
int64_t dot_noalloc_max_n();
}  // namespace dgmi
