// Krylov vector kernels (gfx950): deterministic dot products, AXPYs with device-resident
// scalars, Givens updates of the Hessenberg column, and the basis combination x += V y.
// All reductions use a fixed grid and a fixed tree, so results do not depend on timing.
#include "dgmi_internal.h"

namespace dgmi {

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[RED_THREADS / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < RED_THREADS / 64; ++w) t += sh[w];
  return t;
}

// stage 1: partial[b] = sum over a fixed strided slice of x*y (FP64 accumulation)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_dot_partial(int64_t n, const T* __restrict__ x,
                                                             const T* __restrict__ y,
                                                             double* __restrict__ partial) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s += (double)x[i] * (double)y[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// stage 2: out[0] = sum(partial); mode 1 also stores sqrt into out[1]
__global__ __launch_bounds__(RED_THREADS) void k_dot_final(const double* __restrict__ partial,
                                                           int nb, double* __restrict__ out,
                                                           int mode) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    out[0] = s;
    if (mode == 1) out[1] = sqrt(s);
  }
}

// y += a x  (a from host)
template <typename T>
__global__ void k_axpy(int64_t n, double a, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (T)((double)y[i] + a * (double)x[i]);
}

// Fused classical Gram-Schmidt passes run over up to MDOT_K basis vectors at a time:
//   h[j] = <w, V_j> (one read of w for all j), then w -= sum_j h[j] V_j.


// ---------------------------------------------------------------------------------------
// Segmented (per-component) kernels.  A context may hold several independent solves
// ("components", contiguous DOF ranges).  Vectors are cut into blocks that never cross a
// component (SegBlk); every reduction sums a component's blocks in a fixed order, so a
// component's dot products do not depend on what else is in the batch.  Per-component
// scalars live in device arrays indexed by component.
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_sdot_partial(const SegBlk* __restrict__ sb,
                                                              const T* __restrict__ x,
                                                              const T* __restrict__ y,
                                                              double* __restrict__ partial,
                                                              const int32_t* __restrict__ active) {
  const SegBlk b = sb[blockIdx.x];
  if (active && !active[b.comp]) {  // finished component (uniform over the block)
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  double acc = 0.0;
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x) acc += (double)x[i] * (double)y[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// one work-group per component: out[c] = sum of its partials; sq[c] = sqrt(out[c])
__global__ __launch_bounds__(RED_THREADS) void k_sdot_final(const int32_t* __restrict__ comp_blk,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ out,
                                                            double* __restrict__ sq) {
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int b = comp_blk[c] + threadIdx.x; b < comp_blk[c + 1]; b += blockDim.x) acc += partial[b];
  acc = block_sum(acc);
  if (threadIdx.x == 0) {
    if (out) out[c] = acc;
    if (sq) sq[c] = sqrt(acc);
  }
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_smdot_partial(const SegBlk* __restrict__ sb,
                                                               int k, const T* __restrict__ V,
                                                               int64_t ldv,
                                                               const T* __restrict__ w,
                                                               double* __restrict__ partial,
                                                               const int32_t* __restrict__ active) {
  const SegBlk b = sb[blockIdx.x];
  if (active && !active[b.comp]) {  // finished component (uniform over the block)
    if (threadIdx.x < k) partial[(int64_t)blockIdx.x * MDOT_K + threadIdx.x] = 0.0;
    return;
  }
  double acc[MDOT_K];
#pragma unroll
  for (int j = 0; j < MDOT_K; ++j) acc[j] = 0.0;
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x) {
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < MDOT_K; ++j)
      if (j < k) acc[j] += wi * (double)V[(int64_t)j * ldv + i];
  }
  __shared__ double sh[RED_THREADS / 64][MDOT_K];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < MDOT_K; ++j) {
    if (j < k) {
      const double v = wave_sum(acc[j]);
      if (lane == 0) sh[wid][j] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
    for (int ww = 0; ww < RED_THREADS / 64; ++ww) t += sh[ww][threadIdx.x];
    partial[(int64_t)blockIdx.x * MDOT_K + threadIdx.x] = t;
  }
}

// per component c: coef[c][j] = sum of partials (0 when inactive); h[c*hs + j] (+)= coef
__global__ __launch_bounds__(RED_THREADS) void k_smdot_final(const int32_t* __restrict__ comp_blk,
                                                             const double* __restrict__ partial,
                                                             int k, double* __restrict__ h,
                                                             int64_t hs, double* __restrict__ coef,
                                                             int accumulate,
                                                             const int32_t* __restrict__ active) {
  // all k sums at once: lane group j (8 lanes) sums column j of the component's partials
  // in a fixed strided order, then a fixed shuffle tree
  constexpr int G = RED_THREADS / MDOT_K;  // lanes per column
  static_assert(G * MDOT_K == RED_THREADS && (G & (G - 1)) == 0, "lane groups");
  const int c = blockIdx.x;
  const int j = threadIdx.x / G, sub = threadIdx.x % G;
  double acc = 0.0;
  if (j < k)
    for (int b = comp_blk[c] + sub; b < comp_blk[c + 1]; b += G)
      acc += partial[(int64_t)b * MDOT_K + j];
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (j < k && sub == 0) {
    const bool on = active[c] != 0;
    const double v = on ? acc : 0.0;
    coef[c * MDOT_K + j] = v;
    if (on) h[c * hs + j] = accumulate ? h[c * hs + j] + v : v;
  }
}

// w -= sum_j coef[comp][j] V_j
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_smaxpy(const SegBlk* __restrict__ sb, int k,
                                                        const T* __restrict__ V, int64_t ldv,
                                                        const double* __restrict__ coef,
                                                        T* __restrict__ w,
                                                        const int32_t* __restrict__ active) {
  const SegBlk b = sb[blockIdx.x];
  if (active && !active[b.comp]) return;
  const double* cf = coef + b.comp * MDOT_K;
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x) {
    double s = (double)w[i];
    for (int j = 0; j < k; ++j) s -= cf[j] * (double)V[(int64_t)j * ldv + i];
    w[i] = (T)s;
  }
}

// y = scale * x per component: mode 0 scale = 1/a[c]; 1 scale = a[c].  Inactive
// components (or a[c] = 0 in mode 0) get y = 0.
template <typename Tin, typename Tout>
__global__ __launch_bounds__(RED_THREADS) void k_sscale(const SegBlk* __restrict__ sb,
                                                        const double* __restrict__ a, int mode,
                                                        const int32_t* __restrict__ active,
                                                        const Tin* __restrict__ x,
                                                        Tout* __restrict__ y) {
  const SegBlk b = sb[blockIdx.x];
  const double av = a[b.comp];
  const bool on = (!active || active[b.comp]) && (mode == 1 || av != 0.0);
  const double f = on ? (mode == 0 ? 1.0 / av : av) : 0.0;
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x) y[i] = (Tout)((double)x[i] * f);
}

// y64 += a[c] * x
template <typename Tin>
__global__ __launch_bounds__(RED_THREADS) void k_sadd(const SegBlk* __restrict__ sb,
                                                      const double* __restrict__ a,
                                                      const Tin* __restrict__ x,
                                                      double* __restrict__ y) {
  const SegBlk b = sb[blockIdx.x];
  const double av = a ? a[b.comp] : 1.0;
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x) y[i] += av * (double)x[i];
}

// x = sum_{j < kk[c]} yv[c][j] V_j
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_scombine(const SegBlk* __restrict__ sb,
                                                          const T* __restrict__ V, int64_t ldv,
                                                          const double* __restrict__ yv,
                                                          int64_t ys, const int32_t* __restrict__ kk,
                                                          T* __restrict__ x) {
  const SegBlk b = sb[blockIdx.x];
  const int k = kk[b.comp];
  const double* yc = yv + b.comp * ys;
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += yc[j] * (double)V[(int64_t)j * ldv + i];
    x[i] = (T)s;
  }
}

// Cycle start, one thread per component: active[c] = want[c] and beta[c] finite > 0;
// thr[c] = eps[c] beta[c]; g[c] = (beta[c], 0, ...); kk[c] = 0.
__global__ void k_scycle_start(int ncomp, const double* __restrict__ beta,
                               const double* __restrict__ eps, const int32_t* __restrict__ want,
                               int32_t* __restrict__ active, double* __restrict__ thr,
                               double* __restrict__ g, int64_t gs, int glen,
                               int32_t* __restrict__ kk) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncomp) return;
  const double b = beta[c];
  const bool on = want[c] != 0 && b > 0.0 && isfinite(b);
  active[c] = on ? 1 : 0;
  thr[c] = eps[c] * b;
  for (int i = 0; i < glen; ++i) g[c * gs + i] = (i == 0 && on) ? b : 0.0;
  kk[c] = 0;
}

// Givens step j for every active component (Saad, Alg. 6.9 with rotations); the
// subdiagonal entry comes from hn[c].  A component whose estimate |g[j+1]| falls to
// thr[c] (or whose basis breaks down) turns inactive.  nact[0] counts the still-active ones.
__global__ void k_sgivens(int ncomp, int j, int ldh, int64_t hs, double* __restrict__ H,
                          const double* __restrict__ hn, double* __restrict__ cs,
                          double* __restrict__ sn, int64_t rs, double* __restrict__ g,
                          int64_t gs, double* __restrict__ res, const double* __restrict__ thr,
                          int32_t* __restrict__ active, int32_t* __restrict__ kk,
                          int32_t* __restrict__ nact) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncomp) return;
  if (!active[c]) return;
  double* h = H + c * hs + (int64_t)j * ldh;
  double* csc = cs + c * rs;
  double* snc = sn + c * rs;
  double* gc = g + c * gs;
  h[j + 1] = hn[c];
  for (int i = 0; i < j; ++i) {
    const double t = csc[i] * h[i] + snc[i] * h[i + 1];
    h[i + 1] = -snc[i] * h[i] + csc[i] * h[i + 1];
    h[i] = t;
  }
  const double a = h[j], bb = h[j + 1];
  const double r = hypot(a, bb);
  const double cc = r > 0 ? a / r : 1.0, ss = r > 0 ? bb / r : 0.0;
  csc[j] = cc;
  snc[j] = ss;
  h[j] = r;
  h[j + 1] = 0.0;
  gc[j + 1] = -ss * gc[j];
  gc[j] = cc * gc[j];
  res[c] = fabs(gc[j + 1]);
  kk[c] = j + 1;
  const bool stop = !(res[c] > thr[c]) || !(hn[c] > 0.0) || r == 0.0;
  if (stop) active[c] = 0;
  else atomicAdd(nact, 1);
}

// back substitution R y = g per component over its kk[c] columns
__global__ void k_sbacksolve(int ncomp, int ldh, int64_t hs, const double* __restrict__ H,
                             const double* __restrict__ g, int64_t gs,
                             const int32_t* __restrict__ kk, double* __restrict__ yv,
                             int64_t ys) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncomp) return;
  const double* Hc = H + c * hs;
  const double* gc = g + c * gs;
  double* yc = yv + c * ys;
  const int k = kk[c];
  for (int i = k - 1; i >= 0; --i) {
    double s = gc[i];
    for (int q = i + 1; q < k; ++q) s -= Hc[(int64_t)q * ldh + i] * yc[q];
    yc[i] = s / Hc[(int64_t)i * ldh + i];
  }
}

__global__ void k_expand_active(int64_t nele, int ncomp, const int64_t* __restrict__ eoff,
                                const int32_t* __restrict__ flag, int32_t* __restrict__ act64) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t lo = 64 * b, hi = min(lo + 64, nele);
  if (lo >= nele) return;
  int a = 0, z = ncomp;  // first component whose range ends after lo
  while (a < z) {
    const int m = (a + z) / 2;
    if (eoff[m + 1] <= lo) a = m + 1;
    else z = m;
  }
  int on = 0;
  for (int c = a; c < ncomp && eoff[c] < hi; ++c) on |= flag[c] != 0;
  act64[b] = on;
}

__global__ void k_izero(int n, int32_t* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------------------
// partial sums of the context-free dot (one buffer per host thread, so calls from
// different threads on different streams never share it)
static thread_local double* g_partial = nullptr;

static int ensure_partial() {
  if (!g_partial) DGMI_HIP(hipMalloc(&g_partial, sizeof(double) * RED_BLOCKS));
  return 0;
}

template <typename T>
int launch_dot(int64_t n, const T* x, const T* y, double* out, int mode, hipStream_t s) {
  if (int rc = ensure_partial()) return rc;
  const int nb = (int)std::min<int64_t>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS + 1);
  hipLaunchKernelGGL((k_dot_partial<T>), dim3(nb), dim3(RED_THREADS), 0, s, n, x, y, g_partial);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s, g_partial, nb, out, mode);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_axpy(int64_t n, double a, const T* x, T* y, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_axpy<T>), grd, blk, 0, s, n, a, x, y);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template int launch_dot<double>(int64_t, const double*, const double*, double*, int, hipStream_t);
template int launch_dot<float>(int64_t, const float*, const float*, double*, int, hipStream_t);
template int launch_axpy<double>(int64_t, double, const double*, double*, hipStream_t);
template int launch_axpy<float>(int64_t, double, const float*, float*, hipStream_t);

// exported to the solver translation unit
namespace kry {
// ---- segmented launchers
template <typename T>
int sdot(const SegDev& sg, const T* x, const T* y, double* out, double* sq, hipStream_t s,
         const int32_t* active) {
  hipLaunchKernelGGL((k_sdot_partial<T>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, x, y,
                     sg.partial, active);
  hipLaunchKernelGGL(k_sdot_final, dim3(sg.ncomp), dim3(RED_THREADS), 0, s, sg.comp_blk,
                     (const double*)sg.partial, out, sq);
  DGMI_HIP(hipGetLastError());
  return 0;
}

// One classical Gram-Schmidt pass of w against V_0..V_{k-1} per component; writes (or, with
// accumulate, adds) the coefficients into h[c*hs + j].  Chunks of MDOT_K vectors: within
// a chunk all dots use the same w (classical GS); successive chunks see the updated w.
template <typename T>
int sorth_pass(const SegDev& sg, int k, const T* V, int64_t ldv, T* w, double* h, int64_t hs,
               int accumulate, const int32_t* active, hipStream_t s) {
  for (int c0 = 0; c0 < k; c0 += MDOT_K) {
    const int kc = std::min(MDOT_K, k - c0);
    hipLaunchKernelGGL((k_smdot_partial<T>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, kc,
                       V + (int64_t)c0 * ldv, ldv, (const T*)w, sg.partial, active);
    hipLaunchKernelGGL(k_smdot_final, dim3(sg.ncomp), dim3(RED_THREADS), 0, s, sg.comp_blk,
                       (const double*)sg.partial, kc, h + c0, hs, sg.coef, accumulate, active);
    hipLaunchKernelGGL((k_smaxpy<T>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, kc,
                       V + (int64_t)c0 * ldv, ldv, (const double*)sg.coef, w, active);
  }
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename Tin, typename Tout>
int sscale(const SegDev& sg, const double* a, int mode, const int32_t* active, const Tin* x,
           Tout* y, hipStream_t s) {
  hipLaunchKernelGGL((k_sscale<Tin, Tout>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, a,
                     mode, active, x, y);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename Tin>
int sadd(const SegDev& sg, const double* a, const Tin* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL((k_sadd<Tin>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, a, x, y);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int scombine(const SegDev& sg, const T* V, int64_t ldv, const double* yv, int64_t ys,
             const int32_t* kk, T* x, hipStream_t s) {
  hipLaunchKernelGGL((k_scombine<T>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, V, ldv, yv,
                     ys, kk, x);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int scycle_start(int ncomp, const double* beta, const double* eps, const int32_t* want,
                 int32_t* active, double* thr, double* g, int64_t gs, int glen, int32_t* kk,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_scycle_start, dim3((ncomp + 63) / 64), dim3(64), 0, s, ncomp, beta, eps,
                     want, active, thr, g, gs, glen, kk);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int sgivens(int ncomp, int j, int ldh, int64_t hs, double* H, const double* hn, double* cs,
            double* sn, int64_t rs, double* g, int64_t gs, double* res, const double* thr,
            int32_t* active, int32_t* kk, int32_t* nact, hipStream_t s) {
  hipLaunchKernelGGL(k_izero, dim3(1), dim3(64), 0, s, 1, nact);
  hipLaunchKernelGGL(k_sgivens, dim3((ncomp + 63) / 64), dim3(64), 0, s, ncomp, j, ldh, hs, H,
                     hn, cs, sn, rs, g, gs, res, thr, active, kk, nact);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int sbacksolve(int ncomp, int ldh, int64_t hs, const double* H, const double* g, int64_t gs,
               const int32_t* kk, double* yv, int64_t ys, hipStream_t s) {
  hipLaunchKernelGGL(k_sbacksolve, dim3((ncomp + 63) / 64), dim3(64), 0, s, ncomp, ldh, hs, H, g,
                     gs, kk, yv, ys);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int expand_active(const SegDev& sg, int64_t nele, const int32_t* flag, int32_t* act64,
                  hipStream_t s) {
  const int64_t nb = (nele + 63) / 64;
  hipLaunchKernelGGL(k_expand_active, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, nele,
                     sg.ncomp, sg.eoff, flag, act64);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template int sdot<double>(const SegDev&, const double*, const double*, double*, double*,
                          hipStream_t, const int32_t*);
template int sdot<float>(const SegDev&, const float*, const float*, double*, double*, hipStream_t,
                         const int32_t*);
template int sorth_pass<double>(const SegDev&, int, const double*, int64_t, double*, double*,
                                int64_t, int, const int32_t*, hipStream_t);
template int sorth_pass<float>(const SegDev&, int, const float*, int64_t, float*, double*,
                               int64_t, int, const int32_t*, hipStream_t);
template int sscale<double, double>(const SegDev&, const double*, int, const int32_t*,
                                    const double*, double*, hipStream_t);
template int sscale<float, float>(const SegDev&, const double*, int, const int32_t*,
                                  const float*, float*, hipStream_t);
template int sscale<double, float>(const SegDev&, const double*, int, const int32_t*,
                                   const double*, float*, hipStream_t);
template int sadd<double>(const SegDev&, const double*, const double*, double*, hipStream_t);
template int sadd<float>(const SegDev&, const double*, const float*, double*, hipStream_t);
template int scombine<double>(const SegDev&, const double*, int64_t, const double*, int64_t,
                              const int32_t*, double*, hipStream_t);
template int scombine<float>(const SegDev&, const float*, int64_t, const double*, int64_t,
                             const int32_t*, float*, hipStream_t);

}  // namespace kry

}  // namespace dgmi
