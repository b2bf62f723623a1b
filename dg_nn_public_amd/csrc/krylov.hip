// Krylov vector kernels (gfx950): deterministic dot products, AXPYs with device-resident
// scalars, Givens updates of the Hessenberg column, and the basis combination x += V y.
// All reductions use a fixed grid and a fixed tree, so results do not depend on timing.
#include <mutex>

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

// ---------------------------------------------------------------------------------------
// Fused CGS2 (classical Gram-Schmidt with one re-orthogonalisation) and the norm of the
// result, over wide blocks of GS_CHUNK DOFs that never cross a component:
//   k_gs_dots            p1 = V^T w
//   k_gs_update<1>       w -= V c1, then p2 = V^T w      (c1: pass-1 coefficients)
//   k_gs_update<2>       w -= V c2, then p3 = w^T w      (c2: pass-2 coefficients)
// V is read three times per Arnoldi step instead of four (two passes of dots and updates)
// and w twice less.  Each thread owns 16-byte row groups; the k column sums of a block are
// reduced across the wave by a butterfly that halves the columns held per lane at every
// step (k-1 exchanges in place of 6 per column), then across the block's waves in LDS.
constexpr int GS_THREADS = 256;
constexpr int GS_FINAL = 1024;  // threads of the per-component final reduction

template <typename T>
struct Row16;  // 16 bytes of T
template <>
struct Row16<float> {
  using V = float4;
  static constexpr int N = 4;
};
template <>
struct Row16<double> {
  using V = double2;
  static constexpr int N = 2;
};

// N consecutive values at p: one 16-byte access when N fills it, else N scalar accesses
template <typename T, int N>
__device__ __forceinline__ void load_n(const T* __restrict__ p, T (&o)[N]) {
  if constexpr (N == Row16<T>::N) {
    const typename Row16<T>::V q = *reinterpret_cast<const typename Row16<T>::V*>(p);
    const T* qs = reinterpret_cast<const T*>(&q);
#pragma unroll
    for (int u = 0; u < N; ++u) o[u] = qs[u];
  } else {
#pragma unroll
    for (int u = 0; u < N; ++u) o[u] = p[u];
  }
}
template <typename T, int N>
__device__ __forceinline__ void store_n(T* __restrict__ p, const T (&o)[N]) {
  if constexpr (N == Row16<T>::N) {
    typename Row16<T>::V q;
    T* qs = reinterpret_cast<T*>(&q);
#pragma unroll
    for (int u = 0; u < N; ++u) qs[u] = o[u];
    *reinterpret_cast<typename Row16<T>::V*>(p) = q;
  } else {
#pragma unroll
    for (int u = 0; u < N; ++u) p[u] = o[u];
  }
}

// Sums a[0..KC) over the 64 lanes of the wave, KC a power of two.  On return a[0] of
// lane l is the sum of column l / (64 / KC).  Fixed exchange pattern, so the result does
// not depend on timing.
template <int KC>
__device__ __forceinline__ void wave_reduce_cols(double (&a)[KC]) {
  static_assert(KC >= 1 && KC <= 64 && (KC & (KC - 1)) == 0, "power-of-two column count");
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int S = KC, o = 32; S > 1; S >>= 1, o >>= 1) {
    const bool up = (lane & o) != 0;  // this lane keeps the upper half of the columns
#pragma unroll
    for (int i = 0; i < S / 2; ++i) {
      // both halves cross, and the lane picks a sum: a select between two array elements
      // would be folded into a lane-dependent index, which puts the array in scratch
      const double lo = a[i], hi = a[i + S / 2];
      const double rlo = __shfl_xor(lo, o, 64), rhi = __shfl_xor(hi, o, 64);
      a[i] = up ? hi + rhi : lo + rlo;
    }
  }
#pragma unroll
  for (int o = 32 / KC; o > 0; o >>= 1) a[0] += __shfl_xor(a[0], o, 64);
}

constexpr int pow2_ceil(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}

// out[j] = block sum of column j, j < k.  The butterfly runs over the next power of two
// of columns (the extra ones are zero).
template <int KC>
__device__ __forceinline__ void block_store_cols(const double (&a)[KC], int k,
                                                 double* __restrict__ out) {
  constexpr int P = pow2_ceil(KC);
  static_assert(P <= 64, "at most 64 columns");
  __shared__ double sh[GS_THREADS / 64][P];
  double ap[P];
#pragma unroll
  for (int j = 0; j < P; ++j) ap[j] = j < KC ? a[j] : 0.0;
  wave_reduce_cols<P>(ap);
  constexpr int L = 64 / P;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane % L == 0) sh[wid][lane / L] = ap[0];
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
#pragma unroll
    for (int ww = 0; ww < GS_THREADS / 64; ++ww) t += sh[ww][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// Column j < KC of the block's rows reads V_min(j, k-1): the loads of a row group are
// issued together with no branch, and the columns past k (zero coefficients, unused sums)
// re-read a column that is already in flight.

// acc[j] += <V_j, w> over rows [i, i + N) of the lane
template <typename T, int KC, int N>
__device__ __forceinline__ void gs_dots_rows(int k, const T* __restrict__ V, int64_t ldv,
                                             const T* __restrict__ w, int64_t i,
                                             double (&acc)[KC]) {
  // every column's loads are issued before the first product (the grid is a few waves
  // per SIMD: memory parallelism comes from within the thread)
  T wv[N];
  load_n<T, N>(w + i, wv);
  T vv[KC][N];
#pragma unroll
  for (int j = 0; j < KC; ++j) load_n<T, N>(V + (int64_t)min(j, k - 1) * ldv + i, vv[j]);
#pragma unroll
  for (int j = 0; j < KC; ++j)
#pragma unroll
    for (int u = 0; u < N; ++u) acc[j] += (double)vv[j][u] * (double)wv[u];
}

// w -= sum_j coef_j V_j over rows [i, i + N), then MODE 1: acc[j] += <V_j, w>; MODE 2:
// acc[0] += <w, w> (with w as stored)
// The update runs in T (the precision w is stored in), the dot sums in FP64.
template <typename T, int KC, int MODE, int N, int NO>
__device__ __forceinline__ void gs_update_rows(int k, const T* __restrict__ V, int64_t ldv,
                                               const T* __restrict__ cf, T* __restrict__ w,
                                               int64_t i, double (&acc)[NO]) {
  T wv[N];
  load_n<T, N>(w + i, wv);
  T vv[KC][N];
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    load_n<T, N>(V + (int64_t)min(j, k - 1) * ldv + i, vv[j]);
    const T cj = cf[j];
#pragma unroll
    for (int u = 0; u < N; ++u) wv[u] -= cj * vv[j][u];
  }
  store_n<T, N>(w + i, wv);
  if constexpr (MODE == 1) {
#pragma unroll
    for (int j = 0; j < KC; ++j)
#pragma unroll
      for (int u = 0; u < N; ++u) acc[j] += (double)(vv[j][u] * wv[u]);
  } else {
#pragma unroll
    for (int u = 0; u < N; ++u) acc[0] += (double)wv[u] * (double)wv[u];
  }
}

// partial[blk][j] = <V_j, w> over the block, j < k <= KC.  VEC: 16-byte row groups (the
// block start and V's leading dimension are multiples of the group), else single rows.
template <typename T, int KC, bool VEC>
__global__ __launch_bounds__(GS_THREADS) void k_gs_dots(const SegBlk* __restrict__ sb, int k,
                                                       const T* __restrict__ V, int64_t ldv,
                                                       const T* __restrict__ w,
                                                       double* __restrict__ partial,
                                                       const int32_t* __restrict__ active) {
  constexpr int N = VEC ? Row16<T>::N : 1;
  const SegBlk b = sb[blockIdx.x];
  double* out = partial + (int64_t)blockIdx.x * MDOT_K;
  if (active && !active[b.comp]) {  // finished component (uniform over the block)
    if (threadIdx.x < k) out[threadIdx.x] = 0.0;
    return;
  }
  double acc[KC];
#pragma unroll
  for (int j = 0; j < KC; ++j) acc[j] = 0.0;
  int64_t i = b.beg + N * threadIdx.x;
#pragma unroll 2
  for (; i + N <= b.end; i += N * GS_THREADS) gs_dots_rows<T, KC, N>(k, V, ldv, w, i, acc);
  if constexpr (N > 1)
    for (; i < b.end; ++i) gs_dots_rows<T, KC, 1>(k, V, ldv, w, i, acc);  // ragged end
  block_store_cols<KC>(acc, k, out);
}

// w -= sum_j coef[comp][j] V_j over the block, then MODE 1: partial[blk][j] = <V_j, w>;
// MODE 2: partial[blk][0] = <w, w>.
template <typename T, int KC, int MODE, bool VEC>
__global__ __launch_bounds__(GS_THREADS) void k_gs_update(const SegBlk* __restrict__ sb, int k,
                                                         const T* __restrict__ V, int64_t ldv,
                                                         const double* __restrict__ coef,
                                                         T* __restrict__ w,
                                                         double* __restrict__ partial,
                                                         const int32_t* __restrict__ active) {
  constexpr int N = VEC ? Row16<T>::N : 1;
  constexpr int NO = MODE == 1 ? KC : 1;
  const int kout = MODE == 1 ? k : 1;
  const SegBlk b = sb[blockIdx.x];
  double* out = partial + (int64_t)blockIdx.x * MDOT_K;
  if (active && !active[b.comp]) {
    if (threadIdx.x < kout) out[threadIdx.x] = 0.0;
    return;
  }
  __shared__ T cf[KC];
  if (threadIdx.x < KC) cf[threadIdx.x] = threadIdx.x < k ? (T)coef[b.comp * MDOT_K + threadIdx.x] : T(0);
  __syncthreads();
  double acc[NO];
#pragma unroll
  for (int j = 0; j < NO; ++j) acc[j] = 0.0;
  int64_t i = b.beg + N * threadIdx.x;
#pragma unroll 1
  for (; i + N <= b.end; i += N * GS_THREADS)
    gs_update_rows<T, KC, MODE, N, NO>(k, V, ldv, cf, w, i, acc);
  if constexpr (N > 1)
    for (; i < b.end; ++i) gs_update_rows<T, KC, MODE, 1, NO>(k, V, ldv, cf, w, i, acc);
  block_store_cols<NO>(acc, kout, out);
}

// One work-group per component sums its wide-block partials, column j by 32 lane groups
// in a fixed order and then across the groups.  mode 0: h = coef = sums; 1: h += sums,
// coef = sums; 2: hn[c] = sqrt(sum of column 0).  Inactive components: coef 0 (h kept),
// hn 0.
__global__ __launch_bounds__(GS_FINAL) void k_gs_final(const int32_t* __restrict__ comp_wblk,
                                                       const double* __restrict__ partial,
                                                       int k, int mode, double* __restrict__ h,
                                                       int64_t hs, double* __restrict__ coef,
                                                       double* __restrict__ hn,
                                                       const int32_t* __restrict__ active) {
  constexpr int G = GS_FINAL / MDOT_K;  // lane groups
  const int c = blockIdx.x;
  const int j = threadIdx.x % MDOT_K, grp = threadIdx.x / MDOT_K;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (j < k) {
    int bb = comp_wblk[c] + grp;
    const int e = comp_wblk[c + 1];
    for (; bb + 3 * G < e; bb += 4 * G) {
      a0 += partial[(int64_t)bb * MDOT_K + j];
      a1 += partial[(int64_t)(bb + G) * MDOT_K + j];
      a2 += partial[(int64_t)(bb + 2 * G) * MDOT_K + j];
      a3 += partial[(int64_t)(bb + 3 * G) * MDOT_K + j];
    }
    for (; bb < e; bb += G) a0 += partial[(int64_t)bb * MDOT_K + j];
  }
  __shared__ double sh[G][MDOT_K + 1];
  sh[grp][j] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
    for (int q = 0; q < G; ++q) t += sh[q][threadIdx.x];
    const bool on = !active || active[c] != 0;
    if (mode == 2) {
      hn[c] = on ? sqrt(t) : 0.0;
    } else {
      const double v = on ? t : 0.0;
      coef[c * MDOT_K + threadIdx.x] = v;
      if (on) h[c * hs + threadIdx.x] = mode == 1 ? h[c * hs + threadIdx.x] + v : v;
    }
  }
}

// First stage of the two-stage partial sums: super-block sb sums the partials of its wide
// blocks [sblk[sb], sblk[sb + 1]) column by column in a fixed order.  (A single work-group
// per component walking thousands of partials is latency-bound: ~100 us at 5,000 wide
// blocks.)
__global__ __launch_bounds__(256) void k_gs_stage1(const int32_t* __restrict__ sblk,
                                                  const double* __restrict__ partial, int k,
                                                  double* __restrict__ spartial) {
  constexpr int G = 256 / MDOT_K;
  const int sb = blockIdx.x;
  const int j = threadIdx.x % MDOT_K, grp = threadIdx.x / MDOT_K;
  double a = 0.0;
  if (j < k)
    for (int bb = sblk[sb] + grp; bb < sblk[sb + 1]; bb += G) a += partial[(int64_t)bb * MDOT_K + j];
  __shared__ double sh[G][MDOT_K + 1];
  sh[grp][j] = a;
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
    for (int q = 0; q < G; ++q) t += sh[q][threadIdx.x];
    spartial[(int64_t)sb * MDOT_K + threadIdx.x] = t;
  }
}

// The final sums of the Gram-Schmidt partials, in one or two stages
static void gs_final(const SegDev& sg, int k, int mode, double* h, int64_t hs, double* coef,
                     double* hn, const int32_t* active, hipStream_t s) {
  if (sg.two_stage) {
    hipLaunchKernelGGL(k_gs_stage1, dim3(sg.nsblk), dim3(256), 0, s, sg.sblk,
                       (const double*)sg.partial, k, sg.spartial);
    hipLaunchKernelGGL(k_gs_final, dim3(sg.ncomp), dim3(GS_FINAL), 0, s, sg.comp_sblk,
                       (const double*)sg.spartial, k, mode, h, hs, coef, hn, active);
  } else {
    hipLaunchKernelGGL(k_gs_final, dim3(sg.ncomp), dim3(GS_FINAL), 0, s, sg.comp_wblk,
                       (const double*)sg.partial, k, mode, h, hs, coef, hn, active);
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
  // an inactive component gets exact zeros (also where x is not finite)
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x)
    y[i] = on ? (Tout)((double)x[i] * f) : (Tout)0;
}

// y64 += a[c] * x
template <typename Tin>
__global__ __launch_bounds__(RED_THREADS) void k_sadd(const SegBlk* __restrict__ sb,
                                                      const double* __restrict__ a,
                                                      const Tin* __restrict__ x,
                                                      double* __restrict__ y) {
  const SegBlk b = sb[blockIdx.x];
  const double av = a ? a[b.comp] : 1.0;
  if (!isfinite(av)) return;  // a component isolated for non-finite data keeps its y
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
// thr[c] (or whose basis breaks down) turns inactive.  One work-group: nact[0] and
// status[0] count the still-active components, status[1 + c] = res[c] / thr[c] for them.
constexpr int GIV_THREADS = 256;
__global__ __launch_bounds__(GIV_THREADS) void k_sgivens(
    int ncomp, int j, int ldh, int64_t hs, double* __restrict__ H, const double* __restrict__ hn,
    double* __restrict__ cs, double* __restrict__ sn, int64_t rs, double* __restrict__ g,
    int64_t gs, double* __restrict__ res, const double* __restrict__ thr,
    int32_t* __restrict__ active, int32_t* __restrict__ kk, int32_t* __restrict__ nact,
    double* __restrict__ status) {
  int cnt = 0;
  for (int c = threadIdx.x; c < ncomp; c += GIV_THREADS) {
    double ratio = 0.0;
    if (active[c]) {
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
      if (stop) {
        active[c] = 0;
      } else {
        ++cnt;
        ratio = res[c] / thr[c];
      }
    }
    status[1 + c] = ratio;
  }
  __shared__ int sh[GIV_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < GIV_THREADS / 64; ++w) t += sh[w];
    nact[0] = t;
    status[0] = (double)t;
  }
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

// ---------------------------------------------------------------------------------------
// MG-preconditioned conjugate gradients (z02/volume_mf_diff.py:1875-1931), every component
// at once; a component's scalars (rz = r.z, pap = p.Ap, rr = r.r) live in device arrays,
// and finished components (active[c] = 0) are left untouched by the vector updates.
// x += alpha p, r -= alpha Ap with alpha = rz / pap
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_cg_xr(const SegBlk* __restrict__ sb,
                                                       const double* __restrict__ rz,
                                                       const double* __restrict__ pap,
                                                       const int32_t* __restrict__ active,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ ap,
                                                       T* __restrict__ x, T* __restrict__ r) {
  const SegBlk b = sb[blockIdx.x];
  if (!active[b.comp]) return;
  const double a = rz[b.comp] / pap[b.comp];
  if (!isfinite(a)) return;  // breakdown: k_cg_step stops the component
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x) {
    x[i] = (T)((double)x[i] + a * (double)p[i]);
    r[i] = (T)((double)r[i] - a * (double)ap[i]);
  }
}

// p = z + beta p with beta = rznew / rz
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_cg_p(const SegBlk* __restrict__ sb,
                                                      const double* __restrict__ rznew,
                                                      const double* __restrict__ rz,
                                                      const int32_t* __restrict__ active,
                                                      const T* __restrict__ z,
                                                      T* __restrict__ p) {
  const SegBlk b = sb[blockIdx.x];
  if (!active[b.comp]) return;
  const double beta = rznew[b.comp] / rz[b.comp];
  if (!isfinite(beta)) return;
  for (int64_t i = b.beg + threadIdx.x; i < b.end; i += blockDim.x)
    p[i] = (T)((double)z[i] + beta * (double)p[i]);
}

// End of an iteration, one work-group: its[c] += 1, rz[c] = rznew[c], and the component
// stays active while sqrt(rr[c]) > thr[c], its[c] < max_its and alpha was finite.
// status[0] = the number of components still active.
__global__ __launch_bounds__(GIV_THREADS) void k_cg_step(
    int ncomp, const double* __restrict__ rr, double* __restrict__ rz,
    const double* __restrict__ rznew, const double* __restrict__ pap,
    const double* __restrict__ thr, int max_its, int32_t* __restrict__ its,
    int32_t* __restrict__ active, double* __restrict__ status) {
  int cnt = 0;
  for (int c = threadIdx.x; c < ncomp; c += blockDim.x) {
    if (!active[c]) continue;
    const double a = rz[c] / pap[c];
    const int k = its[c] + 1;
    its[c] = k;
    rz[c] = rznew[c];
    const bool on = isfinite(a) && sqrt(rr[c]) > thr[c] && k < max_its;
    active[c] = on ? 1 : 0;
    cnt += on;
  }
  __shared__ int sh[GIV_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < GIV_THREADS / 64; ++w) t += sh[w];
    status[0] = (double)t;
  }
}

// Start: active[c] = want[c] and sqrt(rr[c]) > thr[c] = tol[c]; its[c] = 0.
__global__ void k_cg_start(int ncomp, const double* __restrict__ rr,
                           const double* __restrict__ tol, const int32_t* __restrict__ want,
                           double* __restrict__ thr, int32_t* __restrict__ its,
                           int32_t* __restrict__ active) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncomp) return;
  thr[c] = tol[c];
  its[c] = 0;
  active[c] = (want[c] && sqrt(rr[c]) > tol[c]) ? 1 : 0;
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

// ---------------------------------------------------------------------------------------
// Allocation-free dot product: the partial sums live in `scratch`, a fixed device array of
// this library (one copy per GPU), instead of a buffer from hipMalloc.  Block b sums
// values [b*NOALLOC_CHUNK, (b+1)*NOALLOC_CHUNK) in a fixed order; one block then sums the
// partials in a fixed order.  Every call on a GPU uses that GPU's array, so calls that
// overlap on one GPU (several streams or threads) race.
constexpr int NOALLOC_BLOCKS = 32768;
constexpr int64_t NOALLOC_CHUNK = 8192;
static __device__ __align__(256) double scratch[NOALLOC_BLOCKS];

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_dot_chunk(int64_t n, const T* __restrict__ x,
                                                           const T* __restrict__ y,
                                                           double* __restrict__ partial) {
  const int64_t lo = (int64_t)blockIdx.x * NOALLOC_CHUNK;
  const int64_t hi = lo + NOALLOC_CHUNK < n ? lo + NOALLOC_CHUNK : n;
  double s = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += RED_THREADS) s += (double)x[i] * (double)y[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// device address of `scratch` on each GPU, looked up once per device
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {};

static int noalloc_scratch(double** out) {
  int dev = -1;
  DGMI_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("device index out of the supported range (0..63)");
    return -2;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    DGMI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("hipGetSymbolAddress returned no address for the dot scratch");
      return -3;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<double*>(g_scratch_ptr[dev]);
  return 0;
}

int64_t dot_noalloc_max_n() { return (int64_t)NOALLOC_BLOCKS * NOALLOC_CHUNK; }

template <typename T>
int launch_dot_noalloc(int64_t n, const T* x, const T* y, double* out, hipStream_t s) {
  if (n < 0 || n > dot_noalloc_max_n()) {
    set_error("dgmi_dot_noalloc: n = " + std::to_string(n) + " is outside 0.." +
              std::to_string(dot_noalloc_max_n()) + " (" + std::to_string(NOALLOC_BLOCKS) +
              " partial sums of " + std::to_string(NOALLOC_CHUNK) + " values)");
    return -2;
  }
  double* part = nullptr;
  if (int rc = noalloc_scratch(&part)) return rc;
  const int nb = n == 0 ? 1 : (int)((n + NOALLOC_CHUNK - 1) / NOALLOC_CHUNK);
  hipLaunchKernelGGL((k_dot_chunk<T>), dim3(nb), dim3(RED_THREADS), 0, s, n, x, y, part);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s, (const double*)part, nb, out,
                     0);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template int launch_dot_noalloc<double>(int64_t, const double*, const double*, double*,
                                        hipStream_t);
template int launch_dot_noalloc<float>(int64_t, const float*, const float*, double*, hipStream_t);

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

// 16-byte row groups are aligned in every wide block of V's columns and w
template <typename T>
static int gs_vec(const SegDev& sg, const T* V, int64_t ldv, const T* w) {
  constexpr int N = Row16<T>::N;
  return sg.wvec4 && ldv % N == 0 && reinterpret_cast<uintptr_t>(V) % 16 == 0 &&
         reinterpret_cast<uintptr_t>(w) % 16 == 0;
}

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
  const bool vec = gs_vec(sg, V, ldv, w);
  for (int c0 = 0; c0 < k; c0 += MDOT_K) {
    const int kc = std::min(MDOT_K, k - c0);
    const T* Vc = V + (int64_t)c0 * ldv;
    if (vec && kc <= 4)
      hipLaunchKernelGGL((k_gs_dots<T, 4, true>), dim3(sg.nwblk), dim3(GS_THREADS), 0, s, sg.wblk,
                         kc, Vc, ldv, (const T*)w, sg.partial, active);
    else if (vec)
      hipLaunchKernelGGL((k_gs_dots<T, MDOT_K, true>), dim3(sg.nwblk), dim3(GS_THREADS), 0, s,
                         sg.wblk, kc, Vc, ldv, (const T*)w, sg.partial, active);
    else
      hipLaunchKernelGGL((k_gs_dots<T, MDOT_K, false>), dim3(sg.nwblk), dim3(GS_THREADS), 0, s,
                         sg.wblk, kc, Vc, ldv, (const T*)w, sg.partial, active);
    gs_final(sg, kc, accumulate ? 1 : 0, h + c0, hs, sg.coef, nullptr, active, s);
    hipLaunchKernelGGL((k_smaxpy<T>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, kc, Vc,
                       ldv, (const double*)sg.coef, w, active);
  }
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T, int KC, bool VEC>
static void launch_cgs2(const SegDev& sg, int k, const T* V, int64_t ldv, T* w, double* h,
                        int64_t hs, double* hn, const int32_t* active, hipStream_t s) {
  const dim3 g(sg.nwblk), t(GS_THREADS);
  hipLaunchKernelGGL((k_gs_dots<T, KC, VEC>), g, t, 0, s, sg.wblk, k, V, ldv, (const T*)w,
                     sg.partial, active);
  gs_final(sg, k, 0, h, hs, sg.coef, nullptr, active, s);
  hipLaunchKernelGGL((k_gs_update<T, KC, 1, VEC>), g, t, 0, s, sg.wblk, k, V, ldv,
                     (const double*)sg.coef, w, sg.partial, active);
  gs_final(sg, k, 1, h, hs, sg.coef, nullptr, active, s);
  hipLaunchKernelGGL((k_gs_update<T, KC, 2, VEC>), g, t, 0, s, sg.wblk, k, V, ldv,
                     (const double*)sg.coef, w, sg.partial, active);
  gs_final(sg, 1, 2, nullptr, 0, sg.coef, hn, active, s);
}

template <typename T>
int sorth_cgs2(const SegDev& sg, int k, const T* V, int64_t ldv, T* w, double* h, int64_t hs,
               double* hn, const int32_t* active, hipStream_t s) {
  if (k < 1 || k > MDOT_K) {
    set_error("fused CGS2 takes 1.." + std::to_string(MDOT_K) + " vectors");
    return -2;
  }
  if (!gs_vec(sg, V, ldv, w)) {
    launch_cgs2<T, MDOT_K, false>(sg, k, V, ldv, w, h, hs, hn, active, s);
  } else {
    // columns rounded up to a multiple of 4
    switch ((k + 3) / 4) {
      case 1: launch_cgs2<T, 4, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
      case 2: launch_cgs2<T, 8, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
      case 3: launch_cgs2<T, 12, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
      case 4: launch_cgs2<T, 16, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
      case 5: launch_cgs2<T, 20, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
      case 6: launch_cgs2<T, 24, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
      case 7: launch_cgs2<T, 28, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
      default: launch_cgs2<T, 32, true>(sg, k, V, ldv, w, h, hs, hn, active, s); break;
    }
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
            int32_t* active, int32_t* kk, int32_t* nact, double* status, hipStream_t s) {
  hipLaunchKernelGGL(k_sgivens, dim3(1), dim3(GIV_THREADS), 0, s, ncomp, j, ldh, hs, H, hn, cs,
                     sn, rs, g, gs, res, thr, active, kk, nact, status);
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

int cg_start(int ncomp, const double* rr, const double* tol, const int32_t* want, double* thr,
             int32_t* its, int32_t* active, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_start, dim3((ncomp + 63) / 64), dim3(64), 0, s, ncomp, rr, tol, want,
                     thr, its, active);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int cg_xr(const SegDev& sg, const double* rz, const double* pap, const int32_t* active,
          const T* p, const T* ap, T* x, T* r, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_xr<T>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, rz, pap,
                     active, p, ap, x, r);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int cg_p(const SegDev& sg, const double* rznew, const double* rz, const int32_t* active,
         const T* z, T* p, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_p<T>), dim3(sg.nblk), dim3(RED_THREADS), 0, s, sg.blk, rznew, rz,
                     active, z, p);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int cg_step(int ncomp, const double* rr, double* rz, const double* rznew, const double* pap,
            const double* thr, int max_its, int32_t* its, int32_t* active, double* status,
            hipStream_t s) {
  hipLaunchKernelGGL(k_cg_step, dim3(1), dim3(GIV_THREADS), 0, s, ncomp, rr, rz, rznew, pap, thr,
                     max_its, its, active, status);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template int cg_xr<double>(const SegDev&, const double*, const double*, const int32_t*,
                           const double*, const double*, double*, double*, hipStream_t);
template int cg_p<double>(const SegDev&, const double*, const double*, const int32_t*,
                          const double*, double*, hipStream_t);
template int sdot<double>(const SegDev&, const double*, const double*, double*, double*,
                          hipStream_t, const int32_t*);
template int sdot<float>(const SegDev&, const float*, const float*, double*, double*, hipStream_t,
                         const int32_t*);
template int sorth_pass<double>(const SegDev&, int, const double*, int64_t, double*, double*,
                                int64_t, int, const int32_t*, hipStream_t);
template int sorth_pass<float>(const SegDev&, int, const float*, int64_t, float*, double*,
                               int64_t, int, const int32_t*, hipStream_t);
template int sorth_cgs2<double>(const SegDev&, int, const double*, int64_t, double*, double*,
                                int64_t, double*, const int32_t*, hipStream_t);
template int sorth_cgs2<float>(const SegDev&, int, const float*, int64_t, float*, double*,
                               int64_t, double*, const int32_t*, hipStream_t);
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
