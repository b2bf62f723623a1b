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

// y += sgn * (*a) x  (a on the device)
template <typename T>
__global__ void k_axpy_dev(int64_t n, const double* __restrict__ a, double sgn,
                           const T* __restrict__ x, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double av = sgn * a[0];
  if (i < n) y[i] = (T)((double)y[i] + av * (double)x[i]);
}

// y = x / (*a)
template <typename T>
__global__ void k_scale_dev(int64_t n, const double* __restrict__ a, const T* __restrict__ x,
                            T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double inv = 1.0 / a[0];
  if (i < n) y[i] = (T)((double)x[i] * inv);
}

template <typename Tin, typename Tout>
__global__ void k_cast(int64_t n, const Tin* __restrict__ x, Tout* __restrict__ y, double scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (Tout)((double)x[i] * scale);
}

// y += scale * x (x of another precision): the refinement update x64 += d32
template <typename Tin>
__global__ void k_add_cast(int64_t n, const Tin* __restrict__ x, double* __restrict__ y,
                           const double* __restrict__ scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double s = scale ? scale[0] : 1.0;
  if (i < n) y[i] += s * (double)x[i];
}

template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ y, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// x += sum_j y[j] V[j]  (V rows of length n, k rows)
template <typename T>
__global__ void k_combine(int64_t n, int k, const T* __restrict__ V, int64_t ldv,
                          const double* __restrict__ yv, T* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = (double)x[i];
  for (int j = 0; j < k; ++j) s += yv[j] * (double)V[(int64_t)j * ldv + i];
  x[i] = (T)s;
}

// Givens step for column j of the Hessenberg matrix (Saad, Alg. 6.9 with rotations):
// H column-major (m+1) x m in `H`; cs/sn rotations; g right-hand side.  Applies the j
// previous rotations to column j, forms rotation j, updates g, stores |g[j+1]| in res.
__global__ void k_givens(int j, int ldh, double* __restrict__ H, double* __restrict__ cs,
                         double* __restrict__ sn, double* __restrict__ g,
                         double* __restrict__ res) {
  if (threadIdx.x != 0) return;
  double* h = H + (int64_t)j * ldh;
  for (int i = 0; i < j; ++i) {
    const double t = cs[i] * h[i] + sn[i] * h[i + 1];
    h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
    h[i] = t;
  }
  const double a = h[j], b = h[j + 1];
  const double r = hypot(a, b);
  const double c = r > 0 ? a / r : 1.0, s = r > 0 ? b / r : 0.0;
  cs[j] = c;
  sn[j] = s;
  h[j] = r;
  h[j + 1] = 0.0;
  g[j + 1] = -s * g[j];
  g[j] = c * g[j];
  res[0] = fabs(g[j + 1]);
}

// back substitution R y = g for the leading k x k block; y into yv
__global__ void k_backsolve(int k, int ldh, const double* __restrict__ H,
                            const double* __restrict__ g, double* __restrict__ yv) {
  if (threadIdx.x != 0) return;
  for (int i = k - 1; i >= 0; --i) {
    double s = g[i];
    for (int c = i + 1; c < k; ++c) s -= H[(int64_t)c * ldh + i] * yv[c];
    yv[i] = s / H[(int64_t)i * ldh + i];
  }
}

// H[i, j] = hh[0] ; used to move a reduced dot into the Hessenberg column
__global__ void k_store_h(const double* __restrict__ src, double* __restrict__ dst, int mode) {
  if (threadIdx.x == 0) dst[0] = mode == 1 ? sqrt(src[0]) : src[0];
}

__global__ void k_set_g(double* __restrict__ g, int len, const double* __restrict__ beta) {
  for (int i = threadIdx.x; i < len; i += blockDim.x) g[i] = i == 0 ? beta[1] : 0.0;
}

// ---------------------------------------------------------------------------------------
// Fused classical Gram-Schmidt pass over up to MDOT_K basis vectors:
//   h[j] = <w, V_j> (one read of w for all j), then w -= sum_j h[j] V_j.
constexpr int MDOT_K = 32;
constexpr int MDOT_BLOCKS = 2048;

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_mdot_partial(int64_t n, int k,
                                                              const T* __restrict__ V,
                                                              int64_t ldv,
                                                              const T* __restrict__ w,
                                                              double* __restrict__ partial) {
  double acc[MDOT_K];
#pragma unroll
  for (int j = 0; j < MDOT_K; ++j) acc[j] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
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
    double s = 0.0;
    for (int ww = 0; ww < RED_THREADS / 64; ++ww) s += sh[ww][threadIdx.x];
    partial[(int64_t)blockIdx.x * MDOT_K + threadIdx.x] = s;
  }
}

// h_out[j] (+)= s_j and tmp[j] = s_j, s_j = sum over blocks of partial[b][j]
__global__ __launch_bounds__(RED_THREADS) void k_mdot_final(const double* __restrict__ partial,
                                                            int nb, int k,
                                                            double* __restrict__ h_out,
                                                            double* __restrict__ tmp,
                                                            int accumulate) {
  for (int j = 0; j < k; ++j) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[(int64_t)b * MDOT_K + j];
    s = block_sum(s);
    if (threadIdx.x == 0) {
      tmp[j] = s;
      h_out[j] = accumulate ? h_out[j] + s : s;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void k_maxpy(int64_t n, int k, const T* __restrict__ V, int64_t ldv,
                        const double* __restrict__ h, T* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = (double)w[i];
  for (int j = 0; j < k; ++j) s -= h[j] * (double)V[(int64_t)j * ldv + i];
  w[i] = (T)s;
}

// ---------------------------------------------------------------------------------------
static double* g_partial = nullptr;
static double* g_mpartial = nullptr;
static double* g_mtmp = nullptr;

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
// One classical Gram-Schmidt pass of w against V_0..V_{k-1}; writes (or, with
// accumulate, adds) the coefficients into h.  Chunks of MDOT_K vectors.
template <typename T>
int cgs_pass(int64_t n, int k, const T* V, int64_t ldv, T* w, double* h, int accumulate,
             hipStream_t s) {
  if (!g_mpartial) DGMI_HIP(hipMalloc(&g_mpartial, sizeof(double) * MDOT_BLOCKS * MDOT_K));
  if (!g_mtmp) DGMI_HIP(hipMalloc(&g_mtmp, sizeof(double) * MDOT_K));
  const int nb = (int)std::min<int64_t>(MDOT_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS);
  const dim3 grd((unsigned)((n + 255) / 256));
  // within a chunk all dots use the same w (classical GS); successive chunks see the
  // updated w (block-modified GS), which only matters for restart lengths above MDOT_K
  for (int c0 = 0; c0 < k; c0 += MDOT_K) {
    const int kc = std::min(MDOT_K, k - c0);
    hipLaunchKernelGGL((k_mdot_partial<T>), dim3(nb), dim3(RED_THREADS), 0, s, n, kc,
                       V + (int64_t)c0 * ldv, ldv, w, g_mpartial);
    hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(RED_THREADS), 0, s, g_mpartial, nb, kc,
                       h + c0, g_mtmp, accumulate);
    hipLaunchKernelGGL((k_maxpy<T>), grd, dim3(256), 0, s, n, kc, V + (int64_t)c0 * ldv, ldv,
                       (const double*)g_mtmp, w);
  }
  DGMI_HIP(hipGetLastError());
  return 0;
}
template int cgs_pass<double>(int64_t, int, const double*, int64_t, double*, double*, int,
                              hipStream_t);
template int cgs_pass<float>(int64_t, int, const float*, int64_t, float*, double*, int,
                             hipStream_t);

template <typename T>
int axpy_dev(int64_t n, const double* a, double sgn, const T* x, T* y, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_axpy_dev<T>), grd, blk, 0, s, n, a, sgn, x, y);
  DGMI_HIP(hipGetLastError());
  return 0;
}
template <typename T>
int scale_dev(int64_t n, const double* a, const T* x, T* y, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_scale_dev<T>), grd, blk, 0, s, n, a, x, y);
  DGMI_HIP(hipGetLastError());
  return 0;
}
template <typename Tin, typename Tout>
int cast(int64_t n, const Tin* x, Tout* y, double scale, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_cast<Tin, Tout>), grd, blk, 0, s, n, x, y, scale);
  DGMI_HIP(hipGetLastError());
  return 0;
}
template <typename Tin>
int add_cast(int64_t n, const Tin* x, double* y, const double* scale, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_add_cast<Tin>), grd, blk, 0, s, n, x, y, scale);
  DGMI_HIP(hipGetLastError());
  return 0;
}
template <typename T>
int fill(int64_t n, T* y, T v, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_fill<T>), grd, blk, 0, s, n, y, v);
  DGMI_HIP(hipGetLastError());
  return 0;
}
template <typename T>
int combine(int64_t n, int k, const T* V, int64_t ldv, const double* y, T* x, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_combine<T>), grd, blk, 0, s, n, k, V, ldv, y, x);
  DGMI_HIP(hipGetLastError());
  return 0;
}
int givens(int j, int ldh, double* H, double* cs, double* sn, double* g, double* res,
           hipStream_t s) {
  hipLaunchKernelGGL(k_givens, dim3(1), dim3(64), 0, s, j, ldh, H, cs, sn, g, res);
  DGMI_HIP(hipGetLastError());
  return 0;
}
int backsolve(int k, int ldh, const double* H, const double* g, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(64), 0, s, k, ldh, H, g, y);
  DGMI_HIP(hipGetLastError());
  return 0;
}
int store_h(const double* src, double* dst, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_store_h, dim3(1), dim3(64), 0, s, src, dst, mode);
  DGMI_HIP(hipGetLastError());
  return 0;
}
int set_g(double* g, int len, const double* beta, hipStream_t s) {
  hipLaunchKernelGGL(k_set_g, dim3(1), dim3(128), 0, s, g, len, beta);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template int axpy_dev<double>(int64_t, const double*, double, const double*, double*, hipStream_t);
template int axpy_dev<float>(int64_t, const double*, double, const float*, float*, hipStream_t);
template int scale_dev<double>(int64_t, const double*, const double*, double*, hipStream_t);
template int scale_dev<float>(int64_t, const double*, const float*, float*, hipStream_t);
template int cast<double, float>(int64_t, const double*, float*, double, hipStream_t);
template int cast<float, double>(int64_t, const float*, double*, double, hipStream_t);
template int cast<double, double>(int64_t, const double*, double*, double, hipStream_t);
template int cast<float, float>(int64_t, const float*, float*, double, hipStream_t);
template int add_cast<float>(int64_t, const float*, double*, const double*, hipStream_t);
template int add_cast<double>(int64_t, const double*, double*, const double*, hipStream_t);
template int fill<double>(int64_t, double*, double, hipStream_t);
template int fill<float>(int64_t, float*, float, hipStream_t);
template int combine<double>(int64_t, int, const double*, int64_t, const double*, double*, hipStream_t);
template int combine<float>(int64_t, int, const float*, int64_t, const double*, float*, hipStream_t);
}  // namespace kry

}  // namespace dgmi
