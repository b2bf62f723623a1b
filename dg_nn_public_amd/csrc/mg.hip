// DG <-> P1CG transfers, coloured probing of RAR, and the SFC V-cycle (gfx950).
//
// P1CG vectors live in SFC order (position p = sfc[node] - 1), so the reference's
// per-V-cycle argsort/permutation (z02/volume_mf_diff.py:1992-2007) is folded into the
// transfer kernels' index lists.  The 1-D SFC levels are CSR matrices; level l+1
// aggregates positions (2I, 2I+1) of level l.
#include <mutex>

#include "dgmi_internal.h"

namespace dgmi {

// P1DG -> P3DG element interpolation (barycentric coordinates of the P3 nodes):
// z02/mesh_init.py:610-631 (3D) and :224-235 (2D).
__constant__ double c_P3[20 * 4] = {
    1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1,
    2.0 / 3, 0, 1.0 / 3, 0, 1.0 / 3, 0, 2.0 / 3, 0, 2.0 / 3, 1.0 / 3, 0, 0,
    1.0 / 3, 2.0 / 3, 0, 0, 0, 2.0 / 3, 1.0 / 3, 0, 0, 1.0 / 3, 2.0 / 3, 0,
    2.0 / 3, 0, 0, 1.0 / 3, 1.0 / 3, 0, 0, 2.0 / 3, 0, 2.0 / 3, 0, 1.0 / 3,
    0, 1.0 / 3, 0, 2.0 / 3, 0, 0, 2.0 / 3, 1.0 / 3, 0, 0, 1.0 / 3, 2.0 / 3,
    0, 1.0 / 3, 1.0 / 3, 1.0 / 3, 1.0 / 3, 1.0 / 3, 1.0 / 3, 0, 1.0 / 3, 0, 1.0 / 3, 1.0 / 3,
    1.0 / 3, 1.0 / 3, 0, 1.0 / 3};
__constant__ double c_P2[10 * 3] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 2.0 / 3, 1.0 / 3, 0,
                                    1.0 / 3, 2.0 / 3, 0, 0, 2.0 / 3, 1.0 / 3, 0, 1.0 / 3, 2.0 / 3,
                                    1.0 / 3, 0, 2.0 / 3, 2.0 / 3, 0, 1.0 / 3, 1.0 / 3, 1.0 / 3,
                                    1.0 / 3};

template <int D> __device__ __forceinline__ double Ptab(int n, int k) {
  if constexpr (D == 3) return c_P3[n * 4 + k];
  else return c_P2[n * 3 + k];
}

// r_sfc[p] = (1/mult) sum over the DG copies (e,k) of vertex p of sum_n P[n][k] r[e][n]
template <typename T, int D>
__global__ __launch_bounds__(256) void k_restrict(int ncg, const int32_t* __restrict__ rptr,
                                                  const int32_t* __restrict__ rent,
                                                  const T* __restrict__ rw,
                                                  const T* __restrict__ r, T* __restrict__ out) {
  constexpr int NL = D == 3 ? 20 : 10, NB = D + 1;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncg) return;
  T s = 0;
  for (int q = rptr[p]; q < rptr[p + 1]; ++q) {
    const int idx = rent[q];
    const int e = idx / NB, k = idx - (idx / NB) * NB;
    const T* re = r + (size_t)e * NL;
    T t = 0;
#pragma unroll
    for (int n = 0; n < NL; ++n) t += (T)Ptab<D>(n, k) * re[n];
    s += t;
  }
  out[p] = s * rw[p];
}

// x[e][n] += sum_k P[n][k] e_sfc[pos(e,k)] / mult(pos(e,k))
template <typename T, int D>
__global__ __launch_bounds__(256) void k_prolong_add(int nele, const int32_t* __restrict__ cellpos,
                                                     const T* __restrict__ rw,
                                                     const T* __restrict__ ec, T* __restrict__ x) {
  constexpr int NL = D == 3 ? 20 : 10, NB = D + 1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele) return;
  T ev[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int p = cellpos[e * NB + k];
    ev[k] = ec[p] * rw[p];
  }
  T* xe = x + (size_t)e * NL;
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    T s = xe[n];
#pragma unroll
    for (int k = 0; k < NB; ++k) s += (T)Ptab<D>(n, k) * ev[k];
    xe[n] = s;
  }
}

// probe vector for colour c: x = P I_fc mask_c (z02/volume_mf_diff.py:1717-1722)
template <int D>
__global__ __launch_bounds__(256) void k_probe_mask(int nele, const int32_t* __restrict__ cellpos,
                                                    const double* __restrict__ rw,
                                                    const int32_t* __restrict__ whichc, int color,
                                                    double* __restrict__ x) {
  constexpr int NL = D == 3 ? 20 : 10, NB = D + 1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele) return;
  double ev[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int p = cellpos[e * NB + k];
    ev[k] = whichc[p] == color ? rw[p] : 0.0;
  }
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) s += Ptab<D>(n, k) * ev[k];
    x[(size_t)e * NL + n] = s;
  }
}

// values[spidx[k]] = -r_sfc[rowpos[k]]  (r_sfc holds R(-A R m))
__global__ void k_probe_collect(int64_t nent, const int32_t* __restrict__ spidx,
                                const int32_t* __restrict__ rowpos,
                                const double* __restrict__ r, double* __restrict__ values) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nent) values[spidx[k]] = -r[rowpos[k]];
}

// ---------------------------------------------------------------------------------------
// SFC level kernels.  Jacobi sweep (z02/multigrid_linearelastic.py:25-39):
//   eout = ein + omega (b - A ein) / diag
// With `ec` set, the input is ein + P ec (nearest prolongation, coarse index i>>1), which
// fuses the up-cycle's e_l += Upsample(e_{l+1}) (:163-164) into the first post-sweep.
// Grid kernels give each row a group of ROW_LANES lanes: the lanes load the row's entries
// side by side (rows hold 14-32 entries), so a sweep costs one round of dependent loads
// instead of one per entry, and the group sums its partial products with shuffles.
constexpr int ROW_LANES = 16;

template <typename T>
__device__ __forceinline__ T level_row_group(int i, int sub, const int32_t* __restrict__ rp,
                                             const int32_t* __restrict__ col,
                                             const T* __restrict__ val,
                                             const T* __restrict__ ein,
                                             const T* __restrict__ ec) {
  T s = 0;
  const int q1 = rp[i + 1];
  for (int q = rp[i] + sub; q < q1; q += ROW_LANES) {
    const int j = col[q];
    T ej = ein[j];
    if (ec) ej += ec[j >> 1];
    s += val[q] * ej;
  }
#pragma unroll
  for (int off = ROW_LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sweep(int n, const int32_t* __restrict__ rp,
                                               const int32_t* __restrict__ col,
                                               const T* __restrict__ val,
                                               const T* __restrict__ diag,
                                               const T* __restrict__ b, const T* __restrict__ ein,
                                               const T* __restrict__ ec, T* __restrict__ eout,
                                               T omega) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / ROW_LANES, sub = t % ROW_LANES;
  if (i >= n) return;  // whole groups leave together
  const T ax = level_row_group<T>(i, sub, rp, col, val, ein, ec);
  if (sub == 0) {
    T ei = ein[i];
    if (ec) ei += ec[i >> 1];
    eout[i] = ei + omega * (b[i] - ax) / diag[i];
  }
}

// Last pre-sweep of level l, fused with the restriction to level l+1: the reference
// evaluates rr = b - A e and applies e += omega rr/diag in the same call (:144-146), then
// restricts rr and e by pairwise sums (:147-153).  Two lane groups per coarse node, one
// per fine row of its pair.
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_restrict(int n, const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ col,
                                                        const T* __restrict__ val,
                                                        const T* __restrict__ diag,
                                                        const T* __restrict__ b,
                                                        const T* __restrict__ ein,
                                                        T* __restrict__ eout, T* __restrict__ rc,
                                                        T* __restrict__ ecn, T omega) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = t / (2 * ROW_LANES);
  const int h = (t / ROW_LANES) & 1, sub = t % ROW_LANES;
  const int nc = (n - 1) / 2 + 1;
  if (I >= nc) return;  // both groups of a pair leave together
  const int i = 2 * I + h;
  T rr = 0, en = 0;
  if (i < n) {  // uniform over the group
    rr = b[i] - level_row_group<T>(i, sub, rp, col, val, ein, (const T*)nullptr);
    en = ein[i] + omega * rr / diag[i];
    if (sub == 0) eout[i] = en;
  }
  // pair sum: lane l of group 0 with lane l of group 1
  const T rs = rr + __shfl_xor(rr, ROW_LANES, 64);
  const T es = en + __shfl_xor(en, ROW_LANES, 64);
  if (h == 0 && sub == 0) {
    rc[I] = rs;
    ecn[I] = es;
  }
}

// One level of one component for the single-work-group kernel.  Row i (0 <= i < n) has
// entries [rp[i], rp[i+1]) of col/val; col - cbase is the row index within r/A/B.
template <typename T> struct LevelDesc {
  int n;
  int cbase;
  const int32_t* rp;
  const int32_t* col;
  const T* val;
  const T* diag;
  T* r;  // right-hand side of the level
  T* A;  // ping
  T* B;  // pong; the level's final correction after the up-cycle
};

// The coarse end of the cycle (levels l0..L-1 of one component per work-group, where a
// grid launch per sweep would be pure launch latency), with the component's coarse
// hierarchy staged in LDS: CSR (int32 rowptr/col, T values), diagonal and the r/A/B
// vectors of every level.  The level table itself is built in LDS by thread 0.  Reads r
// and A of level l0 from global memory and writes B of level l0 back.  Same operations as
// the grid kernels above.
template <typename T> struct LdsLevel {
  int n;
  int* rp;
  int* col;
  T* val;
  T* diag;
  T* r;
  T* A;
  T* B;
};

constexpr int MAX_LEVELS = 40;

__host__ __device__ inline size_t lds_align(size_t x) { return (x + 15) & ~size_t(15); }

template <typename T>
__host__ __device__ inline size_t lds_level_bytes(int n, int nnz) {
  return lds_align(sizeof(int) * (n + 1)) + lds_align(sizeof(int) * nnz) +
         lds_align(sizeof(T) * nnz) + 4 * lds_align(sizeof(T) * n);
}

template <typename T>
__global__ __launch_bounds__(1024) void k_small_levels_lds(const LevelDesc<T>* __restrict__ descs,
                                                           const int32_t* __restrict__ rng,
                                                           T omega, int pre, int post) {
  extern __shared__ __align__(16) unsigned char smem[];
  LdsLevel<T>* tab = reinterpret_cast<LdsLevel<T>*>(smem);
  const int tid = threadIdx.x, nth = blockDim.x;
  const LevelDesc<T>* lvs = descs + (size_t)blockIdx.x * MAX_LEVELS;
  const int l0 = rng[2 * blockIdx.x], L = rng[2 * blockIdx.x + 1];
  if (tid == 0) {
    size_t off = lds_align(sizeof(LdsLevel<T>) * MAX_LEVELS);
    for (int l = l0; l < L; ++l) {
      const int n = lvs[l].n, nnz = lvs[l].rp[n] - lvs[l].rp[0];
      LdsLevel<T>& t = tab[l - l0];
      t.n = n;
      t.rp = reinterpret_cast<int*>(smem + off);
      off += lds_align(sizeof(int) * (n + 1));
      t.col = reinterpret_cast<int*>(smem + off);
      off += lds_align(sizeof(int) * nnz);
      t.val = reinterpret_cast<T*>(smem + off);
      off += lds_align(sizeof(T) * nnz);
      t.diag = reinterpret_cast<T*>(smem + off);
      off += lds_align(sizeof(T) * n);
      t.r = reinterpret_cast<T*>(smem + off);
      off += lds_align(sizeof(T) * n);
      t.A = reinterpret_cast<T*>(smem + off);
      off += lds_align(sizeof(T) * n);
      t.B = reinterpret_cast<T*>(smem + off);
      off += lds_align(sizeof(T) * n);
    }
  }
  __syncthreads();
  for (int l = l0; l < L; ++l) {
    const LevelDesc<T>& g = lvs[l];
    const LdsLevel<T>& t = tab[l - l0];
    const int base = g.rp[0];
    const int n = t.n, nnz = g.rp[n] - base;
    for (int i = tid; i <= n; i += nth) t.rp[i] = g.rp[i] - base;
    for (int q = tid; q < nnz; q += nth) {
      t.col[q] = g.col[base + q] - g.cbase;
      t.val[q] = g.val[base + q];
    }
    for (int i = tid; i < n; i += nth) t.diag[i] = g.diag[i];
    if (l == l0)
      for (int i = tid; i < n; i += nth) {
        t.r[i] = g.r[i];
        t.A[i] = g.A[i];
      }
  }
  __syncthreads();
  // a level of n rows gets G lanes per row (a power of two, at most 16, as many as the
  // work-group can give every row in one pass): the entries of a row are loaded side by
  // side and summed with shuffles, as in the grid kernels
  auto lanes_for = [&](int n) {
    int g = 1;
    while (g < 16 && 2 * g * n <= nth) g *= 2;
    return g;
  };
  auto row = [&](const LdsLevel<T>& t, int i, int sub, int G, const T* ein, const T* ec) {
    T s = 0;
    for (int q = t.rp[i] + sub; q < t.rp[i + 1]; q += G) {
      const int j = t.col[q];
      T ej = ein[j];
      if (ec) ej += ec[j >> 1];
      s += t.val[q] * ej;
    }
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
  };
  auto sweep = [&](const LdsLevel<T>& t, const T* ein, const T* ec, T* eout) {
    const int G = lanes_for(t.n), R = nth / G;
    for (int base = 0; base < t.n; base += R) {
      const int i = base + tid / G, sub = tid % G;
      if (i < t.n) {  // uniform over a lane group
        const T ax = row(t, i, sub, G, ein, ec);
        if (sub == 0) {
          T ei = ein[i];
          if (ec) ei += ec[i >> 1];
          eout[i] = ei + omega * (t.r[i] - ax) / t.diag[i];
        }
      }
    }
    __syncthreads();
  };
  for (int l = l0; l < L - 1; ++l) {
    const LdsLevel<T>& t = tab[l - l0];
    const LdsLevel<T>& tc = tab[l + 1 - l0];
    T* cur = t.A;
    T* nxt = t.B;
    for (int it = 0; it < pre; ++it) {
      sweep(t, cur, (const T*)nullptr, nxt);
      T* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
    if (cur == t.A) {
      for (int i = tid; i < t.n; i += nth) t.B[i] = t.A[i];
      __syncthreads();
      cur = t.B;
    }
    // fused residual + sweep + pairwise restriction: 2G lanes per coarse row, one group
    // of G per fine row of its pair
    {
      const int G = lanes_for(t.n), R = nth / (2 * G);
      for (int base = 0; base < tc.n; base += R) {
        const int I = base + tid / (2 * G);
        const int h = (tid / G) & 1, sub = tid % G;
        if (I < tc.n) {  // uniform over the 2G lanes of the pair
          const int i = 2 * I + h;
          T rr = 0, en = 0;
          if (i < t.n) {  // uniform over a lane group
            rr = t.r[i] - row(t, i, sub, G, cur, (const T*)nullptr);
            en = cur[i] + omega * rr / t.diag[i];
            if (sub == 0) t.A[i] = en;
          }
          const T rs = rr + __shfl_xor(rr, G, 64);
          const T es = en + __shfl_xor(en, G, 64);
          if (h == 0 && sub == 0) {
            tc.r[I] = rs;
            tc.A[I] = es;
          }
        }
      }
    }
    __syncthreads();
  }
  {
    const LdsLevel<T>& t = tab[L - 1 - l0];
    for (int i = tid; i < t.n; i += nth) t.B[i] = t.r[i] / t.diag[i];
    __syncthreads();
  }
  for (int l = L - 2; l >= l0; --l) {
    const LdsLevel<T>& t = tab[l - l0];
    const T* ec = tab[l + 1 - l0].B;
    sweep(t, t.A, ec, t.B);
    const T* cur = t.B;
    T* nxt = t.A;
    for (int it = 1; it < post; ++it) {
      sweep(t, cur, (const T*)nullptr, nxt);
      T* tmp = const_cast<T*>(cur);
      cur = nxt;
      nxt = tmp;
    }
    if (cur != t.B) {
      for (int i = tid; i < t.n; i += nth) t.B[i] = cur[i];
      __syncthreads();
    }
  }
  {
    const LdsLevel<T>& t = tab[0];
    T* out = lvs[l0].B;
    for (int i = tid; i < t.n; i += nth) out[i] = t.B[i];
  }
}

// LDS bytes needed to stage levels l0.. of one component (level table included)
template <typename T>
size_t lds_bytes_from(const std::vector<int64_t>& n, const std::vector<int64_t>& nnz, int l0) {
  size_t b = lds_align(sizeof(LdsLevel<T>) * MAX_LEVELS);
  for (size_t l = l0; l < n.size(); ++l) b += lds_level_bytes<T>((int)n[l], (int)nnz[l]);
  return b;
}

// first level from which the remaining hierarchy fits in `budget` bytes of LDS (FP64)
int lds_start_level(const std::vector<int64_t>& n, const std::vector<int64_t>& nnz, size_t budget) {
  const int L = (int)n.size();
  if (L > MAX_LEVELS) return -1;
  int l0 = L - 1;
  for (int l = L - 1; l >= 0; --l) {
    if (lds_bytes_from<double>(n, nnz, l) <= budget) l0 = l;
    else break;
  }
  return l0;
}

// The dynamic-LDS cap is a property of the kernel, shared by every context: it only
// ever grows, so a context set up later never lowers what another one launches with.
template <typename T>
int set_small_lds_attr(size_t bytes) {
  static std::mutex mu;
  static size_t cap = 0;
  std::lock_guard<std::mutex> lock(mu);
  if (bytes <= cap) return 0;
  DGMI_HIP(hipFuncSetAttribute((const void*)k_small_levels_lds<T>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  cap = bytes;
  return 0;
}
template int set_small_lds_attr<double>(size_t);
template int set_small_lds_attr<float>(size_t);
template size_t lds_bytes_from<double>(const std::vector<int64_t>&, const std::vector<int64_t>&, int);
template size_t lds_bytes_from<float>(const std::vector<int64_t>&, const std::vector<int64_t>&, int);

// ---------------------------------------------------------------------------------------
template <typename T> static T* vec(void* p) { return reinterpret_cast<T*>(p); }

// device copy of a host vector (converted to T); nullptr on failure
template <typename T, typename U>
static T* upload_vec(dgmi_ctx* c, const std::vector<U>& h) {
  std::vector<T> t(h.begin(), h.end());
  T* d = static_cast<T*>(c->alloc(sizeof(T) * std::max<size_t>(t.size(), 1)));
  if (!d) return nullptr;
  if (!t.empty()) {
    const hipError_t e = hipMemcpy(d, t.data(), sizeof(T) * t.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      note_alloc_error(std::string("hipMemcpy to device: ") + hipGetErrorString(e));
      return nullptr;
    }
  }
  return d;
}

template <typename T>
int launch_restrict(dgmi_ctx* c, const T* r_dg, T* r_sfc, hipStream_t s) {
  const int ncg = (int)c->mg.cg_nonods;
  const dim3 blk(256), grd((ncg + 255) / 256);
  const T* rw = std::is_same<T, double>::value ? (const T*)c->mg.rw64 : (const T*)c->mg.rw32;
  if (c->D == 3)
    hipLaunchKernelGGL((k_restrict<T, 3>), grd, blk, 0, s, ncg, c->mg.rptr, c->mg.rent, rw, r_dg, r_sfc);
  else
    hipLaunchKernelGGL((k_restrict<T, 2>), grd, blk, 0, s, ncg, c->mg.rptr, c->mg.rent, rw, r_dg, r_sfc);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_prolong_add(dgmi_ctx* c, const T* e_sfc, T* x_dg, hipStream_t s) {
  const int nele = (int)c->nele;
  const dim3 blk(256), grd((nele + 255) / 256);
  const T* rw = std::is_same<T, double>::value ? (const T*)c->mg.rw64 : (const T*)c->mg.rw32;
  if (c->D == 3)
    hipLaunchKernelGGL((k_prolong_add<T, 3>), grd, blk, 0, s, nele, c->mg.cellpos, rw, e_sfc, x_dg);
  else
    hipLaunchKernelGGL((k_prolong_add<T, 2>), grd, blk, 0, s, nele, c->mg.cellpos, rw, e_sfc, x_dg);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int launch_probe_mask(dgmi_ctx* c, const int32_t* whichc, int color, double* x, hipStream_t s) {
  const int nele = (int)c->nele;
  const dim3 blk(256), grd((nele + 255) / 256);
  if (c->D == 3)
    hipLaunchKernelGGL(k_probe_mask<3>, grd, blk, 0, s, nele, c->mg.cellpos, c->mg.rw64, whichc, color, x);
  else
    hipLaunchKernelGGL(k_probe_mask<2>, grd, blk, 0, s, nele, c->mg.cellpos, c->mg.rw64, whichc, color, x);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int launch_probe_collect(dgmi_ctx*, const double* r_sfc, const int32_t* spidx,
                         const int32_t* rowpos, int64_t nent, double* values, hipStream_t s) {
  if (nent == 0) return 0;
  const dim3 blk(256), grd((unsigned)((nent + 255) / 256));
  hipLaunchKernelGGL(k_probe_collect, grd, blk, 0, s, nent, spidx, rowpos, r_sfc, values);
  DGMI_HIP(hipGetLastError());
  return 0;
}

// One V-cycle (z02/multigrid_linearelastic.py:112-172) of every component.  Level buffers:
// A_l/B_l ping-pong, level 0's B is the caller's output e_sfc and level 0's rhs is the
// caller's r_sfc.  Grid levels 0..small_level-1 sweep the rows of the components that
// start deeper; one work-group per component then runs the rest.
// y[i] = x[i] (or 0 with x null): replaces memset/memcpy nodes so the cycle can be
// captured in a graph
template <typename T>
__global__ void k_copy(int64_t n, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x ? x[i] : T(0);
}

template <typename T>
static int vcycle_body(dgmi_ctx* c, const T* r_sfc, T* e_sfc, hipStream_t s) {
  auto& mg = c->mg;
  auto& lv = mg.lv;
  const bool f64 = std::is_same<T, double>::value;
  const T omega = (T)c->omega;
  const int l0 = mg.small_level;
  auto R = [&](int l) -> T* { return l == 0 ? const_cast<T*>(r_sfc) : vec<T>(lv.r[l]); };
  auto A = [&](int l) -> T* { return vec<T>(lv.e[l]); };
  auto B = [&](int l) -> T* { return l == 0 ? e_sfc : vec<T>(lv.e2[l]); };
  auto val = [&](int l) -> const T* { return f64 ? (const T*)lv.val64[l] : (const T*)lv.val32[l]; };
  auto dia = [&](int l) -> const T* { return f64 ? (const T*)lv.diag64[l] : (const T*)lv.diag32[l]; };
  const dim3 blk(256);
  const dim3 gcp((unsigned)((lv.n[0] + 255) / 256));
  hipLaunchKernelGGL((k_copy<T>), gcp, blk, 0, s, lv.n[0], (const T*)nullptr, A(0));
  // down-cycle on grid-sized levels
  for (int l = 0; l < l0; ++l) {
    const int n = (int)lv.grid_rows[l];
    const dim3 grd((unsigned)(((int64_t)n * ROW_LANES + 255) / 256));
    T* cur = A(l);
    T* nxt = B(l);
    for (int it = 0; it < 3; ++it) {
      hipLaunchKernelGGL((k_sweep<T>), grd, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l), dia(l),
                         R(l), cur, (const T*)nullptr, nxt, omega);
      std::swap(cur, nxt);
    }
    // after 3 sweeps cur == B(l); 4th sweep writes A(l) and the restricted r, e
    const int nc = (n - 1) / 2 + 1;
    const dim3 grc((unsigned)(((int64_t)nc * 2 * ROW_LANES + 255) / 256));
    hipLaunchKernelGGL((k_sweep_restrict<T>), grc, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l),
                       dia(l), R(l), cur, nxt, R(l + 1), A(l + 1), omega);
  }
  // the level tables point level 0 at internal buffers, so components that start at
  // level 0 see the caller's vectors through them
  if (lv.any_start0)
    hipLaunchKernelGGL((k_copy<T>), gcp, blk, 0, s, lv.n[0], r_sfc, vec<T>(lv.r[0]));
  hipLaunchKernelGGL((k_small_levels_lds<T>), dim3(c->ncomp), dim3(1024), f64 ? lv.lds64 : lv.lds32,
                     s, (const LevelDesc<T>*)(f64 ? lv.desc64 : lv.desc32), (const int32_t*)lv.rng,
                     omega, 3, 3);
  if (lv.any_start0)
    hipLaunchKernelGGL((k_copy<T>), gcp, blk, 0, s, lv.n[0], (const T*)vec<T>(lv.e2[0]), e_sfc);
  // up-cycle on grid-sized levels
  for (int l = l0 - 1; l >= 0; --l) {
    const int n = (int)lv.grid_rows[l];
    const dim3 grd((unsigned)(((int64_t)n * ROW_LANES + 255) / 256));
    // (A + P B_{l+1}) -> B, B -> A, A -> B
    hipLaunchKernelGGL((k_sweep<T>), grd, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l), dia(l),
                       R(l), A(l), (const T*)B(l + 1), B(l), omega);
    hipLaunchKernelGGL((k_sweep<T>), grd, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l), dia(l),
                       R(l), B(l), (const T*)nullptr, A(l), omega);
    hipLaunchKernelGGL((k_sweep<T>), grd, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l), dia(l),
                       R(l), A(l), (const T*)nullptr, B(l), omega);
  }
  DGMI_HIP(hipGetLastError());
  return 0;
}

// The cycle is a fixed sequence of ~7 launches per grid level plus one; it is captured
// once per (precision, r, e, omega) into a graph on a private stream and replayed on the
// caller's stream.
template <typename T>
int launch_sfc_vcycle(dgmi_ctx* c, const T* r_sfc, T* e_sfc, hipStream_t s) {
  auto& mg = c->mg;
  if (!mg.use_graph) return vcycle_body<T>(c, r_sfc, e_sfc, s);
  VGraph& vg = std::is_same<T, double>::value ? mg.graph64 : mg.graph32;
  if (!vg.exec || vg.r != (const void*)r_sfc || vg.e != (void*)e_sfc || vg.omega != c->omega) {
    if (vg.exec) {
      DGMI_HIP(hipGraphExecDestroy(vg.exec));
      vg.exec = nullptr;
    }
    if (!mg.capture_stream)
      DGMI_HIP(hipStreamCreateWithFlags(&mg.capture_stream, hipStreamNonBlocking));
    DGMI_HIP(hipStreamBeginCapture(mg.capture_stream, hipStreamCaptureModeRelaxed));
    const int rc = vcycle_body<T>(c, r_sfc, e_sfc, mg.capture_stream);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(mg.capture_stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    DGMI_HIP(ec);
    const hipError_t ei = hipGraphInstantiate(&vg.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DGMI_HIP(ei);
    vg.r = r_sfc;
    vg.e = e_sfc;
    vg.omega = c->omega;
  }
  DGMI_HIP(hipGraphLaunch(vg.exec, s));
  return 0;
}

// Device layout of every component's SFC hierarchy (see SfcLevelsDev) and the level
// tables of the single-work-group kernel.
int upload_sfc_levels(dgmi_ctx* c, const SfcLayout& lay) {
  auto& mg = c->mg;
  auto& lv = mg.lv;
  const int nc = (int)lay.lh.size();
  const int S = lay.smax;
  auto fail = [&]() {
    set_error(alloc_failure());
    return -3;
  };
  // ---- grid levels 0..S
  lv.nlevel = S + 1;
  lv.n.assign(S + 1, 0);
  lv.nnz.assign(S + 1, 0);
  lv.grid_rows.assign(S, 0);
  lv.rowptr.assign(S + 1, nullptr);
  lv.col.assign(S + 1, nullptr);
  lv.val64.assign(S + 1, nullptr);
  lv.diag64.assign(S + 1, nullptr);
  lv.val32.assign(S + 1, nullptr);
  lv.diag32.assign(S + 1, nullptr);
  lv.r.assign(S + 1, nullptr);
  lv.e.assign(S + 1, nullptr);
  lv.e2.assign(S + 1, nullptr);
  std::vector<std::vector<int64_t>> segoff(S + 1, std::vector<int64_t>(nc, -1));
  std::vector<std::vector<int64_t>> rpoff(S + 1, std::vector<int64_t>(nc, 0));
  for (int l = 0; l <= S; ++l) {
    std::vector<int32_t> rp(1, 0), col;
    std::vector<double> val, diag;
    int64_t rows = 0;
    for (int q : lay.order) {
      if (lay.start[q] < l) continue;
      const SfcLevelsHost& h = lay.lh[q];
      const int64_t seg = (q == lay.order.back()) ? h.n[l] : (lay.len0[q] >> l);
      segoff[l][q] = rows;
      rpoff[l][q] = (int64_t)rp.size() - 1;
      if (l < S && lay.start[q] > l) lv.grid_rows[l] = rows + seg;
      for (int64_t i = 0; i < seg; ++i) {
        if (i < h.n[l]) {
          for (int32_t k = h.rowptr[l][i]; k < h.rowptr[l][i + 1]; ++k) {
            col.push_back((int32_t)(h.col[l][k] + rows));
            val.push_back(h.val[l][k]);
          }
          diag.push_back(h.diag[l][i]);
        } else {
          diag.push_back(1.0);  // padding row: no entries, stays zero
        }
        rp.push_back((int32_t)col.size());
      }
      rows += seg;
    }
    lv.n[l] = rows;
    lv.nnz[l] = (int64_t)col.size();
    lv.rowptr[l] = upload_vec<int32_t>(c, rp);
    lv.col[l] = upload_vec<int32_t>(c, col);
    lv.val64[l] = upload_vec<double>(c, val);
    lv.diag64[l] = upload_vec<double>(c, diag);
    lv.val32[l] = upload_vec<float>(c, val);
    lv.diag32[l] = upload_vec<float>(c, diag);
    lv.r[l] = c->alloc(sizeof(double) * std::max<int64_t>(rows, 1));
    lv.e[l] = c->alloc(sizeof(double) * std::max<int64_t>(rows, 1));
    lv.e2[l] = c->alloc(sizeof(double) * std::max<int64_t>(rows, 1));
    if (!lv.rowptr[l] || !lv.col[l] || !lv.val64[l] || !lv.diag64[l] || !lv.val32[l] ||
        !lv.diag32[l] || !lv.r[l] || !lv.e[l] || !lv.e2[l])
      return fail();
  }
  // ---- pooled deeper levels
  std::vector<int32_t> prp, pcol;
  std::vector<double> pval, pdiag;
  struct PoolRef { int64_t rp, diag, vec; };
  std::vector<std::vector<PoolRef>> pref(nc);
  int64_t pvec = 0;
  for (int q = 0; q < nc; ++q) {
    const SfcLevelsHost& h = lay.lh[q];
    pref[q].resize(h.nlevel);
    for (int l = lay.start[q] + 1; l < h.nlevel; ++l) {
      pref[q][l] = {(int64_t)prp.size(), (int64_t)pdiag.size(), pvec};
      const int32_t base = (int32_t)pcol.size();
      for (int64_t i = 0; i <= h.n[l]; ++i) prp.push_back(base + h.rowptr[l][i]);
      pcol.insert(pcol.end(), h.col[l].begin(), h.col[l].end());
      pval.insert(pval.end(), h.val[l].begin(), h.val[l].end());
      pdiag.insert(pdiag.end(), h.diag[l].begin(), h.diag[l].end());
      pvec += 3 * h.n[l];
    }
  }
  lv.pool_rp = upload_vec<int32_t>(c, prp);
  lv.pool_col = upload_vec<int32_t>(c, pcol);
  lv.pool_val64 = upload_vec<double>(c, pval);
  lv.pool_diag64 = upload_vec<double>(c, pdiag);
  lv.pool_val32 = upload_vec<float>(c, pval);
  lv.pool_diag32 = upload_vec<float>(c, pdiag);
  lv.pool_vec = c->alloc(sizeof(double) * std::max<int64_t>(pvec, 1));
  if (!lv.pool_rp || !lv.pool_col || !lv.pool_val64 || !lv.pool_diag64 || !lv.pool_val32 ||
      !lv.pool_diag32 || !lv.pool_vec)
    return fail();
  // ---- level tables, components in layout order (one work-group each)
  std::vector<LevelDesc<double>> d64((size_t)nc * MAX_LEVELS);
  std::vector<LevelDesc<float>> d32((size_t)nc * MAX_LEVELS);
  std::vector<int32_t> rng(2 * nc);
  size_t lds64 = 0, lds32 = 0;
  lv.any_start0 = false;
  for (int b = 0; b < nc; ++b) {
    const int q = lay.order[b];
    const SfcLevelsHost& h = lay.lh[q];
    const int s0 = lay.start[q];
    if (h.nlevel > MAX_LEVELS) {
      set_error("too many SFC levels");
      return -2;
    }
    rng[2 * b] = s0;
    rng[2 * b + 1] = h.nlevel;
    lv.any_start0 |= s0 == 0;
    std::vector<int64_t> nnzq(h.nlevel);
    for (int l = 0; l < h.nlevel; ++l) nnzq[l] = (int64_t)h.col[l].size();
    lds64 = std::max(lds64, lds_bytes_from<double>(h.n, nnzq, s0));
    lds32 = std::max(lds32, lds_bytes_from<float>(h.n, nnzq, s0));
    for (int l = s0; l < h.nlevel; ++l) {
      LevelDesc<double>& a = d64[(size_t)b * MAX_LEVELS + l];
      LevelDesc<float>& f = d32[(size_t)b * MAX_LEVELS + l];
      a.n = f.n = (int)h.n[l];
      if (l == s0) {
        const int64_t so = segoff[l][q];
        a.cbase = f.cbase = (int)so;
        a.rp = f.rp = lv.rowptr[l] + rpoff[l][q];
        a.col = f.col = lv.col[l];
        a.val = lv.val64[l];
        f.val = lv.val32[l];
        a.diag = lv.diag64[l] + so;
        f.diag = lv.diag32[l] + so;
        a.r = (double*)lv.r[l] + so;
        a.A = (double*)lv.e[l] + so;
        a.B = (double*)lv.e2[l] + so;
        f.r = (float*)lv.r[l] + so;
        f.A = (float*)lv.e[l] + so;
        f.B = (float*)lv.e2[l] + so;
      } else {
        const PoolRef& p = pref[q][l];
        a.cbase = f.cbase = 0;
        a.rp = f.rp = lv.pool_rp + p.rp;
        a.col = f.col = lv.pool_col;
        a.val = lv.pool_val64;
        f.val = lv.pool_val32;
        a.diag = lv.pool_diag64 + p.diag;
        f.diag = lv.pool_diag32 + p.diag;
        double* v64 = (double*)lv.pool_vec + p.vec;
        float* v32 = (float*)lv.pool_vec + p.vec;  // FP32 views use the first half
        a.r = v64;
        a.A = v64 + h.n[l];
        a.B = v64 + 2 * h.n[l];
        f.r = v32;
        f.A = v32 + h.n[l];
        f.B = v32 + 2 * h.n[l];
      }
    }
  }
  lv.desc64 = upload_vec<LevelDesc<double>>(c, d64);
  lv.desc32 = upload_vec<LevelDesc<float>>(c, d32);
  lv.rng = upload_vec<int32_t>(c, rng);
  if (!lv.desc64 || !lv.desc32 || !lv.rng) return fail();
  lv.lds64 = lds64;
  lv.lds32 = lds32;
  if (int rc = set_small_lds_attr<double>(lds64)) return rc;
  return set_small_lds_attr<float>(lds32);
}

template int launch_restrict<double>(dgmi_ctx*, const double*, double*, hipStream_t);
template int launch_restrict<float>(dgmi_ctx*, const float*, float*, hipStream_t);
template int launch_prolong_add<double>(dgmi_ctx*, const double*, double*, hipStream_t);
template int launch_prolong_add<float>(dgmi_ctx*, const float*, float*, hipStream_t);
template int launch_sfc_vcycle<double>(dgmi_ctx*, const double*, double*, hipStream_t);
template int launch_sfc_vcycle<float>(dgmi_ctx*, const float*, float*, hipStream_t);

}  // namespace dgmi
