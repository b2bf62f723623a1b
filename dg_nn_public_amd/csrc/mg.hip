// DG <-> P1CG transfers, coloured probing of RAR, and the SFC V-cycle (gfx950).
//
// P1CG vectors live in SFC order (position p = sfc[node] - 1), so the reference's
// per-V-cycle argsort/permutation (z02/volume_mf_diff.py:1992-2007) is folded into the
// transfer kernels' index lists.  The 1-D SFC levels are CSR matrices; level l+1
// aggregates positions (2I, 2I+1) of level l.
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
template <typename T>
__device__ __forceinline__ T level_row(int i, const int32_t* __restrict__ rp,
                                       const int32_t* __restrict__ col,
                                       const T* __restrict__ val, const T* __restrict__ ein,
                                       const T* __restrict__ ec) {
  T s = 0;
  for (int q = rp[i]; q < rp[i + 1]; ++q) {
    const int j = col[q];
    T ej = ein[j];
    if (ec) ej += ec[j >> 1];
    s += val[q] * ej;
  }
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
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T ax = level_row<T>(i, rp, col, val, ein, ec);
  T ei = ein[i];
  if (ec) ei += ec[i >> 1];
  eout[i] = ei + omega * (b[i] - ax) / diag[i];
}

// Last pre-sweep of level l, fused with the restriction to level l+1: the reference
// evaluates rr = b - A e and applies e += omega rr/diag in the same call (:144-146), then
// restricts rr and e by pairwise sums (:147-153).  One thread per coarse node.
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_restrict(int n, const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ col,
                                                        const T* __restrict__ val,
                                                        const T* __restrict__ diag,
                                                        const T* __restrict__ b,
                                                        const T* __restrict__ ein,
                                                        T* __restrict__ eout, T* __restrict__ rc,
                                                        T* __restrict__ ecn, T omega) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = (n - 1) / 2 + 1;
  if (I >= nc) return;
  T rs = 0, es = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = 2 * I + h;
    if (i < n) {
      const T rr = b[i] - level_row<T>(i, rp, col, val, ein, (const T*)nullptr);
      const T en = ein[i] + omega * rr / diag[i];
      eout[i] = en;
      rs += rr;
      es += en;
    }
  }
  rc[I] = rs;
  ecn[I] = es;
}

template <typename T> struct LevelDesc {
  int n;
  const int32_t* rp;
  const int32_t* col;
  const T* val;
  const T* diag;
  T* r;  // right-hand side of the level
  T* A;  // ping
  T* B;  // pong; the level's final correction after the up-cycle
};

// All levels l0..L-1 in one work-group (levels small enough that a grid launch per sweep
// would be pure launch latency).  Same operations as the grid kernels above.
template <typename T>
__global__ __launch_bounds__(1024) void k_small_levels(const LevelDesc<T>* __restrict__ lvs, int l0,
                                                       int L, T omega, int pre, int post) {
  const int tid = threadIdx.x, nth = blockDim.x;
  auto sweep = [&](const LevelDesc<T>& lv, const T* ein, const T* ec, T* eout) {
    for (int i = tid; i < lv.n; i += nth) {
      const T ax = level_row<T>(i, lv.rp, lv.col, lv.val, ein, ec);
      T ei = ein[i];
      if (ec) ei += ec[i >> 1];
      eout[i] = ei + omega * (lv.r[i] - ax) / lv.diag[i];
    }
    __syncthreads();
  };
  for (int l = l0; l < L - 1; ++l) {
    const LevelDesc<T>& lv = lvs[l];
    const T* cur = lv.A;
    T* nxt = lv.B;
    for (int it = 0; it < pre; ++it) {
      sweep(lv, cur, (const T*)nullptr, nxt);
      T* t = const_cast<T*>(cur);
      cur = nxt;
      nxt = t;
    }
    // fused residual+sweep+restrict; result of the 4th sweep lands in lv.A
    const LevelDesc<T>& lc = lvs[l + 1];
    const int nc = lc.n;
    T* fin = lv.A;
    if (cur == lv.A) {  // need a distinct output buffer: copy cur -> B first
      for (int i = tid; i < lv.n; i += nth) lv.B[i] = lv.A[i];
      __syncthreads();
      cur = lv.B;
    }
    for (int I = tid; I < nc; I += nth) {
      T rs = 0, es = 0;
      for (int h = 0; h < 2; ++h) {
        const int i = 2 * I + h;
        if (i < lv.n) {
          const T rr = lv.r[i] - level_row<T>(i, lv.rp, lv.col, lv.val, cur, (const T*)nullptr);
          const T en = cur[i] + omega * rr / lv.diag[i];
          fin[i] = en;
          rs += rr;
          es += en;
        }
      }
      lc.r[I] = rs;
      lc.A[I] = es;
    }
    __syncthreads();
  }
  // coarsest level: one node, direct solve (z02/multigrid_linearelastic.py:156-160)
  {
    const LevelDesc<T>& lv = lvs[L - 1];
    for (int i = tid; i < lv.n; i += nth) lv.B[i] = lv.r[i] / lv.diag[i];
    __syncthreads();
  }
  for (int l = L - 2; l >= l0; --l) {
    const LevelDesc<T>& lv = lvs[l];
    const T* ec = lvs[l + 1].B;
    // post sweeps: (A + P ec) -> B, B -> A, A -> B, ...
    sweep(lv, lv.A, ec, lv.B);
    const T* cur = lv.B;
    T* nxt = lv.A;
    for (int it = 1; it < post; ++it) {
      sweep(lv, cur, (const T*)nullptr, nxt);
      T* t = const_cast<T*>(cur);
      cur = nxt;
      nxt = t;
    }
    if (cur != lv.B) {
      for (int i = tid; i < lv.n; i += nth) lv.B[i] = cur[i];
      __syncthreads();
    }
  }
}

// Same cycle over levels l0..L-1 with the whole coarse hierarchy staged in LDS: CSR
// (int32 rowptr/col, T values), diagonal and the r/A/B vectors of every level.  The
// level table itself is built in LDS by thread 0.  Reads r and A of level l0 from global
// memory and writes B of level l0 back.
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
__global__ __launch_bounds__(1024) void k_small_levels_lds(const LevelDesc<T>* __restrict__ lvs,
                                                           int l0, int L, T omega, int pre,
                                                           int post) {
  extern __shared__ __align__(16) unsigned char smem[];
  LdsLevel<T>* tab = reinterpret_cast<LdsLevel<T>*>(smem);
  const int tid = threadIdx.x, nth = blockDim.x;
  if (tid == 0) {
    size_t off = lds_align(sizeof(LdsLevel<T>) * MAX_LEVELS);
    for (int l = l0; l < L; ++l) {
      const int n = lvs[l].n, nnz = lvs[l].rp[n];
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
    const int n = t.n, nnz = g.rp[n];
    for (int i = tid; i <= n; i += nth) t.rp[i] = g.rp[i];
    for (int q = tid; q < nnz; q += nth) {
      t.col[q] = g.col[q];
      t.val[q] = g.val[q];
    }
    for (int i = tid; i < n; i += nth) t.diag[i] = g.diag[i];
    if (l == l0)
      for (int i = tid; i < n; i += nth) {
        t.r[i] = g.r[i];
        t.A[i] = g.A[i];
      }
  }
  __syncthreads();
  auto row = [&](const LdsLevel<T>& t, int i, const T* ein, const T* ec) {
    T s = 0;
    for (int q = t.rp[i]; q < t.rp[i + 1]; ++q) {
      const int j = t.col[q];
      T ej = ein[j];
      if (ec) ej += ec[j >> 1];
      s += t.val[q] * ej;
    }
    return s;
  };
  auto sweep = [&](const LdsLevel<T>& t, const T* ein, const T* ec, T* eout) {
    for (int i = tid; i < t.n; i += nth) {
      const T ax = row(t, i, ein, ec);
      T ei = ein[i];
      if (ec) ei += ec[i >> 1];
      eout[i] = ei + omega * (t.r[i] - ax) / t.diag[i];
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
    for (int I = tid; I < tc.n; I += nth) {
      T rs = 0, es = 0;
      for (int h = 0; h < 2; ++h) {
        const int i = 2 * I + h;
        if (i < t.n) {
          const T rr = t.r[i] - row(t, i, cur, (const T*)nullptr);
          const T en = cur[i] + omega * rr / t.diag[i];
          t.A[i] = en;
          rs += rr;
          es += en;
        }
      }
      tc.r[I] = rs;
      tc.A[I] = es;
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

// LDS bytes needed to stage levels l0..L-1 (level table included)
template <typename T>
size_t lds_bytes_from(const SfcLevelsDev& lv, const std::vector<int64_t>& nnz, int l0) {
  size_t b = lds_align(sizeof(LdsLevel<T>) * MAX_LEVELS);
  for (int l = l0; l < lv.nlevel; ++l) b += lds_level_bytes<T>((int)lv.n[l], (int)nnz[l]);
  return b;
}

// first level index from which the remaining hierarchy fits in `budget` bytes of LDS
int choose_small_level(const SfcLevelsDev& lv, const std::vector<int64_t>& nnz, size_t budget) {
  const int L = lv.nlevel;
  if (L > MAX_LEVELS) return -1;
  int l0 = L - 1;
  for (int l = L - 1; l >= 0; --l) {
    if (lds_bytes_from<double>(lv, nnz, l) <= budget) l0 = l;
    else break;
  }
  return l0;
}

template <typename T>
int set_small_lds_attr(size_t bytes) {
  DGMI_HIP(hipFuncSetAttribute((const void*)k_small_levels_lds<T>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return 0;
}
template int set_small_lds_attr<double>(size_t);
template int set_small_lds_attr<float>(size_t);
template size_t lds_bytes_from<double>(const SfcLevelsDev&, const std::vector<int64_t>&, int);
template size_t lds_bytes_from<float>(const SfcLevelsDev&, const std::vector<int64_t>&, int);

// ---------------------------------------------------------------------------------------
template <typename T> static T* vec(void* p) { return reinterpret_cast<T*>(p); }

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

// One V-cycle (z02/multigrid_linearelastic.py:112-172).  Level buffers: A_l/B_l ping-pong,
// level 0's B is the caller's output e_sfc and level 0's rhs is the caller's r_sfc.
template <typename T>
int launch_sfc_vcycle(dgmi_ctx* c, const T* r_sfc, T* e_sfc, hipStream_t s) {
  auto& mg = c->mg;
  auto& lv = mg.lv;
  const int L = lv.nlevel;
  const bool f64 = std::is_same<T, double>::value;
  const T omega = (T)c->omega;
  const int l0 = mg.small_level;
  auto R = [&](int l) -> T* { return l == 0 ? const_cast<T*>(r_sfc) : vec<T>(lv.r[l]); };
  auto A = [&](int l) -> T* { return vec<T>(lv.e[l]); };
  auto B = [&](int l) -> T* { return l == 0 ? e_sfc : vec<T>(lv.e2[l]); };
  auto val = [&](int l) -> const T* { return f64 ? (const T*)lv.val64[l] : (const T*)lv.val32[l]; };
  auto dia = [&](int l) -> const T* { return f64 ? (const T*)lv.diag64[l] : (const T*)lv.diag32[l]; };
  const dim3 blk(256);
  if (L == 1) {
    hipLaunchKernelGGL((k_sweep<T>), dim3(1), blk, 0, s, 1, lv.rowptr[0], lv.col[0], val(0),
                       dia(0), R(0), A(0), (const T*)nullptr, B(0), (T)1);  // e = r/a from e=0
    DGMI_HIP(hipGetLastError());
    return 0;
  }
  DGMI_HIP(hipMemsetAsync(A(0), 0, sizeof(T) * lv.n[0], s));
  // down-cycle on grid-sized levels
  for (int l = 0; l < l0 && l < L - 1; ++l) {
    const int n = (int)lv.n[l];
    const dim3 grd((n + 255) / 256);
    T* cur = A(l);
    T* nxt = B(l);
    for (int it = 0; it < 3; ++it) {
      hipLaunchKernelGGL((k_sweep<T>), grd, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l), dia(l),
                         R(l), cur, (const T*)nullptr, nxt, omega);
      std::swap(cur, nxt);
    }
    // after 3 sweeps cur == B(l); 4th sweep writes A(l) and the restricted r, e
    const int nc = (n - 1) / 2 + 1;
    const dim3 grc((nc + 255) / 256);
    hipLaunchKernelGGL((k_sweep_restrict<T>), grc, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l),
                       dia(l), R(l), cur, nxt, R(l + 1), A(l + 1), omega);
  }
  // small levels in one work-group (writes B(l0)).  The device level table points level
  // 0 at internal buffers, so when level 0 itself is small the caller's vectors are
  // staged through them.
  if (l0 == 0) {
    DGMI_HIP(hipMemcpyAsync(lv.r[0], r_sfc, sizeof(T) * lv.n[0], hipMemcpyDeviceToDevice, s));
  }
  if (mg.small_lds) {
    const size_t bytes = lds_bytes_from<T>(lv, lv.nnz, l0);
    hipLaunchKernelGGL((k_small_levels_lds<T>), dim3(1), dim3(1024), bytes, s,
                       (const LevelDesc<T>*)(f64 ? lv.desc64 : lv.desc32), l0, L, omega, 3, 3);
  } else {
    hipLaunchKernelGGL((k_small_levels<T>), dim3(1), dim3(1024), 0, s,
                       (const LevelDesc<T>*)(f64 ? lv.desc64 : lv.desc32), l0, L, omega, 3, 3);
  }
  if (l0 == 0) {
    DGMI_HIP(hipMemcpyAsync(e_sfc, lv.e2[0], sizeof(T) * lv.n[0], hipMemcpyDeviceToDevice, s));
  }
  // up-cycle on grid-sized levels
  for (int l = l0 - 1; l >= 0; --l) {
    const int n = (int)lv.n[l];
    const dim3 grd((n + 255) / 256);
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

// device tables of LevelDesc for the single-work-group kernel; level 0 points at internal
// buffers (the caller's vectors are staged through them when level 0 is small)
int upload_level_descs(dgmi_ctx* c) {
  auto& lv = c->mg.lv;
  const int L = lv.nlevel;
  std::vector<LevelDesc<double>> d64(L);
  std::vector<LevelDesc<float>> d32(L);
  for (int l = 0; l < L; ++l) {
    d64[l] = {(int)lv.n[l], lv.rowptr[l], lv.col[l], lv.val64[l], lv.diag64[l],
              (double*)lv.r[l], (double*)lv.e[l], (double*)lv.e2[l]};
    d32[l] = {(int)lv.n[l], lv.rowptr[l], lv.col[l], lv.val32[l], lv.diag32[l],
              (float*)lv.r[l], (float*)lv.e[l], (float*)lv.e2[l]};
  }
  lv.desc64 = c->alloc(sizeof(LevelDesc<double>) * L);
  lv.desc32 = c->alloc(sizeof(LevelDesc<float>) * L);
  if (!lv.desc64 || !lv.desc32) {
    set_error("device allocation failed");
    return -3;
  }
  DGMI_HIP(hipMemcpy(lv.desc64, d64.data(), sizeof(LevelDesc<double>) * L, hipMemcpyHostToDevice));
  DGMI_HIP(hipMemcpy(lv.desc32, d32.data(), sizeof(LevelDesc<float>) * L, hipMemcpyHostToDevice));
  return 0;
}

template int launch_restrict<double>(dgmi_ctx*, const double*, double*, hipStream_t);
template int launch_restrict<float>(dgmi_ctx*, const float*, float*, hipStream_t);
template int launch_prolong_add<double>(dgmi_ctx*, const double*, double*, hipStream_t);
template int launch_prolong_add<float>(dgmi_ctx*, const float*, float*, hipStream_t);
template int launch_sfc_vcycle<double>(dgmi_ctx*, const double*, double*, hipStream_t);
template int launch_sfc_vcycle<float>(dgmi_ctx*, const float*, float*, hipStream_t);

}  // namespace dgmi
