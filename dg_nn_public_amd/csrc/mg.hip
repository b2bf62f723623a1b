// DG <-> P1CG transfers, coloured probing of RAR, and the SFC V-cycle (gfx950).
//
// P1CG vectors live in SFC order (position p = sfc[node] - 1), so the reference's
// per-V-cycle argsort/permutation (z02/volume_mf_diff.py:1992-2007) is folded into the
// transfer kernels' index lists.  The 1-D SFC levels are CSR matrices; level l+1
// aggregates positions (2I, 2I+1) of level l.
#include <cmath>
#include <cstring>
#include <mutex>

#include "dgmi_internal.h"
#include "faces.h"

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

// r_sfc[p] = (1/mult) sum over the DG copies (e,k) of vertex p of sum_n P[n][k] r[e][n].
// A group of RESTRICT_LANES lanes per vertex takes its copies side by side (about 24 per
// vertex of a tetrahedral mesh, 20 loads each) and sums them with shuffles.  With `e0`
// set it also takes the V-cycle's first sweep from zero on level 0, e0 = omega r / diag.
constexpr int RESTRICT_LANES = 8;
template <typename T, int D>
__global__ __launch_bounds__(256) void k_restrict(int ncg, const int32_t* __restrict__ rptr,
                                                  const int32_t* __restrict__ rent,
                                                  const T* __restrict__ rw,
                                                  const T* __restrict__ r, T* __restrict__ out,
                                                  const T* __restrict__ dg0,
                                                  T* __restrict__ e0, T omega) {
  constexpr int NL = D == 3 ? 20 : 10, NB = D + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = t / RESTRICT_LANES, sub = t % RESTRICT_LANES;
  if (p >= ncg) return;  // whole groups leave together
  T s = 0;
  for (int q = rptr[p] + sub; q < rptr[p + 1]; q += RESTRICT_LANES) {
    const int idx = rent[q];
    const int e = idx / NB, k = idx - (idx / NB) * NB;
    const T* re = r + (size_t)e * NL;
    T u = 0;
#pragma unroll
    for (int n = 0; n < NL; ++n) u += (T)Ptab<D>(n, k) * re[n];
    s += u;
  }
#pragma unroll
  for (int off = RESTRICT_LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (sub == 0) {
    const T v = s * rw[p];
    out[p] = v;
    if (e0) e0[p] = omega * v / dg0[p];
  }
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

// k_prolong_add fused with the face-trace pass of the vector kernels: a work-group of
// nface waves owns 64 elements; every wave forms the element's whole new row
// x + P e (cheap: nloc x (D+1) products), writes its share of the rows (to x and to LDS,
// where the other waves read their face's nodes) and the traces of its face w into
// tout[(w*sngi + g)*nele + e] (same operations as the trace pass).
// fc / fnodes: the face-interpolation tables (faces.h); cf: (nface*D, nele).
template <typename T, int D>
__global__ __launch_bounds__(64 * (D + 1)) void k_prolong_trace(
    int nele, const int32_t* __restrict__ cellpos, const T* __restrict__ rw,
    const T* __restrict__ ec, T* __restrict__ x, const T* __restrict__ fc,
    const int32_t* __restrict__ fnodes, const T* __restrict__ cf,
    typename Vec2<T>::type* __restrict__ tout) {
  using F = FaceDims<D>;
  constexpr int NL = F::NL, NB = D + 1, NF = D + 1, SG = F::SG;
  constexpr int NPW = (NL + NF - 1) / NF;
  __shared__ T xs[NL][64];
  const int lane = threadIdx.x & 63;
  const int f = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int e = xcd_group(blockIdx.x, gridDim.x) * 64 + lane;
  const bool valid = e < nele;
  T xv[NL];
  T* xe = x + (size_t)(valid ? e : 0) * NL;
  if (valid) {
    T ev[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int p = cellpos[e * NB + k];
      ev[k] = ec[p] * rw[p];
    }
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      T s = xe[n];
#pragma unroll
      for (int k = 0; k < NB; ++k) s += (T)Ptab<D>(n, k) * ev[k];
      xv[n] = s;
    }
  }
  // every wave has read the old row before any wave writes its share of the new one
  __syncthreads();
  if (valid) {
#pragma unroll
    for (int k = 0; k < NPW; ++k) {
      const int n = f * NPW + k;
      if (n < NL) {
        xe[n] = xv[n];
        xs[n][lane] = xv[n];
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  T c[D], xf[F::NFN];
#pragma unroll
  for (int d = 0; d < D; ++d) c[d] = cf[(f * D + d) * nele + e];
#pragma unroll
  for (int i = 0; i < F::NFN; ++i) xf[i] = xs[fnodes[f * F::NFN + i]][lane];
  face_traces<T, D>(f, fc, xv, xf, c, [&](int g, T u, T dn) {
    tout[(f * SG + g) * nele + e] = make_trace<typename Vec2<T>::type>(u, dn);
  });
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

// RAR assembled from the element matrices of the continuous P1 hat functions.  For a P1CG
// vector v, P I_fc v is the continuous P1 function of v/mult, and I_cf sums the test
// functions of every vertex over its elements, so RAR[i][j] = a(phi_i, phi_j) / (mult_i
// mult_j) with a the SIPG form: its interior-face terms all carry a jump of u or v and
// vanish, leaving the volume term and the Dirichlet faces' terms.  Per element (one
// thread per entry (a, b)), with the reference tables reduced to the vertex functions
// (rarm: SP[j][k] = P^T Sref_jk P | MP[f] = P^T M_f P | CP[f][d] = P^T C_fd P):
//   ke[a][b] = sum_q G_q (SP_jk + SP_kj)[a][b]
//            + sum_{Dirichlet f} wf_f (pen_f MP_f[a][b] - sum_d c_fd (CP_fd[a][b] + CP_fd[b][a]))
// the terms of k_bdiag's element block with the interior faces left out.
template <int D>
__global__ __launch_bounds__(256) void k_rar_elem(int nele, const double* __restrict__ rarm,
                                                  const double* __restrict__ G,
                                                  const double* __restrict__ cf,
                                                  const double* __restrict__ wf,
                                                  const double* __restrict__ pen,
                                                  const int32_t* __restrict__ fcode,
                                                  double* __restrict__ ke,
                                                  double* __restrict__ miss) {
  constexpr int NB = D + 1, NB2 = NB * NB, NF = D + 1;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) *miss = 0.0;
  if (tid >= (int64_t)nele * NB2) return;
  const int e = (int)(tid / NB2);
  const int ab = (int)(tid - (int64_t)e * NB2);
  const int a = ab / NB, b = ab - (ab / NB) * NB;
  const double* SP = rarm;
  const double* MP = SP + D * D * NB2;
  const double* CP = MP + NF * NB2;
  constexpr int NQ = D == 3 ? 6 : 3;
  constexpr int jj[6] = {0, 1, 2, 0, 0, 1}, kk[6] = {0, 1, 2, 1, 2, 2};
  constexpr int jj2[3] = {0, 1, 0}, kk2[3] = {0, 1, 1};
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int j = D == 3 ? jj[q] : jj2[q], k = D == 3 ? kk[q] : kk2[q];
    double v = SP[(j * D + k) * NB2 + ab];
    if (j != k) v += SP[(k * D + j) * NB2 + ab];
    s += G[q * nele + e] * v;
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    if ((fcode[e * NF + f] & 15) != 1) continue;
    double cs = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d)
      cs += cf[(f * D + d) * nele + e] *
            (CP[(f * D + d) * NB2 + a * NB + b] + CP[(f * D + d) * NB2 + b * NB + a]);
    s += wf[f * nele + e] * (pen[f * nele + e] * MP[f * NB2 + ab] - cs);
  }
  ke[(size_t)e * NB2 + ab] = s;
}

// One thread per P1CG row (by position p): its slots [rlo[p], rhi[p]) with column positions
// colpos are summed over the row vertex's element entries (rptr/rent, in a fixed order, so
// the result does not depend on scheduling) and scaled by rw_p rw_j.
template <int D>
__global__ __launch_bounds__(256) void k_rar_gather(int ncg, const int32_t* __restrict__ rptr,
                                                    const int32_t* __restrict__ rent,
                                                    const int32_t* __restrict__ cellpos,
                                                    const double* __restrict__ rw,
                                                    const double* __restrict__ ke,
                                                    const int32_t* __restrict__ rlo,
                                                    const int32_t* __restrict__ rhi,
                                                    const int32_t* __restrict__ colpos,
                                                    double* __restrict__ vals,
                                                    double* __restrict__ miss) {
  constexpr int NB = D + 1, NB2 = NB * NB;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncg) return;
  const int k0 = rlo[p], k1 = rhi[p];
  for (int k = k0; k < k1; ++k) vals[k] = 0.0;
  for (int t = rptr[p]; t < rptr[p + 1]; ++t) {
    const int q = rent[t];
    const int e = q / NB, a = q - e * NB;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int jp = cellpos[e * NB + b];
      int k = k0;
      while (k < k1 && colpos[k] != jp) ++k;
      if (k < k1)
        vals[k] += ke[(size_t)e * NB2 + a * NB + b];
      else
        *miss = 1.0;
    }
  }
  const double wp = rw[p];
  for (int k = k0; k < k1; ++k) vals[k] *= wp * rw[colpos[k]];
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
  // the row's own operands are requested before the row product (one fewer dependent
  // round trip to memory)
  T ei = ein[i];
  if (ec) ei += ec[i >> 1];
  const T bi = b[i], di = diag[i];
  const T ax = level_row_group<T>(i, sub, rp, col, val, ein, ec);
  if (sub == 0) eout[i] = ei + omega * (bi - ax) / di;
}

// First sweep of a cycle, whose input is zero: eout = omega b / diag (no row product).
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_zero(int n, const T* __restrict__ diag,
                                                    const T* __restrict__ b,
                                                    T* __restrict__ eout, T omega) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) eout[i] = omega * b[i] / diag[i];
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
    const T bi = b[i], ei = ein[i], di = diag[i];  // requested before the row product
    rr = bi - level_row_group<T>(i, sub, rp, col, val, ein, (const T*)nullptr);
    en = ei + omega * rr / di;
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
// hierarchy in LDS: CSR (int32 rowptr/col, T values), diagonal and the r/A/B vectors of
// every level.  The matrices come as a prebuilt byte image of the LDS layout (a table of
// LdsImgLevel offsets, then the arrays), copied with 16-byte loads that are all issued
// before the first LDS store; r and A of level l0 are read from global memory and B of
// level l0 is written back.  Same operations as the grid kernels above.
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
constexpr int SMALL_THREADS = 1024;

// byte offsets (from the start of the LDS image) of one level's arrays
struct LdsImgLevel {
  int n, rp, col, val, diag, r, A, B;
};

__host__ __device__ inline size_t lds_align(size_t x) { return (x + 15) & ~size_t(15); }

// Optional exact coarse solve (cinv != nullptr): cinfo[2b] = offset of the component's
// inverse in cinv (row-major, n x n, n <= EXACT_MAX_N), cinfo[2b+1] = its level c or
// -1.  With c >= 0 the cycle stops at level c and sets B_c = A_c^-1 r_c there; the levels
// below c are not visited.
constexpr int EXACT_MAX_N = 512;    // single-work-group path (packed solves)
// grid path (a context of one solve): the limit is the context's dense_max_n

template <typename T>
__global__ __launch_bounds__(SMALL_THREADS) void k_small_levels_lds(
    const LevelDesc<T>* __restrict__ descs, const int32_t* __restrict__ rng,
    const uint4* __restrict__ img, const int64_t* __restrict__ imginfo, T omega, int pre,
    int post, const T* __restrict__ cinv, const int64_t* __restrict__ cinfo) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ LdsLevel<T> tab[MAX_LEVELS];
  const int tid = threadIdx.x;
  constexpr int nth = SMALL_THREADS;
  const int l0 = rng[2 * blockIdx.x];
  int L = rng[2 * blockIdx.x + 1];
  const T* ainv = nullptr;
  if (cinv) {
    const int c = (int)cinfo[2 * blockIdx.x + 1];
    if (c >= l0 && c < L) {
      L = c + 1;
      ainv = cinv + cinfo[2 * blockIdx.x];
    }
  }
  const LevelDesc<T>& g0 = descs[(size_t)blockIdx.x * MAX_LEVELS + l0];
  {
    const uint4* src = img + imginfo[2 * blockIdx.x];
    const int nw = (int)imginfo[2 * blockIdx.x + 1];
    uint4* dst = reinterpret_cast<uint4*>(smem);
    constexpr int U = 8;
    for (int base = 0; base < nw; base += U * nth) {
      // plain words (a conditionally filled array of vector structs stays in scratch)
      unsigned int t[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * nth + tid;
        const uint4 v = i < nw ? src[i] : make_uint4(0u, 0u, 0u, 0u);
        t[u][0] = v.x;
        t[u][1] = v.y;
        t[u][2] = v.z;
        t[u][3] = v.w;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = base + u * nth + tid;
        if (i < nw) dst[i] = make_uint4(t[u][0], t[u][1], t[u][2], t[u][3]);
      }
    }
  }
  __syncthreads();
  if (tid < L - l0) {
    const LdsImgLevel o = reinterpret_cast<const LdsImgLevel*>(smem)[tid];
    LdsLevel<T>& t = tab[tid];
    t.n = o.n;
    t.rp = reinterpret_cast<int*>(smem + o.rp);
    t.col = reinterpret_cast<int*>(smem + o.col);
    t.val = reinterpret_cast<T*>(smem + o.val);
    t.diag = reinterpret_cast<T*>(smem + o.diag);
    t.r = reinterpret_cast<T*>(smem + o.r);
    t.A = reinterpret_cast<T*>(smem + o.A);
    t.B = reinterpret_cast<T*>(smem + o.B);
  }
  __syncthreads();
  {
    const LdsLevel<T> t = tab[0];
    for (int i = tid; i < t.n; i += nth) {
      t.r[i] = g0.r[i];
      t.A[i] = g0.A[i];
    }
  }
  __syncthreads();
  // a level of n rows gets G lanes per row (a power of two, at most 16, as many as the
  // work-group can give every row in one pass): the entries of a row are loaded side by
  // side and summed with shuffles, as in the grid kernels
  auto lanes_for = [=](int n) {
    int g = 1;
    while (g < 16 && 2 * g * n <= nth) g *= 2;
    return g;
  };
  // a lane's entries of a row are taken four at a time, so their LDS loads are in flight
  // together (the single work-group is latency-bound)
  // level descriptors are copied out of LDS into registers (by value), so the pointer
  // loads are not repeated after every LDS store
  auto row = [=](const LdsLevel<T> t, int i, int sub, int G, const T* ein, const T* ec) {
    T s = 0;
    int q = t.rp[i] + sub;
    const int q1 = t.rp[i + 1];
    for (; q + 3 * G < q1; q += 4 * G) {
      const int j0 = t.col[q], j1 = t.col[q + G], j2 = t.col[q + 2 * G], j3 = t.col[q + 3 * G];
      const T v0 = t.val[q], v1 = t.val[q + G], v2 = t.val[q + 2 * G], v3 = t.val[q + 3 * G];
      T e0 = ein[j0], e1 = ein[j1], e2 = ein[j2], e3 = ein[j3];
      if (ec) {
        e0 += ec[j0 >> 1];
        e1 += ec[j1 >> 1];
        e2 += ec[j2 >> 1];
        e3 += ec[j3 >> 1];
      }
      s += v0 * e0;
      s += v1 * e1;
      s += v2 * e2;
      s += v3 * e3;
    }
    for (; q < q1; q += G) {
      const int j = t.col[q];
      T ej = ein[j];
      if (ec) ej += ec[j >> 1];
      s += t.val[q] * ej;
    }
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
  };
  auto sweep = [=](const LdsLevel<T> t, const T* ein, const T* ec, T* eout) {
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
    const LdsLevel<T> t = tab[l - l0];
    const LdsLevel<T> tc = tab[l + 1 - l0];
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
  if (ainv) {
    // B = A^-1 r, one wave per row, in the same order as k_dense_rows (so a solve gives
    // the same bits packed with others or alone)
    const LdsLevel<T> t = tab[L - 1 - l0];
    const int n = t.n, lane = tid & 63;
    for (int row = tid >> 6; row < n; row += nth / 64) {
      const T* a = ainv + (size_t)row * n;
      T acc = 0;
      for (int j = lane; j < n; j += 64) acc += a[j] * t.r[j];
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if (lane == 0) t.B[row] = acc;
    }
    __syncthreads();
  } else {
    const LdsLevel<T> t = tab[L - 1 - l0];
    for (int i = tid; i < t.n; i += nth) t.B[i] = t.r[i] / t.diag[i];
    __syncthreads();
  }
  for (int l = L - 2; l >= l0; --l) {
    const LdsLevel<T> t = tab[l - l0];
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
    const LdsLevel<T> t = tab[0];
    T* out = g0.B;
    for (int i = tid; i < t.n; i += nth) out[i] = t.B[i];
  }
}

// LDS image of levels l0.. of one component: the LdsImgLevel table, then per level rp,
// col, val, diag (copied from `bytes` below) and the r, A, B work vectors (not copied).
// Returns the LDS bytes of the whole layout; with out != nullptr also writes the copied
// part (16-byte multiple) and sets *copy_bytes.
template <typename T>
size_t lds_image(const std::vector<int64_t>& n, const std::vector<int64_t>& nnz, int l0,
                 const SfcLevelsHost* h, std::vector<unsigned char>* out, size_t* copy_bytes) {
  const int L = (int)n.size();
  std::vector<LdsImgLevel> tab(MAX_LEVELS, LdsImgLevel{0, 0, 0, 0, 0, 0, 0, 0});
  size_t off = lds_align(sizeof(LdsImgLevel) * MAX_LEVELS);
  for (int l = l0; l < L; ++l) {
    LdsImgLevel& t = tab[l - l0];
    t.n = (int)n[l];
    t.rp = (int)off;
    off += lds_align(sizeof(int) * (n[l] + 1));
    t.col = (int)off;
    off += lds_align(sizeof(int) * nnz[l]);
    t.val = (int)off;
    off += lds_align(sizeof(T) * nnz[l]);
    t.diag = (int)off;
    off += lds_align(sizeof(T) * n[l]);
  }
  const size_t copied = off;
  for (int l = l0; l < L; ++l) {
    LdsImgLevel& t = tab[l - l0];
    t.r = (int)off;
    off += lds_align(sizeof(T) * n[l]);
    t.A = (int)off;
    off += lds_align(sizeof(T) * n[l]);
    t.B = (int)off;
    off += lds_align(sizeof(T) * n[l]);
  }
  if (out) {
    out->assign(copied, 0);
    std::memcpy(out->data(), tab.data(), sizeof(LdsImgLevel) * MAX_LEVELS);
    for (int l = l0; l < L; ++l) {
      const LdsImgLevel& t = tab[l - l0];
      int32_t* rp = reinterpret_cast<int32_t*>(out->data() + t.rp);
      int32_t* col = reinterpret_cast<int32_t*>(out->data() + t.col);
      T* val = reinterpret_cast<T*>(out->data() + t.val);
      T* diag = reinterpret_cast<T*>(out->data() + t.diag);
      for (int64_t i = 0; i <= n[l]; ++i) rp[i] = h->rowptr[l][i];
      for (int64_t k = 0; k < nnz[l]; ++k) {
        col[k] = h->col[l][k];
        val[k] = (T)h->val[l][k];
      }
      for (int64_t i = 0; i < n[l]; ++i) diag[i] = (T)h->diag[l][i];
    }
    *copy_bytes = copied;
  }
  return off;
}

template <typename T>
size_t lds_bytes_from(const std::vector<int64_t>& n, const std::vector<int64_t>& nnz, int l0) {
  return lds_image<T>(n, nnz, l0, nullptr, nullptr, nullptr);
}

// first level from which the remaining hierarchy fits in `budget` bytes of LDS in
// precision T
template <typename T>
int lds_start_level(const std::vector<int64_t>& n, const std::vector<int64_t>& nnz, size_t budget) {
  const int L = (int)n.size();
  if (L > MAX_LEVELS) return -1;
  int l0 = L - 1;
  for (int l = L - 1; l >= 0; --l) {
    if (lds_bytes_from<T>(n, nnz, l) <= budget) l0 = l;
    else break;
  }
  return l0;
}
template int lds_start_level<double>(const std::vector<int64_t>&, const std::vector<int64_t>&, size_t);
template int lds_start_level<float>(const std::vector<int64_t>&, const std::vector<int64_t>&, size_t);

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

// device copy of a host vector (converted to T), queued on s; nullptr on failure
template <typename T, typename U>
static T* upload_vec(dgmi_ctx* c, const std::vector<U>& h, hipStream_t s) {
  T* d = static_cast<T*>(c->alloc(sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!d) return nullptr;
  if (!h.empty()) {
    int rc;
    if constexpr (std::is_same<T, U>::value) {
      rc = h2d(c, d, h.data(), sizeof(T) * h.size(), s);
    } else {
      const std::vector<T> t(h.begin(), h.end());
      rc = h2d(c, d, t.data(), sizeof(T) * t.size(), s);
    }
    if (rc) {
      note_alloc_error("copy to device failed");
      return nullptr;
    }
  }
  return d;
}

template <typename T>
int launch_restrict(dgmi_ctx* c, const T* r_dg, T* r_sfc, hipStream_t s, T* e_first) {
  const int ncg = (int)c->mg.cg_nonods;
  const dim3 blk(256), grd((unsigned)(((int64_t)ncg * RESTRICT_LANES + 255) / 256));
  const bool f64 = std::is_same<T, double>::value;
  const T* rw = f64 ? (const T*)c->mg.rw64 : (const T*)c->mg.rw32;
  const T* dg0 = nullptr;
  if (e_first) {
    if (!vcycle_fuses_first<T>(c)) {
      set_error("first sweep not fusable into the restriction");
      return -2;
    }
    dg0 = f64 ? (const T*)c->mg.lv.diag64[0] : (const T*)c->mg.lv.diag32[0];
  }
  const T om = (T)c->omega;
  if (c->D == 3)
    hipLaunchKernelGGL((k_restrict<T, 3>), grd, blk, 0, s, ncg, c->mg.rptr, c->mg.rent, rw, r_dg,
                       r_sfc, dg0, e_first, om);
  else
    hipLaunchKernelGGL((k_restrict<T, 2>), grd, blk, 0, s, ncg, c->mg.rptr, c->mg.rent, rw, r_dg,
                       r_sfc, dg0, e_first, om);
  DGMI_HIP(hipGetLastError());
  return 0;
}

// FP32 vector family: x += P e and the traces of the new x into buffer tout (0 or 1)
int launch_prolong_trace(dgmi_ctx* c, const float* e_sfc, float* x_dg, int tout, hipStream_t s) {
  const int nele = (int)c->nele;
  DevGeom<float>& g = c->g32;
  if (!ensure_traces(c, g)) return -3;
  typename Vec2<float>::type* to = tout == 0 ? g.trace : g.trace2;
  if (!to || (tout != 0 && tout != 1)) {
    set_error("bad trace buffer");
    return -2;
  }
  const dim3 blk(64 * (c->D + 1)), grd((nele + 63) / 64);
  const float* rw = (const float*)c->mg.rw32;
  const float* fc = g.tab + g.off_fc;
  if (c->D == 3)
    hipLaunchKernelGGL((k_prolong_trace<float, 3>), grd, blk, 0, s, nele, c->mg.cellpos, rw, e_sfc,
                       x_dg, fc, c->fnodes, g.cf, to);
  else
    hipLaunchKernelGGL((k_prolong_trace<float, 2>), grd, blk, 0, s, nele, c->mg.cellpos, rw, e_sfc,
                       x_dg, fc, c->fnodes, g.cf, to);
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

int launch_rar_assemble(dgmi_ctx* c, const double* rarm, double* ke, const int32_t* rlo,
                        const int32_t* rhi, const int32_t* colpos, int64_t nslot,
                        double* vals, hipStream_t s) {
  const int nele = (int)c->nele, nb = c->D + 1;
  const int ncg = (int)c->mg.cg_nonods;
  double* miss = vals + nslot;
  const dim3 blk(256), grd((unsigned)(((int64_t)nele * nb * nb + 255) / 256));
  const dim3 grd2((unsigned)((ncg + 255) / 256));
  if (c->D == 3) {
    hipLaunchKernelGGL(k_rar_elem<3>, grd, blk, 0, s, nele, rarm, c->g64.G, c->g64.cf,
                       c->g64.wf, c->g64.pen, c->fcode, ke, miss);
    hipLaunchKernelGGL(k_rar_gather<3>, grd2, blk, 0, s, ncg, c->mg.rptr, c->mg.rent,
                       c->mg.cellpos, c->mg.rw64, ke, rlo, rhi, colpos, vals, miss);
  } else {
    hipLaunchKernelGGL(k_rar_elem<2>, grd, blk, 0, s, nele, rarm, c->g64.G, c->g64.cf,
                       c->g64.wf, c->g64.pen, c->fcode, ke, miss);
    hipLaunchKernelGGL(k_rar_gather<2>, grd2, blk, 0, s, ncg, c->mg.rptr, c->mg.rent,
                       c->mg.cellpos, c->mg.rw64, ke, rlo, rhi, colpos, vals, miss);
  }
  DGMI_HIP(hipGetLastError());
  return 0;
}

// One V-cycle (z02/multigrid_linearelastic.py:112-172) of every component.  Level buffers:
// A_l/B_l ping-pong, level 0's B is the caller's output e_sfc and level 0's rhs is the
// caller's r_sfc.  Grid levels 0..small_level-1 sweep the rows of the components that
// start deeper; one work-group per component then runs the rest.
// y[i] = x[i] (or 0 with x null): replaces memset/memcpy nodes so the cycle can be
// captured in a graph
// out = A^-1 r for one dense level (row-major inverse): one wave per row, lanes over the
// columns, a fixed shuffle tree
template <typename T>
__global__ __launch_bounds__(256) void k_dense_rows(int n, const T* __restrict__ ainv,
                                                    const T* __restrict__ r, T* __restrict__ out) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;  // whole wave
  const T* a = ainv + (size_t)row * n;
  T acc = 0;
  // unrolled for loads in flight; the additions keep their order
#pragma unroll 16
  for (int j = lane; j < n; j += 64) acc += a[j] * r[j];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) out[row] = acc;
}

template <typename T>
__global__ void k_copy(int64_t n, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x ? x[i] : T(0);
}

// Grid-sized levels of a V-cycle (those before the single-work-group or dense end).
template <typename T> int vcycle_grid_levels(const dgmi_ctx* c) {
  const auto& mg = c->mg;
  const bool f64 = std::is_same<T, double>::value;
  const bool dense = !f64 && mg.coarse_exact && mg.lv.dense_rm32;
  return f64 ? mg.small_level : (dense ? mg.lv.dense_level32 : mg.small_level32);
}

// The first sweep from zero can ride on the restriction when level 0 is a grid level of
// every component and the cycle is launched on the stream.
template <typename T> bool vcycle_fuses_first(const dgmi_ctx* c) {
  const auto& lv = c->mg.lv;
  const bool f64 = std::is_same<T, double>::value;
  const bool any_start0 = f64 ? lv.any_start0 : lv.any_start0_32;
  const auto& grid_rows = f64 ? lv.grid_rows : lv.grid_rows32;
  return !c->mg.use_graph && vcycle_grid_levels<T>(c) > 0 && !any_start0 &&
         !grid_rows.empty() && grid_rows[0] == c->mg.cg_nonods && lv.n[0] == c->mg.cg_nonods;
}

template <typename T>
static int vcycle_body(dgmi_ctx* c, const T* r_sfc, T* e_sfc, hipStream_t s, bool first_done) {
  auto& mg = c->mg;
  auto& lv = mg.lv;
  const bool f64 = std::is_same<T, double>::value;
  const T omega = (T)c->omega;
  const bool dense = !f64 && mg.coarse_exact && lv.dense_rm32;
  const int l0 = vcycle_grid_levels<T>(c);
  const std::vector<int64_t>& grid_rows = f64 ? lv.grid_rows : lv.grid_rows32;
  const bool any_start0 = f64 ? lv.any_start0 : lv.any_start0_32;
  auto R = [&](int l) -> T* { return l == 0 ? const_cast<T*>(r_sfc) : vec<T>(lv.r[l]); };
  auto A = [&](int l) -> T* { return vec<T>(lv.e[l]); };
  auto B = [&](int l) -> T* { return l == 0 ? e_sfc : vec<T>(lv.e2[l]); };
  auto val = [&](int l) -> const T* { return f64 ? (const T*)lv.val64[l] : (const T*)lv.val32[l]; };
  auto dia = [&](int l) -> const T* { return f64 ? (const T*)lv.diag64[l] : (const T*)lv.diag32[l]; };
  const dim3 blk(256);
  const dim3 gcp((unsigned)((lv.n[0] + 255) / 256));
  // the cycle starts from e = 0 on level 0: a grid level 0 takes it in its first sweep,
  // components whose level 0 is small read it from A(0)
  if (l0 == 0 || any_start0)
    hipLaunchKernelGGL((k_copy<T>), gcp, blk, 0, s, lv.n[0], (const T*)nullptr, A(0));
  // down-cycle on grid-sized levels
  for (int l = 0; l < l0; ++l) {
    const int n = (int)grid_rows[l];
    const dim3 grd((unsigned)(((int64_t)n * ROW_LANES + 255) / 256));
    T* cur = A(l);
    T* nxt = B(l);
    for (int it = 0; it < 3; ++it) {
      if (l == 0 && it == 0 && first_done) {
        // the restriction has written e = omega r / diag into nxt (== e_sfc)
      } else if (l == 0 && it == 0)
        hipLaunchKernelGGL((k_sweep_zero<T>), dim3((unsigned)((n + 255) / 256)), blk, 0, s, n,
                           dia(l), (const T*)R(l), nxt, omega);
      else
        hipLaunchKernelGGL((k_sweep<T>), grd, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l),
                           dia(l), R(l), cur, (const T*)nullptr, nxt, omega);
      std::swap(cur, nxt);
    }
    // after 3 sweeps cur == B(l); 4th sweep writes A(l) and the restricted r, e
    const int nc = (n - 1) / 2 + 1;
    const dim3 grc((unsigned)(((int64_t)nc * 2 * ROW_LANES + 255) / 256));
    hipLaunchKernelGGL((k_sweep_restrict<T>), grc, blk, 0, s, n, lv.rowptr[l], lv.col[l], val(l),
                       dia(l), R(l), cur, nxt, R(l + 1), A(l + 1), omega);
  }
  if (dense) {
    // one solve whose exact level is its first non-grid level: B = A^-1 r across the GPU
    const int n = lv.dense_n32;
    hipLaunchKernelGGL((k_dense_rows<T>), dim3((unsigned)((n + 3) / 4)), blk, 0, s, n,
                       (const T*)lv.dense_rm32, (const T*)R(l0), B(l0));
  } else {
    // the level tables point level 0 at internal buffers, so components that start at
    // level 0 see the caller's vectors through them
    if (any_start0)
      hipLaunchKernelGGL((k_copy<T>), gcp, blk, 0, s, lv.n[0], r_sfc, vec<T>(lv.r[0]));
    hipLaunchKernelGGL((k_small_levels_lds<T>), dim3(c->ncomp), dim3(SMALL_THREADS),
                       f64 ? lv.lds64 : lv.lds32, s,
                       (const LevelDesc<T>*)(f64 ? lv.desc64 : lv.desc32),
                       (const int32_t*)(f64 ? lv.rng64 : lv.rng32),
                       (const uint4*)(f64 ? lv.img64 : lv.img32),
                       (const int64_t*)(f64 ? lv.imginfo64 : lv.imginfo32), omega, 3, 3,
                       (const T*)(f64 || !mg.coarse_exact ? nullptr : lv.cinv32),
                       (const int64_t*)lv.cinfo32);
    if (any_start0)
      hipLaunchKernelGGL((k_copy<T>), gcp, blk, 0, s, lv.n[0], (const T*)vec<T>(lv.e2[0]), e_sfc);
  }
  // up-cycle on grid-sized levels
  for (int l = l0 - 1; l >= 0; --l) {
    const int n = (int)grid_rows[l];
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
int launch_sfc_vcycle(dgmi_ctx* c, const T* r_sfc, T* e_sfc, hipStream_t s, bool first_done) {
  auto& mg = c->mg;
  if (!mg.use_graph) return vcycle_body<T>(c, r_sfc, e_sfc, s, first_done);
  if (first_done) {
    set_error("a captured V-cycle takes its own first sweep");
    return -2;
  }
  VGraph& vg = std::is_same<T, double>::value ? mg.graph64 : mg.graph32;
  if (!vg.exec || vg.r != (const void*)r_sfc || vg.e != (void*)e_sfc || vg.omega != c->omega) {
    if (vg.exec) {
      DGMI_HIP(hipGraphExecDestroy(vg.exec));
      vg.exec = nullptr;
    }
    if (!mg.capture_stream)
      DGMI_HIP(hipStreamCreateWithFlags(&mg.capture_stream, hipStreamNonBlocking));
    DGMI_HIP(hipStreamBeginCapture(mg.capture_stream, hipStreamCaptureModeRelaxed));
    const int rc = vcycle_body<T>(c, r_sfc, e_sfc, mg.capture_stream, false);
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

// In-place blocked Gauss-Jordan inversion of dense N x N matrices (row-major, FP64, N a
// multiple of GJB) on the device, without pivoting: the SFC level operators are Galerkin
// projections of a symmetric positive definite operator, so every pivot block is positive
// definite (a non-positive or non-finite pivot sets the matrix's fail flag).  Block column
// K = [k0, k0+GJB), R = the rest:  P = A_KK^-1;  A_KR <- P A_KR;  A_RR <- A_RR - A_RK A_KR;
// A_RK <- -A_RK P;  A_KK <- P  (the block form of the scalar steps a_kk <- 1/a_kk, row k
// /= a_kk, a_ij -= a_ik a_kj, column k *= -1/a_kk).  Several matrices of different sizes
// are inverted together: blockIdx.z selects the matrix, and a launch for block column k0
// skips the matrices already done.
constexpr int GJB = 64;

struct GjMat {
  double* a;        // N x N work matrix
  double* p;        // GJB x GJB pivot block
  float* out;       // n x n FP32 row-major inverse
  const int32_t* rp;  // CSR of the level (n rows)
  const int32_t* col;
  const double* val;
  int n, N, fail;   // fail: index into the fail flags
};

// Entries of one CSR level stored into the zeroed dense matrix, and the identity on the
// padding rows i >= n: one thread per row of the padded matrix.
__global__ __launch_bounds__(256) void k_csr_to_dense(const GjMat* __restrict__ mats) {
  const GjMat m = mats[blockIdx.z];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  double* row = m.a + (size_t)i * m.N;
  if (i >= m.n) {
    row[i] = 1.0;
    return;
  }
  for (int32_t q = m.rp[i]; q < m.rp[i + 1]; ++q) row[m.col[q]] += m.val[q];
}

// P = A_KK^-1 by unblocked Gauss-Jordan in LDS; written to A_KK and to p (GJB x GJB).
// Wave q updates columns [q GJB/4, (q+1) GJB/4) of every row, lane r row r, from the
// scaled pivot row that wave 0 publishes: two barriers and GJB/4 FMAs per lane a step.
constexpr int GJP_WAVES = 4, GJP_COLS = GJB / GJP_WAVES;
__global__ __launch_bounds__(64 * GJP_WAVES) void k_bgj_pivot(const GjMat* __restrict__ mats,
                                                              int k0, int* __restrict__ fail) {
  static_assert(GJB == 64, "one lane per row");
  const GjMat m = mats[blockIdx.z];
  if (k0 >= m.N) return;
  const int N = m.N;
  double* A = m.a;
  __shared__ double a[GJB][GJB + 1];
  __shared__ double prow[GJB];
  const int t = threadIdx.x & 63, q = threadIdx.x >> 6, j0 = q * GJP_COLS;
  for (int x = threadIdx.x; x < GJB * GJB; x += 64 * GJP_WAVES)
    a[x / GJB][x % GJB] = A[(size_t)(k0 + x / GJB) * N + k0 + x % GJB];
  __syncthreads();
  bool bad = false;
  for (int k = 0; k < GJB; ++k) {
    const double piv = a[k][k];
    bad = bad || !(piv > 0.0) || !isfinite(piv);
    const double inv = 1.0 / piv;
    const double f = a[t][k];
    // row k scaled (its column k becomes 1/piv), read by every lane after the barrier
    if (q == 0) prow[t] = t == k ? inv : a[k][t] * inv;
    __syncthreads();
    // row k: prow (sc = 0, fx = -1 give it exactly); the others: a - f prow
    const double sc = t == k ? 0.0 : 1.0, fx = t == k ? -1.0 : f;
    double rw[GJP_COLS];
#pragma unroll
    for (int j = 0; j < GJP_COLS; ++j) rw[j] = fma(-fx, prow[j0 + j], a[t][j0 + j] * sc);
#pragma unroll
    for (int j = 0; j < GJP_COLS; ++j) a[t][j0 + j] = rw[j];
    if (t != k && k >= j0 && k < j0 + GJP_COLS) a[t][k] = -f * inv;
    __syncthreads();
  }
  if (threadIdx.x == 0 && bad) fail[m.fail] = 1;
  for (int x = threadIdx.x; x < GJB * GJB; x += 64 * GJP_WAVES) {
    const double v = a[x / GJB][x % GJB];
    A[(size_t)(k0 + x / GJB) * N + k0 + x % GJB] = v;
    m.p[x] = v;
  }
}

// A_KJ <- P A_KJ for the column tile J (one work-group per tile, the pivot tile skipped)
__global__ __launch_bounds__(256) void k_bgj_row(const GjMat* __restrict__ mats, int k0) {
  const GjMat m = mats[blockIdx.z];
  const int j0 = blockIdx.x * GJB;
  if (k0 >= m.N || j0 >= m.N || j0 == k0) return;
  const int N = m.N;
  double* A = m.a;
  __shared__ double ps[GJB][GJB + 1], ys[GJB][GJB + 1];
  const int t = threadIdx.x;
  for (int q = t; q < GJB * GJB; q += 256) {
    ps[q / GJB][q % GJB] = m.p[q];
    ys[q / GJB][q % GJB] = A[(size_t)(k0 + q / GJB) * N + j0 + q % GJB];
  }
  __syncthreads();
  const int tx = t % 16, ty = t / 16;
  double c[4][4] = {};
  for (int k = 0; k < GJB; ++k) {
    double u[4], v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = ps[ty * 4 + i][k];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = ys[k][tx * 4 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) c[i][j] += u[i] * v[j];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) A[(size_t)(k0 + ty * 4 + i) * N + j0 + tx * 4 + j] = c[i][j];
}

// A_IJ -= A_IK A_KJ for tiles I, J outside the pivot block (A_KJ already updated)
__global__ __launch_bounds__(256) void k_bgj_rest(const GjMat* __restrict__ mats, int k0) {
  const GjMat m = mats[blockIdx.z];
  const int i0 = blockIdx.y * GJB, j0 = blockIdx.x * GJB;
  if (k0 >= m.N || i0 >= m.N || j0 >= m.N || i0 == k0 || j0 == k0) return;
  const int N = m.N;
  double* A = m.a;
  __shared__ double ls[GJB][GJB + 1], rs[GJB][GJB + 1];
  const int t = threadIdx.x;
  for (int q = t; q < GJB * GJB; q += 256) {
    ls[q / GJB][q % GJB] = A[(size_t)(i0 + q / GJB) * N + k0 + q % GJB];
    rs[q / GJB][q % GJB] = A[(size_t)(k0 + q / GJB) * N + j0 + q % GJB];
  }
  __syncthreads();
  const int tx = t % 16, ty = t / 16;
  double c[4][4] = {};
  for (int k = 0; k < GJB; ++k) {
    double u[4], v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = ls[ty * 4 + i][k];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = rs[k][tx * 4 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) c[i][j] += u[i] * v[j];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) A[(size_t)(i0 + ty * 4 + i) * N + j0 + tx * 4 + j] -= c[i][j];
}

// A_IK <- -A_IK P for the row tile I (the pivot tile skipped)
__global__ __launch_bounds__(256) void k_bgj_col(const GjMat* __restrict__ mats, int k0) {
  const GjMat m = mats[blockIdx.z];
  const int i0 = blockIdx.x * GJB;
  if (k0 >= m.N || i0 >= m.N || i0 == k0) return;
  const int N = m.N;
  double* A = m.a;
  __shared__ double ls[GJB][GJB + 1], ps[GJB][GJB + 1];
  const int t = threadIdx.x;
  for (int q = t; q < GJB * GJB; q += 256) {
    ps[q / GJB][q % GJB] = m.p[q];
    ls[q / GJB][q % GJB] = A[(size_t)(i0 + q / GJB) * N + k0 + q % GJB];
  }
  __syncthreads();
  const int tx = t % 16, ty = t / 16;
  double c[4][4] = {};
  for (int k = 0; k < GJB; ++k) {
    double u[4], v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = ls[ty * 4 + i][k];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = ps[k][tx * 4 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) c[i][j] += u[i] * v[j];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) A[(size_t)(i0 + ty * 4 + i) * N + k0 + tx * 4 + j] = -c[i][j];
}

// out (n x n, FP32 row-major) = the leading n x n block of a (N x N, FP64)
__global__ void k_d2f_block(const GjMat* __restrict__ mats) {
  const GjMat m = mats[blockIdx.z];
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)m.n * m.n) return;
  const int i = (int)(q / m.n), j = (int)(q - (int64_t)i * m.n);
  m.out[q] = (float)m.a[(size_t)i * m.N + j];
}

// One level inverse to form: level l of hierarchy h, FP32 result written to out.
struct LevelInverse {
  const SfcLevelsHost* h;
  int l;
  float* out;
};

// FP64 blocked Gauss-Jordan inverses of the given levels, formed together on stream s and
// written as FP32 row-major; returns 0 and fails[i] = 1 for a level that is not positive
// definite, or an error code (the stream is synchronised on return).
static int invert_levels(dgmi_ctx* c, const std::vector<LevelInverse>& jobs,
                         std::vector<int>& fails, hipStream_t s) {
  const int nm = (int)jobs.size();
  fails.assign(nm, 0);
  if (nm == 0) return 0;
  std::vector<GjMat> mats(nm);
  size_t asz = 0, nnz = 0, nrp = 0;
  int nbmax = 0, nmax = 0;
  for (int i = 0; i < nm; ++i) {
    const int n = (int)jobs[i].h->n[jobs[i].l];
    const int N = (n + GJB - 1) / GJB * GJB;
    mats[i].n = n;
    mats[i].N = N;
    mats[i].fail = i;
    asz += (size_t)N * N;
    nnz += jobs[i].h->col[jobs[i].l].size();
    nrp += (size_t)n + 1;
    nbmax = std::max(nbmax, N / GJB);
    nmax = std::max(nmax, N);
  }
  double *da = nullptr, *dp = nullptr, *dval = nullptr;
  int32_t *drp = nullptr, *dcol = nullptr;
  int* dfail = nullptr;
  GjMat* dm = nullptr;
  int rc = 0;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && !rc) {
      set_error(std::string("coarse inverse: ") + hipGetErrorString(e));
      rc = -3;
    }
    return rc == 0;
  };
  auto take = [&](auto** d, size_t bytes) {
    *d = static_cast<std::remove_pointer_t<decltype(d)>>(dev_alloc(bytes));
    return chk(*d ? hipSuccess : hipErrorOutOfMemory);
  };
  if (take(&da, sizeof(double) * asz) && take(&dp, sizeof(double) * GJB * GJB * nm) &&
      take(&dval, sizeof(double) * std::max<size_t>(nnz, 1)) &&
      take(&dcol, sizeof(int32_t) * std::max<size_t>(nnz, 1)) &&
      take(&drp, sizeof(int32_t) * nrp) && take(&dfail, sizeof(int) * nm) &&
      take(&dm, sizeof(GjMat) * nm)) {
    size_t ao = 0, vo = 0, ro = 0;
    for (int i = 0; i < nm && rc == 0; ++i) {
      const SfcLevelsHost& h = *jobs[i].h;
      const int l = jobs[i].l;
      const size_t k = h.col[l].size();
      mats[i].a = da + ao;
      mats[i].p = dp + (size_t)GJB * GJB * i;
      mats[i].out = jobs[i].out;
      mats[i].rp = drp + ro;
      mats[i].col = dcol + vo;
      mats[i].val = dval + vo;
      if (h2d(c, drp + ro, h.rowptr[l].data(), sizeof(int32_t) * (mats[i].n + 1), s) && !rc)
        rc = -3;
      if (k && rc == 0 && (h2d(c, dcol + vo, h.col[l].data(), sizeof(int32_t) * k, s) ||
                           h2d(c, dval + vo, h.val[l].data(), sizeof(double) * k, s)))
        rc = -3;
      ao += (size_t)mats[i].N * mats[i].N;
      vo += k;
      ro += (size_t)mats[i].n + 1;
    }
    if (rc == 0 && (h2d(c, dm, mats.data(), sizeof(GjMat) * nm, s) == 0 || (rc = -3, false)) &&
        chk(hipMemsetAsync(dfail, 0, sizeof(int) * nm, s)) &&
        chk(hipMemsetAsync(da, 0, sizeof(double) * asz, s))) {
      const dim3 blk(256);
      hipLaunchKernelGGL(k_csr_to_dense, dim3((nmax + 255) / 256, 1, nm), blk, 0, s,
                         (const GjMat*)dm);
      for (int kb = 0; kb < nbmax; ++kb) {
        const int k0 = kb * GJB;
        hipLaunchKernelGGL(k_bgj_pivot, dim3(1, 1, nm), dim3(64 * GJP_WAVES), 0, s,
                           (const GjMat*)dm, k0, dfail);
        hipLaunchKernelGGL(k_bgj_row, dim3(nbmax, 1, nm), blk, 0, s, (const GjMat*)dm, k0);
        hipLaunchKernelGGL(k_bgj_rest, dim3(nbmax, nbmax, nm), blk, 0, s, (const GjMat*)dm, k0);
        hipLaunchKernelGGL(k_bgj_col, dim3(nbmax, 1, nm), blk, 0, s, (const GjMat*)dm, k0);
      }
      const int64_t nn = (int64_t)nmax * nmax;
      hipLaunchKernelGGL(k_d2f_block, dim3((unsigned)((nn + 255) / 256), 1, nm), blk, 0, s,
                         (const GjMat*)dm);
      chk(hipGetLastError());
      if (rc == 0 && d2h(fails.data(), dfail, sizeof(int) * nm, s)) rc = -3;
    }
  }
  (void)hipStreamSynchronize(s);  // nothing in flight reads the work space any more
  for (void* q : {(void*)da, (void*)dp, (void*)dval, (void*)dcol, (void*)drp, (void*)dfail,
                  (void*)dm})
    dev_free(q);
  return rc;
}

// FP32 row-major inverse of the dense level-l operator of one component, formed on the
// device on stream s (synchronous with s only); nullptr on failure.
static float* dense_inverse_device(dgmi_ctx* c, const SfcLevelsHost& h, int l, hipStream_t s) {
  const int n = (int)h.n[l];
  float* out = static_cast<float*>(c->alloc(sizeof(float) * (size_t)n * n));
  if (!out) return nullptr;
  std::vector<int> fails;
  if (invert_levels(c, {LevelInverse{&h, l, out}}, fails, s) != 0 || fails[0]) return nullptr;
  return out;
}

// Device layout of every component's SFC hierarchy (see SfcLevelsDev) and the level
// tables of the single-work-group kernel.
int upload_sfc_levels(dgmi_ctx* c, const SfcLayout& lay, hipStream_t s) {
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
  lv.grid_rows32.assign(S, 0);
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
    // components swept at level l, in layout order: their rows and entries side by side,
    // each segment padded to its layout length with empty rows (diagonal 1)
    std::vector<int> comps;
    std::vector<int64_t> rowbase(1, 0), nzbase(1, 0);
    for (int q : lay.order) {
      if (lay.start[q] < l) continue;
      const SfcLevelsHost& h = lay.lh[q];
      const int64_t seg = (q == lay.order.back()) ? h.n[l] : (lay.len0[q] >> l);
      const int64_t rows = rowbase.back();
      segoff[l][q] = rows;
      rpoff[l][q] = rows;
      if (l < S && lay.start[q] > l) lv.grid_rows[l] = rows + seg;
      if (l < S && lay.start32[q] > l) lv.grid_rows32[l] = rows + seg;
      comps.push_back(q);
      rowbase.push_back(rows + seg);
      nzbase.push_back(nzbase.back() + (int64_t)h.col[l].size());
    }
    const int64_t R = rowbase.back(), NZ = nzbase.back();
    std::vector<int32_t> rp(R + 1), col(NZ);
    std::vector<double> val(NZ), diag(R);
    rp[R] = (int32_t)NZ;
    parallel_chunks(R, [&](int, int64_t a, int64_t b) {
      int k = (int)(std::upper_bound(rowbase.begin(), rowbase.end(), a) - rowbase.begin()) - 1;
      for (int64_t r = a; r < b; ++r) {
        while (r >= rowbase[k + 1]) ++k;
        const SfcLevelsHost& h = lay.lh[comps[k]];
        const int64_t i = r - rowbase[k];
        const int64_t n = h.n[l];
        if (i < n) {
          const int32_t q0 = h.rowptr[l][i], q1 = h.rowptr[l][i + 1];
          rp[r] = (int32_t)(nzbase[k] + q0);
          for (int32_t q = q0; q < q1; ++q) {
            col[nzbase[k] + q] = (int32_t)(h.col[l][q] + rowbase[k]);
            val[nzbase[k] + q] = h.val[l][q];
          }
          diag[r] = h.diag[l][i];
        } else {
          rp[r] = (int32_t)(nzbase[k] + h.rowptr[l][n]);
          diag[r] = 1.0;  // padding row: no entries, stays zero
        }
      }
    });
    const int64_t rows = R;
    lv.n[l] = rows;
    lv.nnz[l] = NZ;
    lv.rowptr[l] = upload_vec<int32_t>(c, rp, s);
    lv.col[l] = upload_vec<int32_t>(c, col, s);
    lv.val64[l] = upload_vec<double>(c, val, s);
    lv.diag64[l] = upload_vec<double>(c, diag, s);
    lv.val32[l] = static_cast<float*>(c->alloc(sizeof(float) * std::max<int64_t>(NZ, 1)));
    lv.diag32[l] = static_cast<float*>(c->alloc(sizeof(float) * std::max<int64_t>(rows, 1)));
    lv.r[l] = c->alloc(sizeof(double) * std::max<int64_t>(rows, 1));
    lv.e[l] = c->alloc(sizeof(double) * std::max<int64_t>(rows, 1));
    lv.e2[l] = c->alloc(sizeof(double) * std::max<int64_t>(rows, 1));
    if (!lv.rowptr[l] || !lv.col[l] || !lv.val64[l] || !lv.diag64[l] || !lv.val32[l] ||
        !lv.diag32[l] || !lv.r[l] || !lv.e[l] || !lv.e2[l])
      return fail();
    // FP32 copies converted on the device
    if (NZ) {
      if (int rc = launch_convert<float>(lv.val64[l], lv.val32[l], NZ, s)) return rc;
    }
    if (rows) {
      if (int rc = launch_convert<float>(lv.diag64[l], lv.diag32[l], rows, s)) return rc;
    }
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
  lv.pool_rp = upload_vec<int32_t>(c, prp, s);
  lv.pool_col = upload_vec<int32_t>(c, pcol, s);
  lv.pool_val64 = upload_vec<double>(c, pval, s);
  lv.pool_diag64 = upload_vec<double>(c, pdiag, s);
  lv.pool_val32 = upload_vec<float>(c, pval, s);
  lv.pool_diag32 = upload_vec<float>(c, pdiag, s);
  lv.pool_vec = c->alloc(sizeof(double) * std::max<int64_t>(pvec, 1));
  if (!lv.pool_rp || !lv.pool_col || !lv.pool_val64 || !lv.pool_diag64 || !lv.pool_val32 ||
      !lv.pool_diag32 || !lv.pool_vec)
    return fail();
  // ---- level tables, components in layout order (one work-group each).  A level from
  // the start up to the component's FP64 start is read from the grid layout, deeper ones
  // from the pool.
  std::vector<LevelDesc<double>> d64((size_t)nc * MAX_LEVELS);
  std::vector<LevelDesc<float>> d32((size_t)nc * MAX_LEVELS);
  std::vector<int32_t> rng64(2 * nc), rng32(2 * nc);
  size_t lds64 = 0, lds32 = 0;
  lv.any_start0 = false;
  lv.any_start0_32 = false;
  auto fill = [&](auto& d, int q, int l, auto* val_grid, auto* diag_grid, auto* val_pool,
                  auto* diag_pool) {
    using TT = std::remove_pointer_t<decltype(val_grid)>;
    const SfcLevelsHost& h = lay.lh[q];
    d.n = (int)h.n[l];
    if (l <= lay.start[q]) {
      const int64_t so = segoff[l][q];
      d.cbase = (int)so;
      d.rp = lv.rowptr[l] + rpoff[l][q];
      d.col = lv.col[l];
      d.val = val_grid;
      d.diag = diag_grid + so;
      d.r = (TT*)lv.r[l] + so;
      d.A = (TT*)lv.e[l] + so;
      d.B = (TT*)lv.e2[l] + so;
    } else {
      const PoolRef& p = pref[q][l];
      d.cbase = 0;
      d.rp = lv.pool_rp + p.rp;
      d.col = lv.pool_col;
      d.val = val_pool;
      d.diag = diag_pool + p.diag;
      TT* v = (TT*)lv.pool_vec + p.vec;  // FP32 views use the first half
      d.r = v;
      d.A = v + h.n[l];
      d.B = v + 2 * h.n[l];
    }
  };
  std::vector<unsigned char> img64, img32, piece;
  std::vector<int64_t> info64(2 * nc), info32(2 * nc);
  for (int b = 0; b < nc; ++b) {
    const int q = lay.order[b];
    const SfcLevelsHost& h = lay.lh[q];
    const int s0 = lay.start[q], s32 = lay.start32[q];
    if (h.nlevel > MAX_LEVELS) {
      set_error("too many SFC levels");
      return -2;
    }
    rng64[2 * b] = s0;
    rng64[2 * b + 1] = h.nlevel;
    rng32[2 * b] = s32;
    rng32[2 * b + 1] = h.nlevel;
    lv.any_start0 |= s0 == 0;
    lv.any_start0_32 |= s32 == 0;
    std::vector<int64_t> nnzq(h.nlevel);
    for (int l = 0; l < h.nlevel; ++l) nnzq[l] = (int64_t)h.col[l].size();
    size_t cb = 0;
    lds64 = std::max(lds64, lds_image<double>(h.n, nnzq, s0, &h, &piece, &cb));
    info64[2 * b] = (int64_t)(img64.size() / 16);
    info64[2 * b + 1] = (int64_t)(cb / 16);
    img64.insert(img64.end(), piece.begin(), piece.end());
    lds32 = std::max(lds32, lds_image<float>(h.n, nnzq, s32, &h, &piece, &cb));
    info32[2 * b] = (int64_t)(img32.size() / 16);
    info32[2 * b + 1] = (int64_t)(cb / 16);
    img32.insert(img32.end(), piece.begin(), piece.end());
    for (int l = s0; l < h.nlevel; ++l)
      fill(d64[(size_t)b * MAX_LEVELS + l], q, l, lv.val64[std::min(l, S)],
           lv.diag64[std::min(l, S)], lv.pool_val64, lv.pool_diag64);
    for (int l = s32; l < h.nlevel; ++l)
      fill(d32[(size_t)b * MAX_LEVELS + l], q, l, lv.val32[std::min(l, S)],
           lv.diag32[std::min(l, S)], lv.pool_val32, lv.pool_diag32);
  }
  // FP32 exact coarse solve: the first level at or below a component's FP32 start with at
  // most EXACT_MAX_N rows (and a coarser level below it) is inverted on the device, every
  // component's in turn on the stream
  std::vector<int64_t> cinfo(2 * nc, -1);
  std::vector<int> lvl(nc, -1);
  int64_t ctotal = 0;
  for (int b = 0; b < nc; ++b) {
    const int q = lay.order[b];
    const SfcLevelsHost& h = lay.lh[q];
    cinfo[2 * b] = ctotal;
    for (int l = lay.start32[q]; l < h.nlevel - 1; ++l)
      if (h.n[l] <= EXACT_MAX_N) {
        lvl[b] = l;
        ctotal += h.n[l] * h.n[l];
        break;
      }
  }
  lv.cinv32 = static_cast<float*>(c->alloc(sizeof(float) * std::max<int64_t>(ctotal, 1)));
  if (!lv.cinv32) return fail();
  {
    std::vector<LevelInverse> jobs;
    std::vector<int> which;
    for (int b = 0; b < nc; ++b)
      if (lvl[b] >= 0) {
        jobs.push_back({&lay.lh[lay.order[b]], lvl[b], lv.cinv32 + cinfo[2 * b]});
        which.push_back(b);
      }
    std::vector<int> fails;
    if (int rc = invert_levels(c, jobs, fails, s)) return rc;
    for (size_t i = 0; i < which.size(); ++i)
      if (fails[i]) lvl[which[i]] = -1;  // not positive definite: this component sweeps every level
  }
  for (int b = 0; b < nc; ++b) cinfo[2 * b + 1] = lvl[b];
  lv.cinfo32 = upload_vec<int64_t>(c, cinfo, s);
  if (!lv.cinfo32) return fail();
  mg.exact32.assign(nc, -1);
  for (int b = 0; b < nc; ++b) mg.exact32[lay.order[b]] = (int)cinfo[2 * b + 1];
  // one solve: the exact level is the first of at most dense_max_n rows (inverted on the
  // device) and a grid kernel applies its inverse, in place of the single-work-group
  // kernel and of the grid levels from there down
  lv.dense_rm32 = nullptr;
  lv.dense_n32 = 0;
  lv.dense_level32 = -1;
  if (nc == 1) {
    const SfcLevelsHost& h = lay.lh[lay.order[0]];
    for (int l = 0; l <= std::min(lay.start32[lay.order[0]], h.nlevel - 2); ++l)
      if (h.n[l] <= c->dense_max_n) {
        // the level the packed path inverts: reuse that inverse, so that a solve alone
        // and packed with others gives the same bits
        if (float* inv = l == lvl[0] ? lv.cinv32 : dense_inverse_device(c, h, l, s)) {
          lv.dense_rm32 = inv;
          lv.dense_n32 = (int)h.n[l];
          lv.dense_level32 = l;
          mg.exact32[lay.order[0]] = l;
        }
        break;
      }
  }
  lv.desc64 = upload_vec<LevelDesc<double>>(c, d64, s);
  lv.desc32 = upload_vec<LevelDesc<float>>(c, d32, s);
  lv.rng64 = upload_vec<int32_t>(c, rng64, s);
  lv.rng32 = upload_vec<int32_t>(c, rng32, s);
  lv.img64 = upload_vec<unsigned char>(c, img64, s);
  lv.img32 = upload_vec<unsigned char>(c, img32, s);
  lv.imginfo64 = upload_vec<int64_t>(c, info64, s);
  lv.imginfo32 = upload_vec<int64_t>(c, info32, s);
  if (!lv.desc64 || !lv.desc32 || !lv.rng64 || !lv.rng32 || !lv.img64 || !lv.img32 ||
      !lv.imginfo64 || !lv.imginfo32)
    return fail();
  lv.lds64 = lds64;
  lv.lds32 = lds32;
  if (int rc = set_small_lds_attr<double>(lds64)) return rc;
  return set_small_lds_attr<float>(lds32);
}

// ---------------------------------------------------------------------------------------
// Stand-alone level operations on caller-owned CSR levels and vectors: the pieces of the
// reference's Python V-cycle (z02/multigrid_linearelastic.py:25-39, 112-172) for a caller
// that drives the cycle itself.  The context's own V-cycle fuses them (above).
// rr = b - A e (row groups as k_sweep)
template <typename T>
__global__ __launch_bounds__(256) void k_level_rr(int n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const T* __restrict__ val,
                                                  const T* __restrict__ b,
                                                  const T* __restrict__ e,
                                                  T* __restrict__ rr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / ROW_LANES, sub = t % ROW_LANES;
  if (i >= n) return;  // whole groups leave together
  const T bi = b[i];
  const T ax = level_row_group<T>(i, sub, rp, col, val, e, (const T*)nullptr);
  if (sub == 0) rr[i] = bi - ax;
}

// e += omega rr / diag
template <typename T>
__global__ __launch_bounds__(256) void k_level_update(int n, const T* __restrict__ diag,
                                                      const T* __restrict__ rr,
                                                      T* __restrict__ e, T omega) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) e[i] += rr[i] / diag[i] * omega;
}

// coarse[I] = fine[2I] + fine[2I+1], the fine vector padded with one zero
// (F.pad + conv1d [1, 1] stride 2, :147-153)
template <typename T>
__global__ __launch_bounds__(256) void k_level_pairsum(int n, const T* __restrict__ fine,
                                                       T* __restrict__ coarse) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = (n - 1) / 2 + 1;
  if (I >= nc) return;
  T s = fine[2 * I];
  if (2 * I + 1 < n) s += fine[2 * I + 1];
  coarse[I] = s;
}

// fine[i] += coarse[i >> 1] (nearest upsampling, :163-164)
template <typename T>
__global__ __launch_bounds__(256) void k_level_upsample_add(int n, const T* __restrict__ coarse,
                                                            T* __restrict__ fine) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fine[i] += coarse[i >> 1];
}

// P3DG -> P1DG per element, y[e][k] = sum_n P[n][k] x[e][n]
// (vel_pndg_to_p1dg_restrictor, z02/multigrid_linearelastic.py:286-290)
template <typename T, int D>
__global__ __launch_bounds__(256) void k_pn_to_p1dg(int nele, const T* __restrict__ x,
                                                    T* __restrict__ y) {
  constexpr int NL = D == 3 ? 20 : 10, NB = D + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nele * NB) return;
  const int e = t / NB, k = t - (t / NB) * NB;
  const T* xe = x + (size_t)e * NL;
  T s = 0;
#pragma unroll
  for (int n = 0; n < NL; ++n) s += (T)Ptab<D>(n, k) * xe[n];
  y[t] = s;
}

// P1DG -> P3DG per element, y[e][n] = sum_k P[n][k] x[e][k]
// (vel_p1dg_to_pndg_prolongator, :293-297)
template <typename T, int D>
__global__ __launch_bounds__(256) void k_p1dg_to_pn(int nele, const T* __restrict__ x,
                                                    T* __restrict__ y) {
  constexpr int NL = D == 3 ? 20 : 10, NB = D + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nele * NL) return;
  const int e = t / NL, n = t - (t / NL) * NL;
  const T* xe = x + (size_t)e * NB;
  T s = 0;
#pragma unroll
  for (int k = 0; k < NB; ++k) s += (T)Ptab<D>(n, k) * xe[k];
  y[t] = s;
}

template <typename T>
int launch_level_smooth(int n, const int32_t* rp, const int32_t* col, const T* val,
                        const T* diag, T omega, T* e, const T* b, T* rr, hipStream_t s) {
  const int64_t th = (int64_t)n * ROW_LANES;
  hipLaunchKernelGGL((k_level_rr<T>), dim3((unsigned)((th + 255) / 256)), dim3(256), 0, s, n,
                     rp, col, val, b, (const T*)e, rr);
  hipLaunchKernelGGL((k_level_update<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                     diag, (const T*)rr, e, omega);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_level_pairsum(int n, const T* fine, T* coarse, hipStream_t s) {
  const int nc = (n - 1) / 2 + 1;
  hipLaunchKernelGGL((k_level_pairsum<T>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s,
                     n, fine, coarse);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_level_upsample_add(int n, const T* coarse, T* fine, hipStream_t s) {
  hipLaunchKernelGGL((k_level_upsample_add<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     s, n, coarse, fine);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_pn_p1dg(int D, int dir, int nele, const T* x, T* y, hipStream_t s) {
  const int64_t nt = (int64_t)nele * (dir == 0 ? D + 1 : (D == 3 ? 20 : 10));
  const dim3 g((unsigned)((nt + 255) / 256)), b(256);
  if (D == 3) {
    if (dir == 0) hipLaunchKernelGGL((k_pn_to_p1dg<T, 3>), g, b, 0, s, nele, x, y);
    else hipLaunchKernelGGL((k_p1dg_to_pn<T, 3>), g, b, 0, s, nele, x, y);
  } else {
    if (dir == 0) hipLaunchKernelGGL((k_pn_to_p1dg<T, 2>), g, b, 0, s, nele, x, y);
    else hipLaunchKernelGGL((k_p1dg_to_pn<T, 2>), g, b, 0, s, nele, x, y);
  }
  DGMI_HIP(hipGetLastError());
  return 0;
}

template int launch_level_smooth<double>(int, const int32_t*, const int32_t*, const double*,
                                         const double*, double, double*, const double*, double*,
                                         hipStream_t);
template int launch_level_smooth<float>(int, const int32_t*, const int32_t*, const float*,
                                        const float*, float, float*, const float*, float*,
                                        hipStream_t);
template int launch_level_pairsum<double>(int, const double*, double*, hipStream_t);
template int launch_level_pairsum<float>(int, const float*, float*, hipStream_t);
template int launch_level_upsample_add<double>(int, const double*, double*, hipStream_t);
template int launch_level_upsample_add<float>(int, const float*, float*, hipStream_t);
template int launch_pn_p1dg<double>(int, int, int, const double*, double*, hipStream_t);
template int launch_pn_p1dg<float>(int, int, int, const float*, float*, hipStream_t);

template int launch_restrict<double>(dgmi_ctx*, const double*, double*, hipStream_t, double*);
template int launch_restrict<float>(dgmi_ctx*, const float*, float*, hipStream_t, float*);
template bool vcycle_fuses_first<double>(const dgmi_ctx*);
template bool vcycle_fuses_first<float>(const dgmi_ctx*);
template int launch_prolong_add<double>(dgmi_ctx*, const double*, double*, hipStream_t);
template int launch_prolong_add<float>(dgmi_ctx*, const float*, float*, hipStream_t);
template int launch_sfc_vcycle<double>(dgmi_ctx*, const double*, double*, hipStream_t, bool);
template int launch_sfc_vcycle<float>(dgmi_ctx*, const float*, float*, hipStream_t, bool);

}  // namespace dgmi
