// SIPG operator apply, block diagonal and block-Jacobi smoother kernels (gfx950).
//
// Data layout in HBM (element e, nele elements):
//   DOF vectors        x[e*nloc + n]                       (reference layout, element-major)
//   geometry (SoA)     G[c*nele + e], cf[(f*D+d)*nele + e], wf/pen[f*nele + e]
//   face records       nbface[e*nface + f] = N*nface + f' (or -1), fcode = type | alnmt<<4
//   face traces (SoA)  trace[(f*sngi + g)*nele + e] = (u, grad u . n_f) at face point g
//                      (lanes are elements here; the matrix-core kernels, whose lanes
//                      hold points, use the element-major layout instead)
//   smoother (SoA)     S[(n*nloc + m)*nele + e]
//
// The apply is two launches:
//   k_trace_w     u and du/dn at every element's own face points (4 faces x 12 points in
//                 3D) -> trace buffer;
//   k_residual_w  volume term, then for every face the SIPG flux from the own trace and
//                 the neighbour's trace at the aligned points, lifted to the test
//                 functions; writes r = b - A x, or applies the smoother x += S r.
// MODE 0: r = -A x; 1: r = b - A x; 2: x += S (b - A x); 3: as 2 and also store r;
// 4: r = +A x; 5: x = (1 - omega) x + S (b - A_off x), where A_off is A without its
// diagonal blocks; 6: x = S b (a sweep from x = 0).  Mode 5 equals mode 2 when S = omega A_ee^-1 (the exact block
// smoother): it needs no volume term and only the neighbour half of the face terms.
#include <utility>

#include "dd.h"
#include "dgmi.h"
#include "dgmi_internal.h"
#include "faces.h"
#include "rowvec.h"

namespace dgmi {

template <int D> struct Dm;
template <> struct Dm<3> { static constexpr int NL = 20, NGI = 24, SG = 12, NF = 4, NG = 6; };
template <> struct Dm<2> { static constexpr int NL = 10, NGI = 12, SG = 4, NF = 3, NG = 3; };

// ---------------------------------------------------------------------------------------
// Wave-split variants.  A block of nface waves owns 64 elements (one per lane); wave w
// handles face w and the volume points [w*ngi/nface, (w+1)*ngi/nface) of all 64
// elements, so the tables stay wave-uniform (scalar loads) while each element gets nface
// lanes of work.  Partial residuals are summed through LDS.
template <typename T, int D>
__global__ __launch_bounds__(64 * (D + 1)) void k_trace_w(int nele, const T* __restrict__ x,
                                                          const T* __restrict__ fc,
                                                          const int32_t* __restrict__ fnodes,
                                                          const T* __restrict__ cf,
                                                          typename Vec2<T>::type* __restrict__ trace,
                                                          const int32_t* __restrict__ act) {
  using M = Dm<D>;
  using F = FaceDims<D>;
  const int grp = act ? (int)blockIdx.x : xcd_group(blockIdx.x, gridDim.x);
  if (act && !act[grp]) return;  // 64 elements of finished solves
  const int lane = threadIdx.x & 63;
  const int f = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int e = grp * 64 + lane;
  if (e >= nele) return;
  const T* xe = x + (size_t)e * M::NL;
  T xv[M::NL], xf[F::NFN];
  load_row<T, M::NL>(xe, xv);
#pragma unroll
  for (int i = 0; i < F::NFN; ++i) xf[i] = xe[fnodes[f * F::NFN + i]];  // same lines: cached
  T c[D];
#pragma unroll
  for (int d = 0; d < D; ++d) c[d] = cf[(f * D + d) * nele + e];
  face_traces<T, D>(f, fc, xv, xf, c, [&](int g, T u, T dn) {
    typename Vec2<T>::type tv;
    tv.x = u;
    tv.y = dn;
    trace[(f * M::SG + g) * nele + e] = tv;
  });
}

// EMIT (mode 5 only): also write the face traces of the new x into tout (another buffer
// than `trace`, which other work-groups are still reading), so the next sweep needs no
// trace pass.
// SS: storage of the smoother (T, or uint16_t = the bfloat16 packed symmetric copy of the
// bf16 inner cycles).  Arithmetic is in T throughout.
// FP32 neighbour-form sweeps: at least 6 waves per SIMD, so that a 24^3x6 mesh (1,296
// work-groups) is resident in one round on 256 CUs; within that register budget the
// compiler issues the smoother rows' loads ahead of the products.  The single-pass sweep
// needs more registers for the neighbour's row and traces: 4 waves (128 registers, no
// spills; 6 waves spill and measured 10% slower at 48^3x6).
// NBX (single pass): the face traces are not read from a trace buffer but computed here,
// the element's own from its row of x and the neighbour's from the neighbour's row (80 B in
// FP32, read through L2), so no trace pass runs before and no trace buffer is written or
// read.  Modes 0, 1, 4 (x is read only) and 5 (x_new = the output, x read only: the caller
// passes the new rows in r).
template <typename T, int D, int MODE, bool EMIT = false, typename SS = T, bool NBX = false>
__global__ __launch_bounds__(64 * (D + 1),
                             sizeof(T) == 4 && (MODE == 5 || MODE == 6) ? (NBX ? 4 : 6) : 1) void k_residual_w(
    int nele, T* __restrict__ x, const T* b, T* r, const T* __restrict__ vtab,
    const T* __restrict__ wgt, const T* __restrict__ sw,
    const int32_t* __restrict__ aln, const T* __restrict__ G, const T* __restrict__ cf,
    const T* __restrict__ wf, const T* __restrict__ pen, const int32_t* __restrict__ nbface,
    const int32_t* __restrict__ fcode, const typename Vec2<T>::type* __restrict__ trace,
    const SS* __restrict__ S, const int32_t* __restrict__ act, T omega,
    typename Vec2<T>::type* __restrict__ tout, const T* __restrict__ fc,
    const int32_t* __restrict__ fnodes, const T* __restrict__ cnb,
    const uint8_t* __restrict__ nperm) {
  using M = Dm<D>;
  using F = FaceDims<D>;
  static_assert(!EMIT || MODE == 5 || MODE == 6, "traces are emitted by the sweeps");
  static_assert(!NBX || (!EMIT && (MODE <= 1 || MODE == 4 || MODE == 5)),
                "single-pass form: applies and neighbour-form sweeps");
  const int grp = act ? (int)blockIdx.x : xcd_group(blockIdx.x, gridDim.x);
  if (act && !act[grp]) return;  // whole block: no barrier is reached
  constexpr bool OFFD = MODE == 5;
  constexpr bool ZERO = MODE == 6;  // sweep from x = 0: x = S b (x is not read)
  constexpr int NL = M::NL, NF = M::NF;
  constexpr int GPW = M::NGI / NF;          // volume points per wave
  constexpr int NPW = (NL + NF - 1) / NF;   // output rows per wave
  constexpr bool ROWS_FIT = NL % NF == 0;   // every wave's rows exist (3D): no row guards
  static_assert(M::NGI % NF == 0, "volume points must split evenly over the waves");
  // partial residual rows; in the single-pass form wave w first relabels its face's
  // neighbour row through its own rows red[w]
  __shared__ T red[NF][NL][64];
  // node relabellings of the single-pass form (NPERM vertex permutations x NL nodes)
  constexpr int NPERM = D == 3 ? 24 : 6;
  __shared__ uint8_t ptab[NBX ? NPERM * NL : 1];
  if constexpr (NBX) {
    for (int i = threadIdx.x; i < NPERM * NL; i += 64 * NF) ptab[i] = nperm[i];
    __syncthreads();
  }
  // the new rows for the trace output (other waves may still read red when a wave writes)
  __shared__ T xs[EMIT ? NL : 1][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int e = grp * 64 + lane;
  const bool valid = e < nele;
  T xv[NL], acc[NL];
#pragma unroll
  for (int n = 0; n < NL; ++n) acc[n] = xv[n] = 0;
  // this wave's rows of b and of x (kept or damped by the sweeps), requested before the
  // face and volume work; loaded by row index, so no register array is indexed by w
  constexpr bool KEEPX = MODE == 2 || MODE == 3 || OFFD;
  T bv[NPW], xo[NPW];
#pragma unroll
  for (int k = 0; k < NPW; ++k) {
    const int n = w * NPW + k;
    bv[k] = (MODE != 0 && MODE != 4 && !ZERO && valid && n < NL) ? b[(size_t)e * NL + n] : T(0);
    xo[k] = (KEEPX && valid && n < NL) ? x[(size_t)e * NL + n] : T(0);
  }
  // the smoothing modes read the whole row of b: every wave forms the whole residual row
  // from the partial sums, so the rows need no second exchange through LDS
  constexpr bool SMOOTH = MODE == 2 || MODE == 3 || OFFD || ZERO;
  // requested before the face and volume work in FP32 (FP64 loads it after the barrier:
  // the row would cost 40 more registers over the volume work)
  constexpr bool BEARLY = ZERO || (SMOOTH && sizeof(T) == 4);
  T bfull[NL];
#pragma unroll
  for (int n = 0; n < NL; ++n) bfull[n] = 0;
  if constexpr (BEARLY) {
    if (valid) load_row<T, NL>(b + (size_t)e * NL, bfull);
  }
  if (valid && !ZERO && !OFFD) {
    load_row<T, NL>(x + (size_t)e * NL, xv);
  }
  if (valid && !OFFD && !ZERO) {
    // ---- volume points of this wave
    T gm[M::NG];
#pragma unroll
    for (int k = 0; k < M::NG; ++k) gm[k] = G[k * nele + e];
#pragma unroll 2
    for (int gg = 0; gg < GPW; ++gg) {
      const int g = w * GPW + gg;
      const T* t = vtab + g * D * NL;
      T gd[D];
#pragma unroll
      for (int d = 0; d < D; ++d) gd[d] = 0;
#pragma unroll
      for (int n = 0; n < NL; ++n)
#pragma unroll
        for (int d = 0; d < D; ++d) gd[d] += t[d * NL + n] * xv[n];
      const T wq = wgt[g];
      T y[D];
      if constexpr (D == 3) {
        y[0] = wq * (gm[0] * gd[0] + gm[3] * gd[1] + gm[4] * gd[2]);
        y[1] = wq * (gm[3] * gd[0] + gm[1] * gd[1] + gm[5] * gd[2]);
        y[2] = wq * (gm[4] * gd[0] + gm[5] * gd[1] + gm[2] * gd[2]);
      } else {
        y[0] = wq * (gm[0] * gd[0] + gm[2] * gd[1]);
        y[1] = wq * (gm[2] * gd[0] + gm[1] * gd[1]);
      }
#pragma unroll
      for (int n = 0; n < NL; ++n) {
        T s = acc[n];
#pragma unroll
        for (int d = 0; d < D; ++d) s += t[d * NL + n] * y[d];
        acc[n] = s;
      }
    }
  }
  // face w's lift onto its own nodes, added to red[w] after the row sums are stored
  FaceLift<T, D> fl;
  fl.zero();
  if constexpr (NBX) {
    // ---- face w in one pass, without lane-divergent branches: a lane whose face has no
    // term here (or whose element is past the end) gets weight 0; a boundary face reads its
    // own element as the "neighbour" with the neighbour's values multiplied by 0, and the
    // interior and Dirichlet fluxes share one formula (a = 1/2 and the neighbour's trace on
    // an interior face, a = 1 and no neighbour on a Dirichlet face):
    //   q_s = w (-a (du/dn - du'/dn) + pen (u - u')),  q_d = -w a (u - u')
    // and in the neighbour form (mode 5) only the neighbour's part.
    // The neighbour's traces are formed in this face's frame: its vertices relabelled so
    // that the shared face is its face f with this element's vertex order (node n of the
    // relabelled neighbour is its node nperm[pid][n]; cnb = its J^-T n in that frame).
    // Then its face point g is this side's point g, and the trace tables are those of face
    // f for every lane (wave-uniform reads), whatever face the neighbour meets it with.
    const int f = w;
    const int ec = valid ? e : nele - 1;
    const int z = ec * NF + f;
    const int code = fcode[z];
    const int type = code & 15;
    const bool interior = type == 0;
    const bool on = valid && (OFFD ? interior : type < 2);
    T c[D];
#pragma unroll
    for (int d = 0; d < D; ++d) c[d] = cf[(f * D + d) * nele + ec];
    const T wfe = on ? wf[f * nele + ec] : T(0);
    const T pe = pen[f * nele + ec];
    const T half = interior ? T(0.5) : T(1);
    const T keep = interior ? T(1) : T(0);
    const int nbe = interior ? nbface[z] / NF : ec;
    const int pid = interior ? (code >> 8) & 31 : 0;  // 0: the identity
    T tnu[M::SG], tndn[M::SG];
    {
      // the neighbour's row, relabelled through this lane's column of red[w]
      T xn[NL];
      load_row<T, NL>(x + (size_t)nbe * NL, xn);
#pragma unroll
      for (int n = 0; n < NL; ++n) red[w][n][lane] = xn[n];
      const uint8_t* pp = ptab + pid * NL;
      T xr[NL], xrf[F::NFN], cn[D];
#pragma unroll
      for (int n = 0; n < NL; ++n) xr[n] = red[w][pp[n]][lane];
#pragma unroll
      for (int i = 0; i < F::NFN; ++i) xrf[i] = red[w][pp[fnodes[f * F::NFN + i]]][lane];
#pragma unroll
      for (int d = 0; d < D; ++d) cn[d] = cnb[(f * D + d) * nele + ec];
      face_traces<T, D>(f, fc, xr, xrf, cn, [&](int g, T u, T dn) {
        tnu[g] = keep * u;
        tndn[g] = keep * dn;
      });
    }
    T tu[M::SG], tdn[M::SG];
    if constexpr (!OFFD) {
      T xf[F::NFN];
#pragma unroll
      for (int i = 0; i < F::NFN; ++i) xf[i] = x[(size_t)ec * NL + fnodes[f * F::NFN + i]];
      face_traces<T, D>(f, fc, xv, xf, c, [&](int g, T u, T dn) {
        tu[g] = u;
        tdn[g] = dn;
      });
    }
#pragma unroll
    for (int g = 0; g < M::SG; ++g) {
      const T ws = wfe * sw[g];
      T qsn, qdn;
      if constexpr (OFFD) {
        qsn = ws * (T(0.5) * tndn[g] - pe * tnu[g]);
        qdn = ws * (T(0.5) * tnu[g]);
      } else {
        const T jump = tu[g] - tnu[g];
        qsn = ws * (-half * (tdn[g] - tndn[g]) + pe * jump);
        qdn = ws * (-half * jump);
      }
      fl.add(f, g, fc, qsn, qdn);
    }
    fl.finish(f, fc, c, acc);
    // (red[w] is next written by this lane only, after its reads above: program order)
  } else if (valid && !ZERO) {
    // ---- face w (mode 5: interior faces, neighbour terms only)
    const int f = w;
    const int z = e * NF + f;
    const int code = fcode[z];
    const int type = code & 15;
    if (OFFD ? type == 0 : type < 2) {
      T c[D];
#pragma unroll
      for (int d = 0; d < D; ++d) c[d] = cf[(f * D + d) * nele + e];
      const int a = (code >> 4) & 15;
      const T wfe = wf[f * nele + e];
      const T pe = pen[f * nele + e];
      int nbe = 0, nbf = 0;
      if (type == 0) {
        const int nbz = nbface[z];
        nbe = nbz / NF;
        nbf = nbz - nbe * NF;
      }
      {
#pragma unroll 2
      for (int g = 0; g < M::SG; ++g) {
        const T ws = wfe * sw[g];
        T qsn, qdn;
        if constexpr (OFFD) {
          const int ga = aln[a * M::SG + g];
          const auto tn = trace_value(trace[(nbf * M::SG + ga) * nele + nbe]);
          qsn = ws * (T(0.5) * tn.y - pe * tn.x);
          qdn = ws * (T(0.5) * tn.x);
        } else if (type == 0) {
          const auto tr = trace_value(trace[(f * M::SG + g) * nele + e]);
          const int ga = aln[a * M::SG + g];
          const auto tn = trace_value(trace[(nbf * M::SG + ga) * nele + nbe]);
          const T jump = tr.x - tn.x;
          qsn = ws * (T(-0.5) * (tr.y - tn.y) + pe * jump);
          qdn = ws * (T(-0.5) * jump);
        } else {
          const auto tr = trace_value(trace[(f * M::SG + g) * nele + e]);
          qsn = ws * (pe * tr.x - tr.y);
          qdn = -ws * tr.x;
        }
        fl.add(f, g, fc, qsn, qdn);
      }
      }
      fl.finish(f, fc, c, acc);
    }
  }
  T rv[NPW];
  if constexpr (!ZERO) {
#pragma unroll
    for (int n = 0; n < NL; ++n) red[w][n][lane] = acc[n];
    // the face's own nodes (a wave-uniform list): read-modify-write of this lane's column
#pragma unroll
    for (int i = 0; i < F::NFN; ++i) red[w][fnodes[w * F::NFN + i]][lane] += fl.accf[i];
    __syncthreads();
    if constexpr (SMOOTH) {
      if constexpr (!BEARLY) {
        if (valid) load_row<T, NL>(b + (size_t)e * NL, bfull);
      }
      // the whole residual row, in bfull (b - sum of the partials)
#pragma unroll
      for (int n = 0; n < NL; ++n) {
        T s = 0;
#pragma unroll
        for (int ww = 0; ww < NF; ++ww) s += red[ww][n][lane];
        bfull[n] = valid ? bfull[n] - s : T(0);
      }
      if constexpr (MODE == 3) {
        // this wave's rows of r, indexed statically (bfull[w * NPW + k] would not be)
#pragma unroll
        for (int k = 0; k < NPW; ++k) {
          const int n = w * NPW + k;
          if (ROWS_FIT || n < NL) {
            T s = 0;
#pragma unroll
            for (int ww = 0; ww < NF; ++ww) s += red[ww][n][lane];
            rv[k] = valid ? bv[k] - s : T(0);
          }
        }
      }
    } else {
      // rows n of this wave: sum the partials, form the residual
#pragma unroll
      for (int k = 0; k < NPW; ++k) {
        const int n = w * NPW + k;
        if (ROWS_FIT || n < NL) {
          T s = 0;
#pragma unroll
          for (int ww = 0; ww < NF; ++ww) s += red[ww][n][lane];
          if constexpr (MODE == 0) rv[k] = -s;
          else if constexpr (MODE == 4) rv[k] = s;
          else rv[k] = valid ? bv[k] - s : T(0);
        }
      }
    }
  }
  if constexpr (MODE <= 1 || MODE == 4) {
    if (valid) {
#pragma unroll
      for (int k = 0; k < NPW; ++k) {
        const int n = w * NPW + k;
        if (ROWS_FIT || n < NL) r[(size_t)e * NL + n] = rv[k];
      }
    }
  } else {
    // smoothing: bfull holds the whole residual row (b itself from x = 0)
    if (valid) {
#pragma unroll
      for (int k = 0; k < NPW; ++k) {
        const int n = w * NPW + k;
        if (ROWS_FIT || n < NL) {
          T srow[NL];
          load_smoother_row<T, NL, SS>(S, n, e, nele, srow);
          T s = ZERO ? T(0) : (OFFD ? (T(1) - omega) * xo[k] : xo[k]);
#pragma unroll
          for (int m = 0; m < NL; ++m) s += srow[m] * bfull[m];
          if constexpr (NBX) r[(size_t)e * NL + n] = s;  // the new rows go to another vector
          else x[(size_t)e * NL + n] = s;
          if constexpr (MODE == 3) r[(size_t)e * NL + n] = rv[k];
          if constexpr (EMIT) xs[n][lane] = s;
        }
      }
    }
    if constexpr (EMIT) {
      // wave w: the traces of face w of the new x (same operations as k_trace_w)
      __syncthreads();
      if (valid) {
        const int f = w;
        T cc[D], xn[NL], xf[F::NFN];
#pragma unroll
        for (int d = 0; d < D; ++d) cc[d] = cf[(f * D + d) * nele + e];
#pragma unroll
        for (int n = 0; n < NL; ++n) xn[n] = xs[n][lane];
#pragma unroll
        for (int i = 0; i < F::NFN; ++i) xf[i] = xs[fnodes[f * F::NFN + i]][lane];
        face_traces<T, D>(f, fc, xn, xf, cc, [&](int g, T u, T dn) {
          tout[(f * M::SG + g) * nele + e] = make_trace<typename Vec2<T>::type>(u, dn);
        });
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// FP32 smoothing sweep of the exact block smoother in one pass, the form the FP32
// preconditioner runs:  x_new = (1 - omega) x + S (b - A_off x),  S = omega A_ee^-1 packed
// symmetric, A_off = the face coupling to the neighbours.  A block of NF waves owns 64
// elements (one per lane):
//   0. the block's rows of b and x (64 x NL contiguous values each) are read with
//      coalesced vector loads and transposed into LDS, so no wave reads a row of its own
//      from global memory (a lane's row is NL contiguous values, 80 B in 3D);
//   1. wave w forms the neighbour term of face w: the neighbour's row (the only gather),
//      relabelled into the face's frame, its traces, the flux and its lift (as the
//      single-pass k_residual_w);
//   2. every wave forms the element's residual row r = b - sum of the face partials, then
//      its share of S r from a contiguous quarter of the packed entries (each entry read
//      once per block, for both S[a][b] and S[b][a]);
//   3. the partial products are summed per row, damped x added, and the block's new rows
//      written with coalesced vector stores.
// x_new must be another vector than x (the neighbours read x).
template <int NL, int IDX>
__device__ __forceinline__ void sym_term(const float* __restrict__ S, int nele, int e,
                                         const float* r, float* y) {
  constexpr int A = [] {
    int i = IDX, a = 0;
    while (i >= NL - a) i -= NL - a++;
    return a;
  }();
  constexpr int B = [] {
    int i = IDX, a = 0;
    while (i >= NL - a) i -= NL - a++;
    return a + i;
  }();
  static_assert(sym_index<NL>(A, B) == IDX, "packed index");
  const float s = S[(size_t)IDX * nele + e];
  y[A] += s * r[B];
  if constexpr (A != B) y[B] += s * r[A];
}
template <int NL, int IDX>
__device__ __forceinline__ void sym_term_s(float s, const float* r, float* y) {
  constexpr int A = [] {
    int i = IDX, a = 0;
    while (i >= NL - a) i -= NL - a++;
    return a;
  }();
  constexpr int B = [] {
    int i = IDX, a = 0;
    while (i >= NL - a) i -= NL - a++;
    return a + i;
  }();
  y[A] += s * r[B];
  if constexpr (A != B) y[B] += s * r[A];
}
template <int NL, int P0, int... J>
__device__ __forceinline__ void sym_part(const float* __restrict__ S, int nele, int e,
                                         const float* r, float* y,
                                         std::integer_sequence<int, J...>) {
  (sym_term<NL, P0 + J>(S, nele, e, r, y), ...);
}
// wave W of NF: packed entries [W NS / NF, (W+1) NS / NF)
template <int NL, int NF, int W>
__device__ __forceinline__ void sym_share(const float* __restrict__ S, int nele, int e,
                                          const float* r, float* y) {
  constexpr int NS = sym_size<NL>();
  constexpr int P0 = W * NS / NF, P1 = (W + 1) * NS / NF;
  sym_part<NL, P0>(S, nele, e, r, y, std::make_integer_sequence<int, P1 - P0>{});
}

// The same share from the FP16 copy: packed entries in pairs (2p, 2p+1) per 32-bit word,
// wave W of NF taking pairs [W NP / NF, (W+1) NP / NF); the caller multiplies y by the
// element's scale
__device__ __forceinline__ float half_lo(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}
template <int NL, int PAIR>
__device__ __forceinline__ void sym_pair_h(const uint32_t* __restrict__ H, int nele, int e,
                                           const float* r, float* y) {
  const uint32_t wv = H[(size_t)PAIR * nele + e];
  sym_term_s<NL, 2 * PAIR>(half_lo(wv), r, y);
  if constexpr (2 * PAIR + 1 < sym_size<NL>()) sym_term_s<NL, 2 * PAIR + 1>(half_hi(wv), r, y);
}
template <int NL, int P0, int... J>
__device__ __forceinline__ void sym_part_h(const uint32_t* __restrict__ H, int nele, int e,
                                           const float* r, float* y,
                                           std::integer_sequence<int, J...>) {
  (sym_pair_h<NL, P0 + J>(H, nele, e, r, y), ...);
}
template <int NL, int NF, int W>
__device__ __forceinline__ void sym_share_h(const uint32_t* __restrict__ H, int nele, int e,
                                            const float* r, float* y) {
  constexpr int NP = (sym_size<NL>() + 1) / 2;
  constexpr int P0 = W * NP / NF, P1 = (W + 1) * NP / NF;
  sym_part_h<NL, P0>(H, nele, e, r, y, std::make_integer_sequence<int, P1 - P0>{});
}

// The same from the wave's words already in registers (hw[j] = pair P0 + j)
template <int NL, int PAIR>
__device__ __forceinline__ void sym_pair_hw(uint32_t wv, const float* r, float* y) {
  sym_term_s<NL, 2 * PAIR>(half_lo(wv), r, y);
  if constexpr (2 * PAIR + 1 < sym_size<NL>()) sym_term_s<NL, 2 * PAIR + 1>(half_hi(wv), r, y);
}
template <int NL, int P0, int... J>
__device__ __forceinline__ void sym_part_hw(const uint32_t* hw, const float* r, float* y,
                                            std::integer_sequence<int, J...>) {
  (sym_pair_hw<NL, P0 + J>(hw[J], r, y), ...);
}
template <int NL, int NF, int W>
__device__ __forceinline__ void sym_share_hw(const uint32_t* hw, const float* r, float* y) {
  constexpr int NP = (sym_size<NL>() + 1) / 2;
  constexpr int P0 = W * NP / NF, P1 = (W + 1) * NP / NF;
  sym_part_hw<NL, P0>(hw, r, y, std::make_integer_sequence<int, P1 - P0>{});
}

template <int D, int MODE, bool HALF = false>
__global__ __launch_bounds__(64 * (D + 1), MODE == 5 || MODE == 6 ? 5 : 4) void k_staged(
    int nele, const float* __restrict__ x, const float* __restrict__ b, float* __restrict__ out,
    const float* __restrict__ vtab, const float* __restrict__ wgt, const float* __restrict__ G,
    const float* __restrict__ sw, const float* __restrict__ cf, const float* __restrict__ wf,
    const float* __restrict__ pen, const int32_t* __restrict__ nbface,
    const int32_t* __restrict__ fcode, const float* __restrict__ S,
    const int32_t* __restrict__ act, float omega, const float* __restrict__ fc,
    const int32_t* __restrict__ fnodes, const float* __restrict__ cnb,
    const uint8_t* __restrict__ nperm, const uint32_t* __restrict__ Sh = nullptr,
    const float* __restrict__ Shs = nullptr) {
  using M = Dm<D>;
  using F = FaceDims<D>;
  using RV = RowVec<float, M::NL>;
  using V = typename RV::V;
  static_assert(MODE == 0 || MODE == 1 || MODE == 4 || MODE == 5 || MODE == 6, "mode");
  constexpr bool SWEEP = MODE == 5 || MODE == 6;  // x_new = (1 - omega) x + S r
  constexpr bool ZERO = MODE == 6;                 // from x = 0: x_new = S b
  constexpr bool APPLY = !SWEEP;                   // the whole operator (volume and faces)
  constexpr bool READB = MODE == 1 || SWEEP;
  constexpr int NL = M::NL, NF = M::NF, W = RV::W;
  constexpr int GPW = M::NGI / NF;          // volume points per wave
  constexpr int NPW = (NL + NF - 1) / NF;   // output rows per wave
  constexpr int NPERM = D == 3 ? 24 : 6;    // vertex permutations
  constexpr int NT = 64 * NF;               // threads
  constexpr int NVEC = 64 * NL / W;         // vectors of a block's rows
  static_assert(M::NGI % NF == 0, "volume points must split evenly over the waves");
  const int grp = act ? (int)blockIdx.x : xcd_group(blockIdx.x, gridDim.x);
  if (act && !act[grp]) return;  // whole block: no barrier is reached
  __shared__ float red[NF][NL][64];       // face relabelling, partial rows, partial S r
  __shared__ float bs[READB ? NL : 1][65];  // the block's rows of b, element per column
  __shared__ float xs[NL][65];            // rows of x, then of the output
  __shared__ uint8_t ptab[ZERO ? 1 : NPERM * NL];
  // (applies) the P2 mass matrix times the P2 coefficients of d u / d xi_d, per direction
  __shared__ float vtmp[APPLY ? D : 1][APPLY ? F::NP2 : 1][64];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int e0 = grp * 64;
  const int e = e0 + lane;
  const bool valid = e < nele;
  const int ec = valid ? e : nele - 1;
  const int nvec = (nele - e0 < 64 ? nele - e0 : 64) * NL / W;  // vectors of real elements
  // ---- 0. rows of b and x into LDS (zeros past the last element)
  {
    const V* bv = reinterpret_cast<const V*>(b + (size_t)e0 * NL);
    const V* xv = reinterpret_cast<const V*>(x + (size_t)e0 * NL);
#pragma unroll
    for (int i0 = 0; i0 < NVEC; i0 += NT) {
      const int i = i0 + tid;
      if (i < NVEC) {
        float tb[W], tx[W];
#pragma unroll
        for (int k = 0; k < W; ++k) tb[k] = tx[k] = 0.0f;
        if (i < nvec) {
          if constexpr (READB) {
            const V v = bv[i];
            const float* pv = reinterpret_cast<const float*>(&v);
#pragma unroll
            for (int k = 0; k < W; ++k) tb[k] = pv[k];
          }
          if constexpr (!ZERO) {
            const V v = xv[i];
            const float* pv = reinterpret_cast<const float*>(&v);
#pragma unroll
            for (int k = 0; k < W; ++k) tx[k] = pv[k];
          }
        }
        const int el = (W * i) / NL, n = (W * i) % NL;
#pragma unroll
        for (int k = 0; k < W; ++k) {
          if constexpr (READB) bs[n + k][el] = tb[k];
          xs[n + k][el] = tx[k];
        }
      }
    }
    if constexpr (!ZERO)
      for (int i = tid; i < NPERM * NL; i += NT) ptab[i] = nperm[i];
  }
  __syncthreads();
  // (FP16 smoother) this wave's words of the packed smoother, loaded before the face work
  // so that their latency hides behind it
  constexpr int NPH = (sym_size<NL>() + 1) / 2, HW = (NPH + NF - 1) / NF;
  uint32_t hw[HALF ? HW : 1];
  if constexpr (HALF) {
    const int p0 = w * NPH / NF, p1 = (w + 1) * NPH / NF;
#pragma unroll
    for (int j = 0; j < HW; ++j) hw[j] = p0 + j < p1 ? Sh[(size_t)(p0 + j) * nele + ec] : 0u;
  }
  if constexpr (!ZERO) {
    float acc[NL];
#pragma unroll
    for (int n = 0; n < NL; ++n) acc[n] = 0.0f;
    // ---- 1a. (applies) volume term, first half: d u / d xi_w is a P2 polynomial; its
    // P2 coefficients g_w = VD_w x, times the P2 mass matrix, go to the other waves
    // (wave w < D takes direction w; the volume integral is
    //   r_n = sum_{d,e} G_de sum_ij VD_d[i][n] VM[i][j] VD_e[j][m] x_m,
    // exact for the affine element, the rule being exact for P2 x P2)
    if constexpr (APPLY) {
      if (w < D) {
        const float* vd = fc + F::VD + w * F::NP2 * NL;
        const float* vm = fc + F::VM;
        float xv[NL], gco[F::NP2];
#pragma unroll
        for (int n = 0; n < NL; ++n) xv[n] = xs[n][lane];
#pragma unroll
        for (int i = 0; i < F::NP2; ++i) {
          float s = 0.0f;
#pragma unroll
          for (int n = 0; n < NL; ++n) s += vd[i * NL + n] * xv[n];
          gco[i] = s;
        }
#pragma unroll
        for (int i = 0; i < F::NP2; ++i) {
          float s = 0.0f;
#pragma unroll
          for (int j = 0; j < F::NP2; ++j) s += vm[i * F::NP2 + j] * gco[j];
          vtmp[w][i][lane] = s;
        }
      }
    }
    // ---- 1b. face w, without lane-divergent branches: a face without a term here (or a
    // lane past the last element) gets weight 0, and a boundary face reads its own element
    // as the "neighbour" with the neighbour's values multiplied by 0.  The neighbour's
    // traces are formed in this face's frame: its row relabelled so that the shared face
    // is its face f with this element's vertex order (node n of the relabelled neighbour
    // is its node nperm[pid][n]; cnb = its J^-T n in that frame), so its face point g is
    // this side's point g and every trace table read is wave-uniform.  Interior and
    // Dirichlet fluxes share one formula (a = 1/2 and the neighbour's traces on an
    // interior face, a = 1 and no neighbour on a Dirichlet face):
    //   q_s = w (-a (du/dn - du'/dn) + pen (u - u')),  q_d = -w a (u - u'),
    // and the sweep (A without its diagonal blocks) keeps the neighbour's part only:
    //   q_s = w (a du'/dn - pen u'),  q_d = w a u'.
    FaceLift<float, D> fl;
    fl.zero();
    const int f = w;
    {
      const int z = ec * NF + f;
      const int code = fcode[z];
      const int type = code & 15;
      const bool interior = type == 0;
      const bool on = valid && (SWEEP ? interior : type < 2);
      float c[D], cn[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        c[d] = cf[(f * D + d) * nele + ec];
        cn[d] = cnb[(f * D + d) * nele + ec];
      }
      const float wfe = on ? wf[f * nele + ec] : 0.0f;
      const float pe = pen[f * nele + ec];
      const float half = interior ? 0.5f : 1.0f;
      const float keep = interior ? 1.0f : 0.0f;
      const int nbe = interior ? nbface[z] / NF : ec;
      const int pid = interior ? (code >> 8) & 31 : 0;  // 0: the identity
      // the neighbour's row, relabelled through this lane's column of red[w]: its
      // face-node values and c' . grad(u') at the face's quadratic nodes
      float xrf[F::NFN], dkn[F::NK];
      {
        float xn[NL];
        load_row<float, NL>(x + (size_t)nbe * NL, xn);
#pragma unroll
        for (int n = 0; n < NL; ++n) red[w][n][lane] = xn[n];
        const uint8_t* pp = ptab + pid * NL;
        float xr[NL];
#pragma unroll
        for (int n = 0; n < NL; ++n) xr[n] = red[w][pp[n]][lane];
#pragma unroll
        for (int i = 0; i < F::NFN; ++i) xrf[i] = keep * red[w][pp[fnodes[f * F::NFN + i]]][lane];
        face_dk<float, D>(f, fc, xr, cn, dkn);
#pragma unroll
        for (int k2 = 0; k2 < F::NK; ++k2) dkn[k2] *= keep;
      }
      if constexpr (APPLY) {
        // own traces in node form: the jump u - u' at the face's nodes and
        // du/dn - du'/dn at its quadratic nodes
        float xv[NL], jump[F::NFN], ddk[F::NK];
#pragma unroll
        for (int n = 0; n < NL; ++n) xv[n] = xs[n][lane];
        face_dk<float, D>(f, fc, xv, c, ddk);
#pragma unroll
        for (int k2 = 0; k2 < F::NK; ++k2) ddk[k2] -= dkn[k2];
#pragma unroll
        for (int i = 0; i < F::NFN; ++i) jump[i] = xs[fnodes[f * F::NFN + i]][lane] - xrf[i];
        fl.set_nodes(f, fc, jump, ddk, wfe * pe, -half * wfe, -half * wfe);
      } else {
        fl.set_nodes(f, fc, xrf, dkn, -pe * wfe, 0.5f * wfe, 0.5f * wfe);
      }
      {
        // the faces' lifts are summed at the P2 nodes and applied once (the applies with
        // the volume term in step 1c, the sweeps in their own step 1c): this face's Q and
        // c go to the slots of red[w] at the nodes off face w, which the rows never read
        using O = FaceOff<D>;
        static_assert(O::NO >= F::NK + D, "room for Q and c off the face");
#pragma unroll
        for (int k2 = 0; k2 < F::NK; ++k2) red[w][O::tab.off[w][k2]][lane] = fl.Q[k2];
#pragma unroll
        for (int d = 0; d < D; ++d) red[w][O::tab.off[w][F::NK + d]][lane] = c[d];
      }
    }
    if constexpr (APPLY) {
      // ---- 1c. volume term, second half, with the faces' lifts in direction w:
      //   h_w = sum_e G_we (VM g_e) + W[.][w],  W[p][w] = sum_{f, k: node k of face f is p}
      //   c_fw Q_fk,  r += VD_w^T h_w
      __syncthreads();  // vtmp of every direction, every face's Q and c
      if (w < D) {
        using O = FaceOff<D>;
        const float* vd = fc + F::VD + w * F::NP2 * NL;
        float gr[D];
#pragma unroll
        for (int d = 0; d < D; ++d) gr[d] = G[(w == d ? w : D - 1 + w + d) * nele + ec];
        float hw[F::NP2];
#pragma unroll
        for (int i = 0; i < F::NP2; ++i) hw[i] = 0.0f;
#pragma unroll
        for (int ff = 0; ff < NF; ++ff) {
          const float cw = red[ff][O::tab.off[ff][F::NK + w]][lane];
#pragma unroll
          for (int k2 = 0; k2 < F::NK; ++k2)
            hw[FaceP2<D>::p[ff][k2]] += cw * red[ff][O::tab.off[ff][k2]][lane];
        }
#pragma unroll
        for (int i = 0; i < F::NP2; ++i) {
          float h = hw[i];
#pragma unroll
          for (int d = 0; d < D; ++d) h += gr[d] * vtmp[d][i][lane];
#pragma unroll
          for (int n = 0; n < NL; ++n) acc[n] += vd[i * NL + n] * h;
        }
      }
      __syncthreads();  // every wave has read the parked Q and c before red[w] is rewritten
    }
    // (red[w] is next written by this lane only, after its reads above: program order)
    if constexpr (APPLY) {
#pragma unroll
      for (int n = 0; n < NL; ++n) red[w][n][lane] = acc[n];
#pragma unroll
      for (int i = 0; i < F::NFN; ++i) red[w][fnodes[w * F::NFN + i]][lane] += fl.accf[i];
    } else {
#pragma unroll
      for (int i = 0; i < F::NFN; ++i) red[w][fnodes[w * F::NFN + i]][lane] = fl.accf[i];
    }
    __syncthreads();
    if constexpr (SWEEP) {
      // ---- 1c. (sweeps) the faces' lifts summed at the P2 nodes,
      //   W[p][d] = sum_{f, k: node k of face f is p} c_fd Q_fk,
      // then for this wave's rows r_n = b_n - sum_{faces f on n} accf_f(n)
      //   - sum_{p,d} W[p][d] dphi_n/dxi_d(p)  (one lift instead of one per face)
      using O = FaceOff<D>;
      float Wp[F::NP2][D];
#pragma unroll
      for (int i = 0; i < F::NP2; ++i)
#pragma unroll
        for (int d = 0; d < D; ++d) Wp[i][d] = 0.0f;
#pragma unroll
      for (int ff = 0; ff < NF; ++ff) {
        float cc[D];
#pragma unroll
        for (int d = 0; d < D; ++d) cc[d] = red[ff][O::tab.off[ff][F::NK + d]][lane];
#pragma unroll
        for (int k2 = 0; k2 < F::NK; ++k2) {
          const float q = red[ff][O::tab.off[ff][k2]][lane];
#pragma unroll
          for (int d = 0; d < D; ++d) Wp[FaceP2<D>::p[ff][k2]][d] += cc[d] * q;
        }
      }
      const float* vd = fc + F::VD;
#pragma unroll
      for (int k = 0; k < NPW; ++k) {
        const int n = w * NPW + k;
        if (NL % NF == 0 || n < NL) {
          float s = bs[n][lane];
          const int mk = O::tab.mask[n];
#pragma unroll
          for (int ff = 0; ff < NF; ++ff)
            if (mk & (1 << ff)) s -= red[ff][n][lane];
#pragma unroll
          for (int i = 0; i < F::NP2; ++i)
#pragma unroll
            for (int d = 0; d < D; ++d) s -= Wp[i][d] * vd[(d * F::NP2 + i) * NL + n];
          bs[n][lane] = s;  // this wave's rows only
        }
      }
      __syncthreads();
    }
  }
  if constexpr (APPLY) {
    // ---- 2. rows of this wave: r = -A x, b - A x or A x
#pragma unroll
    for (int k = 0; k < NPW; ++k) {
      const int n = w * NPW + k;
      if (NL % NF == 0 || n < NL) {
        float s = 0.0f;
#pragma unroll
        for (int ww = 0; ww < NF; ++ww) s += red[ww][n][lane];
        if constexpr (MODE == 0) s = -s;
        else if constexpr (MODE == 1) s = bs[n][lane] - s;
        xs[n][lane] = s;
      }
    }
  } else {
    // ---- 2. the residual row (b itself from x = 0), then this wave's share of S r
    float r[NL], y[NL];
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      const float s = bs[n][lane];  // b, or b - A_off x from step 1c
      r[n] = s;
      y[n] = 0.0f;
    }
    if constexpr (HALF) {
      if (w == 0) sym_share_hw<NL, NF, 0>(hw, r, y);
      else if (w == 1) sym_share_hw<NL, NF, 1>(hw, r, y);
      else if (w == 2) sym_share_hw<NL, NF, 2>(hw, r, y);
      else if constexpr (NF > 3) sym_share_hw<NL, NF, 3>(hw, r, y);
      const float sc = Shs[ec];
#pragma unroll
      for (int n = 0; n < NL; ++n) y[n] *= sc;
    } else {
      if (w == 0) sym_share<NL, NF, 0>(S, nele, ec, r, y);
      else if (w == 1) sym_share<NL, NF, 1>(S, nele, ec, r, y);
      else if (w == 2) sym_share<NL, NF, 2>(S, nele, ec, r, y);
      else if constexpr (NF > 3) sym_share<NL, NF, 3>(S, nele, ec, r, y);
    }
    if constexpr (!ZERO) __syncthreads();  // every wave has read the face partials
#pragma unroll
    for (int n = 0; n < NL; ++n) red[w][n][lane] = y[n];
    __syncthreads();
    // ---- 3. rows of this wave: x_new = (1 - omega) x + S r
#pragma unroll
    for (int k = 0; k < NPW; ++k) {
      const int n = w * NPW + k;
      if (NL % NF == 0 || n < NL) {
        float s = ZERO ? 0.0f : (1.0f - omega) * xs[n][lane];
#pragma unroll
        for (int ww = 0; ww < NF; ++ww) s += red[ww][n][lane];
        xs[n][lane] = s;
      }
    }
  }
  __syncthreads();
  // ---- 4. the block's output rows, coalesced
  {
    V* ov = reinterpret_cast<V*>(out + (size_t)e0 * NL);
#pragma unroll
    for (int i0 = 0; i0 < NVEC; i0 += NT) {
      const int i = i0 + tid;
      if (i < nvec) {
        const int el = (W * i) / NL, n = (W * i) % NL;
        V v;
        float* pv = reinterpret_cast<float*>(&v);
#pragma unroll
        for (int k = 0; k < W; ++k) pv[k] = xs[n + k][el];
        ov[i] = v;
      }
    }
  }
}

// r = b - A x with every operation in double-double arithmetic (dd.h) and the result
// rounded once to FP64.  The operator is the FP64 one of k_residual_w (same geometry,
// tables and face terms; its coefficients are taken as exact), but the sum is carried to
// about 106 bits: in FP64 the volume and face contributions of a smooth x cancel from
// |A||x| down to |b - A x|, and the rounding of those terms alone can reach 1e-10 ||b||
// for strongly varying diffusivity on fine meshes.  One thread per element computes its
// own and its neighbours' face traces directly from x (no trace buffer); the solvers call
// it only for the few true residuals that decide convergence.
// The face terms are formed without lane-divergent branches (with this kernel's register
// pressure a branching version returned non-finite values on faces that part of a wave
// skipped): a face of another type than
// interior (0) or Dirichlet (1) gets weight 0, a boundary face reads its own element as
// the "neighbour" with that neighbour's values multiplied out, and the interior and
// Dirichlet fluxes share one formula,
//   q_s = w (-a (du/dn - du'/dn) + pen (u - u')),  q_d = -w a (u - u'),
// with a = 1/2 and the neighbour's u', du'/dn (outward normal of the neighbour) on an
// interior face, a = 1 and u' = du'/dn = 0 on a Dirichlet face.
template <int D>
__global__ __launch_bounds__(64) void k_residual_dd(
    int nele, const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r,
    const double* __restrict__ vtab, const double* __restrict__ wgt,
    const double* __restrict__ sw, const int32_t* __restrict__ aln,
    const double* __restrict__ G, const double* __restrict__ cf, const double* __restrict__ wf,
    const double* __restrict__ pen, const int32_t* __restrict__ nbface,
    const int32_t* __restrict__ fcode, const double* __restrict__ fc,
    const int32_t* __restrict__ fnodes) {
  using M = Dm<D>;
  using F = FaceDims<D>;
  constexpr int NL = M::NL, NF = M::NF, SG = M::SG;
  const int e = min((int)(blockIdx.x * blockDim.x + threadIdx.x), nele - 1);
  const bool store = (int)(blockIdx.x * blockDim.x + threadIdx.x) < nele;
  dd xv[NL], acc[NL];
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    xv[n] = dd(x[(size_t)e * NL + n]);
    acc[n] = dd(0.0);
  }
  // ---- volume: acc_n = sum_g w_g t_g(n) . (G t_g(x))
  double gm[M::NG];
#pragma unroll
  for (int k = 0; k < M::NG; ++k) gm[k] = G[k * nele + e];
#pragma unroll 1
  for (int g = 0; g < M::NGI; ++g) {
    const double* t = vtab + g * D * NL;
    dd gd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) gd[d] = dd(0.0);
#pragma unroll
    for (int n = 0; n < NL; ++n)
#pragma unroll
      for (int d = 0; d < D; ++d) gd[d] += t[d * NL + n] * xv[n];
    const double wq = wgt[g];
    dd y[D];
    if constexpr (D == 3) {
      y[0] = wq * (gm[0] * gd[0] + gm[3] * gd[1] + gm[4] * gd[2]);
      y[1] = wq * (gm[3] * gd[0] + gm[1] * gd[1] + gm[5] * gd[2]);
      y[2] = wq * (gm[4] * gd[0] + gm[5] * gd[1] + gm[2] * gd[2]);
    } else {
      y[0] = wq * (gm[0] * gd[0] + gm[2] * gd[1]);
      y[1] = wq * (gm[2] * gd[0] + gm[1] * gd[1]);
    }
#pragma unroll
    for (int n = 0; n < NL; ++n)
#pragma unroll
      for (int d = 0; d < D; ++d) acc[n] += t[d * NL + n] * y[d];
  }
  // ---- faces
#pragma unroll 1
  for (int f = 0; f < NF; ++f) {
    const int z = e * NF + f;
    const int code = fcode[z];
    const int type = code & 15;
    const bool interior = type == 0;
    const double wfe = (type < 2) ? wf[f * nele + e] : 0.0;
    const double pe = pen[f * nele + e];
    const double half = interior ? 0.5 : 1.0;
    const double keep = interior ? 1.0 : 0.0;  // multiplies the neighbour's values
    const int nbz = interior ? nbface[z] : z;   // a boundary face reads itself
    const int nbe = nbz / NF, nbf = nbz - (nbz / NF) * NF;
    const int a = interior ? ((code >> 4) & 15) : 0;
    dd c[D], xf[F::NFN], tu[SG], tdn[SG], nu[SG], ndn[SG];
#pragma unroll
    for (int d = 0; d < D; ++d) c[d] = dd(cf[(f * D + d) * nele + e]);
#pragma unroll
    for (int i = 0; i < F::NFN; ++i) xf[i] = dd(x[(size_t)e * NL + fnodes[f * F::NFN + i]]);
    face_traces<dd, D>(f, fc, xv, xf, c, [&](int g, dd u, dd dn) {
      tu[g] = u;
      tdn[g] = dn;
    });
    {
      dd xn[NL], xnf[F::NFN], cn[D];
#pragma unroll
      for (int n = 0; n < NL; ++n) xn[n] = dd(x[(size_t)nbe * NL + n]);
#pragma unroll
      for (int i = 0; i < F::NFN; ++i)
        xnf[i] = dd(x[(size_t)nbe * NL + fnodes[nbf * F::NFN + i]]);
#pragma unroll
      for (int d = 0; d < D; ++d) cn[d] = dd(cf[(nbf * D + d) * nele + nbe]);
      face_traces<dd, D>(nbf, fc, xn, xnf, cn, [&](int g, dd u, dd dn) {
        nu[g] = keep * u;
        ndn[g] = keep * dn;
      });
    }
    FaceLift<dd, D> fl;
    fl.zero();
#pragma unroll
    for (int g = 0; g < SG; ++g) {
      const dd ws = wfe * dd(sw[g]);
      // the neighbour's point aligned with g, selected without indexing a register array
      const int ga = aln[a * SG + g];
      dd tnu(0.0), tndn(0.0);
#pragma unroll
      for (int g2 = 0; g2 < SG; ++g2) {
        tnu = g2 == ga ? nu[g2] : tnu;
        tndn = g2 == ga ? ndn[g2] : tndn;
      }
      const dd jump = tu[g] - tnu;
      const dd qsn = ws * (-half * (tdn[g] - tndn) + pe * jump);
      const dd qdn = ws * (-half * jump);
      fl.add(f, g, fc, qsn, qdn);
    }
    fl.finish(f, fc, c, acc);
#pragma unroll
    for (int i = 0; i < F::NFN; ++i) {
      const int node = fnodes[f * F::NFN + i];  // wave-uniform
#pragma unroll
      for (int n = 0; n < NL; ++n)
        if (n == node) acc[n] += fl.accf[i];
    }
  }
  if (store) {
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      const dd s = dd(b[(size_t)e * NL + n]) - acc[n];
      r[(size_t)e * NL + n] = s.hi;
    }
  }
}

// The same residual with the face terms in node form and the neighbour in the face's
// frame (as k_staged), which the single-pass kernels need every interior face to have:
// per face, c . grad(u) at the quadratic nodes on both sides, the jump of u at the face's
// nodes, then accf = w (pen MF jump - a ML ddk) onto the face's nodes and
// Q = -a w ML^T jump lifted through the gradient table.  Far fewer double-double values
// live at once than in the per-point form, so the kernel holds them in registers.
template <int D>
__device__ __forceinline__ void dd_face_dk(int f, const double* __restrict__ fc,
                                           const double* xv, const dd* c, dd* dk) {
  using F = FaceDims<D>;
  const double* gk = fc + F::GK + f * F::NK * D * F::NL;
#pragma unroll
  for (int k = 0; k < F::NK; ++k) {
    dd gd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) gd[d] = dd(0.0);
#pragma unroll
    for (int n = 0; n < F::NL; ++n)
#pragma unroll
      for (int d = 0; d < D; ++d) gd[d] += two_prod(gk[(k * D + d) * F::NL + n], xv[n]);
    dd s(0.0);
#pragma unroll
    for (int d = 0; d < D; ++d) s += c[d] * gd[d];
    dk[k] = s;
  }
}

template <int D>
__global__ __launch_bounds__(128, 2) void k_residual_dd_nodes(
    int nele, const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r,
    const double* __restrict__ vtab, const double* __restrict__ wgt,
    const double* __restrict__ G, const double* __restrict__ cf, const double* __restrict__ cnb,
    const double* __restrict__ wf, const double* __restrict__ pen,
    const int32_t* __restrict__ nbface, const int32_t* __restrict__ fcode,
    const double* __restrict__ fc, const uint8_t* __restrict__ nperm) {
  using M = Dm<D>;
  using F = FaceDims<D>;
  using FN = FaceNodes<D>;
  constexpr int NL = M::NL, NF = M::NF, NFN = F::NFN, NK = F::NK;
  constexpr int NPERM = D == 3 ? 24 : 6;
  __shared__ uint8_t ptab[NPERM * NL];
  // the rows accumulate in LDS (each thread its own column), which keeps the row of
  // double-double sums out of the registers the volume and face work need
  __shared__ double ahi[NL][128], alo[NL][128];
  for (int i = threadIdx.x; i < NPERM * NL; i += blockDim.x) ptab[i] = nperm[i];
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = min(t, nele - 1);
  const bool store = t < nele;
  double xv[NL];
  load_row<double, NL>(x + (size_t)e * NL, xv);
  const int tl = threadIdx.x;
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    ahi[n][tl] = 0.0;
    alo[n][tl] = 0.0;
  }
  auto add_row = [&](int n, dd v) {
    const dd a = dd(ahi[n][tl], alo[n][tl]) + v;
    ahi[n][tl] = a.hi;
    alo[n][tl] = a.lo;
  };
  // ---- volume: acc_n = sum_g w_g t_g(n) . (G t_g(x))
  double gm[M::NG];
#pragma unroll
  for (int k = 0; k < M::NG; ++k) gm[k] = G[k * nele + e];
#pragma unroll 1
  for (int g = 0; g < M::NGI; ++g) {
    const double* tt = vtab + g * D * NL;
    dd gd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) gd[d] = dd(0.0);
#pragma unroll
    for (int n = 0; n < NL; ++n)
#pragma unroll
      for (int d = 0; d < D; ++d) gd[d] += two_prod(tt[d * NL + n], xv[n]);
    const double wq = wgt[g];
    dd y[D];
    if constexpr (D == 3) {
      y[0] = wq * (gm[0] * gd[0] + gm[3] * gd[1] + gm[4] * gd[2]);
      y[1] = wq * (gm[3] * gd[0] + gm[1] * gd[1] + gm[5] * gd[2]);
      y[2] = wq * (gm[4] * gd[0] + gm[5] * gd[1] + gm[2] * gd[2]);
    } else {
      y[0] = wq * (gm[0] * gd[0] + gm[2] * gd[1]);
      y[1] = wq * (gm[2] * gd[0] + gm[1] * gd[1]);
    }
#pragma unroll 4
    for (int n = 0; n < NL; ++n) {
      dd v = tt[n] * y[0];
#pragma unroll
      for (int d = 1; d < D; ++d) v += tt[d * NL + n] * y[d];
      add_row(n, v);
    }
  }
  // ---- faces (unrolled: the face-node positions are compile-time)
  auto face = [&](auto fi) {
    constexpr int f = decltype(fi)::value;
    const int z = e * NF + f;
    const int code = fcode[z];
    const int type = code & 15;
    const bool interior = type == 0;
    const double wfe = (type < 2) ? wf[f * nele + e] : 0.0;
    const double pe = pen[f * nele + e];
    const double half = interior ? 0.5 : 1.0;
    const double keep = interior ? 1.0 : 0.0;  // multiplies the neighbour's values
    const int nbe = interior ? nbface[z] / NF : e;
    const int pid = interior ? (code >> 8) & 31 : 0;
    dd c[D], cn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      c[d] = dd(cf[(f * D + d) * nele + e]);
      cn[d] = dd(cnb[(f * D + d) * nele + e]);
    }
    // the neighbour relabelled into this face's frame: c' . grad(u') at the quadratic
    // nodes and u' at the face's nodes
    dd ddk[NK];
    double xrf[NFN];
    {
      double xr[NL];
      const uint8_t* pp = ptab + pid * NL;
#pragma unroll
      for (int n = 0; n < NL; ++n) xr[n] = x[(size_t)nbe * NL + pp[n]];
#pragma unroll
      for (int i = 0; i < NFN; ++i) xrf[i] = keep * xr[FN::v[f][i]];
      dd_face_dk<D>(f, fc, xr, cn, ddk);
    }
    {
      dd dk[NK];
      dd_face_dk<D>(f, fc, xv, c, dk);
#pragma unroll
      for (int k = 0; k < NK; ++k) ddk[k] = dk[k] - keep * ddk[k];
    }
    dd jump[NFN];
#pragma unroll
    for (int i = 0; i < NFN; ++i) jump[i] = two_sum(xv[FN::v[f][i]], -xrf[i]);
    const double* mf = fc + F::MF + f * NFN * NFN;
    const double* ml = fc + F::ML + f * NFN * NK;
    const double sv = wfe * pe, sd = -half * wfe;  // (products rounded once: coefficients)
    dd Q[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) Q[k] = dd(0.0);
#pragma unroll
    for (int i = 0; i < NFN; ++i) {
      dd a(0.0), bq(0.0);
#pragma unroll
      for (int j = 0; j < NFN; ++j) a += mf[i * NFN + j] * jump[j];
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        bq += ml[i * NK + k] * ddk[k];
        Q[k] += ml[i * NK + k] * jump[i];
      }
      add_row(FN::v[f][i], sv * a + sd * bq);
    }
    // lift of the normal-derivative part: acc_n += sum_{k,d} gk[k][d][n] c_d sd Q_k
    const double* gk = fc + F::GK + f * NK * D * NL;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const dd qk = sd * Q[k];
      dd wk[D];
#pragma unroll
      for (int d = 0; d < D; ++d) wk[d] = c[d] * qk;
#pragma unroll 4
      for (int n = 0; n < NL; ++n) {
        dd v = gk[(k * D) * NL + n] * wk[0];
#pragma unroll
        for (int d = 1; d < D; ++d) v += gk[(k * D + d) * NL + n] * wk[d];
        add_row(n, v);
      }
    }
  };
  face(std::integral_constant<int, 0>{});
  face(std::integral_constant<int, 1>{});
  face(std::integral_constant<int, 2>{});
  if constexpr (NF > 3) face(std::integral_constant<int, 3>{});
  if (store) {
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      const dd s = dd(b[(size_t)e * NL + n]) - dd(ahi[n][tl], alo[n][tl]);
      r[(size_t)e * NL + n] = s.hi;
    }
  }
}

// x = S b  (a smoothing step from x = 0, where the residual is b itself)
template <typename T, int NL>
__global__ __launch_bounds__(256) void k_smooth_zero(int nele, T* __restrict__ x,
                                                     const T* __restrict__ b,
                                                     const T* __restrict__ S,
                                                     const int32_t* __restrict__ act) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele || (act && !act[e >> 6])) return;
  T bv[NL], xv[NL];
  load_row<T, NL>(b + (size_t)e * NL, bv);
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    T srow[NL];
    load_smoother_row<T, NL>(S, n, e, nele, srow);
    T s = 0;
#pragma unroll
    for (int m = 0; m < NL; ++m) s += srow[m] * bv[m];
    xv[n] = s;
  }
  store_row<T, NL>(x + (size_t)e * NL, xv);
}

// bdiagK[e][n][m] = sum_{j,k} G_jk Sref_jk[n][m]
//   + sum_f wf_f (pen_f M_f[n][m] - alpha_f sum_d c_fd (C_fd[n][m] + C_fd[m][n]))
// alpha = 1/2 on interior faces, 1 on Dirichlet faces (z02/get_bdiag_diag.py:105-284).
// Written as a sum of NT fixed reference matrices with per-element coefficients,
//   bdiagK[e] = sum_t coef[e][t] T_t,
// T: the symmetrised Sref_jk (NQ), then per face M_f and C_fd + C_fd^T (1 + D each), with
// coef = G_q | wf pen, -alpha wf c_fd (zero on faces without a term).  A work-group takes
// EB elements: their coefficients go to LDS, and thread nm keeps column nm of the NT
// tables in registers, so each output costs NT FMAs with broadcast LDS reads and the
// rows are written contiguously.
template <int D> struct BdiagDims {
  static constexpr int NL = D == 3 ? 20 : 10, NN = NL * NL, NF = D + 1;
  static constexpr int NQ = D == 3 ? 6 : 3, NT = NQ + NF * (1 + D);
  static constexpr int THREADS = (NN + 63) / 64 * 64, EB = 64;
};
template <int D>
__global__ __launch_bounds__(BdiagDims<D>::THREADS) void k_bdiag(
    int nele, const double* __restrict__ tab, const double* __restrict__ G,
    const double* __restrict__ cf, const double* __restrict__ wf,
    const double* __restrict__ pen, const int32_t* __restrict__ fcode,
    double* __restrict__ bdiag) {
  using B = BdiagDims<D>;
  constexpr int NN = B::NN, NF = B::NF, NQ = B::NQ, NT = B::NT, EB = B::EB;
  __shared__ double cs[EB][NT + 1];
  const int e0 = blockIdx.x * EB;
  const int t = threadIdx.x;
  if (t < EB && e0 + t < nele) {
    const int e = e0 + t;
#pragma unroll
    for (int q = 0; q < NQ; ++q) cs[t][q] = G[(size_t)q * nele + e];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const int type = fcode[(size_t)e * NF + f] & 15;
      double* o = &cs[t][NQ + f * (1 + D)];
      if (type >= 2) {
#pragma unroll
        for (int d = 0; d <= D; ++d) o[d] = 0.0;
        continue;
      }
      const double alpha = type == 0 ? 0.5 : 1.0;
      const double w = wf[(size_t)f * nele + e];
      o[0] = w * pen[(size_t)f * nele + e];
#pragma unroll
      for (int d = 0; d < D; ++d) o[1 + d] = -alpha * w * cf[(size_t)(f * D + d) * nele + e];
    }
  }
  __syncthreads();
  if (t >= NN) return;
  double tv[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) tv[k] = tab[k * NN + t];
  const int ne = min(EB, nele - e0);
  for (int j = 0; j < ne; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NT; ++k) acc = fma(cs[j][k], tv[k], acc);
    bdiag[(size_t)(e0 + j) * NN + t] = acc;
  }
}

// Right-hand side (z02/volume_mf_diff.py:195-239 volume, :298-357 Dirichlet faces with the
// owner's kappa as z05/volume_mf_um_diff_varprop.py:320-342, :360-400 Neumann faces):
//   b_n = detJ sum_m Mref[n][m] (f_m + alpha_m / dt)
//       + sum_{Dirichlet f} wf_f (pen_f sum_m M_f[n][m] u_m - sum_d c_fd sum_m C_fd[m][n] u_m)
//       + sum_{Neumann f} sdet_f sum_m M_f[n][m] g_m
// alpha (may be null) is the transient term's nodal data (z02/volume_mf_diff.py:230-238),
// g (may be null: no Neumann data) the nodal Neumann flux; on Neumann faces wf = sdet.
template <int D>
__global__ __launch_bounds__(256) void k_rhs(int nele, const double* __restrict__ refm,
                                             const double* __restrict__ detJ,
                                             const double* __restrict__ cf,
                                             const double* __restrict__ wf,
                                             const double* __restrict__ pen,
                                             const int32_t* __restrict__ fcode,
                                             const double* __restrict__ fsrc,
                                             const double* __restrict__ ubc,
                                             const double* __restrict__ gneu,
                                             const double* __restrict__ alpha, double inv_dt,
                                             double* __restrict__ b) {
  using M = Dm<D>;
  constexpr int NL = M::NL, NN = NL * NL;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (int64_t)nele * NL) return;
  const int e = (int)(tid / NL);
  const int n = (int)(tid - (int64_t)e * NL);
  const double* Mf = refm + D * D * NN;
  const double* Cf = Mf + M::NF * NN;
  const double* Mref = Cf + M::NF * D * NN;
  const double* fe = fsrc + (size_t)e * NL;
  const double* ue = ubc + (size_t)e * NL;
  double s = 0.0;
  for (int m = 0; m < NL; ++m) s += Mref[n * NL + m] * fe[m];
  if (alpha) {
    const double* ae = alpha + (size_t)e * NL;
    double t = 0.0;
    for (int m = 0; m < NL; ++m) t += Mref[n * NL + m] * ae[m];
    s += t * inv_dt;
  }
  s *= detJ[e];
  for (int f = 0; f < M::NF; ++f) {
    const int type = fcode[e * M::NF + f] & 15;
    if (type == 1) {
      double pm = 0.0, cm = 0.0;
      for (int m = 0; m < NL; ++m) {
        pm += Mf[f * NN + n * NL + m] * ue[m];
        double t = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d)
          t += cf[(f * D + d) * nele + e] * Cf[(f * D + d) * NN + m * NL + n];
        cm += t * ue[m];
      }
      s += wf[f * nele + e] * (pen[f * nele + e] * pm - cm);
    } else if (type == 3 && gneu) {
      const double* ge = gneu + (size_t)e * NL;
      double pm = 0.0;
      for (int m = 0; m < NL; ++m) pm += Mf[f * NN + n * NL + m] * ge[m];
      s += wf[f * nele + e] * pm;
    }
  }
  b[tid] = s;
}

int launch_rhs(dgmi_ctx* c, const double* refm, const double* detJ, const double* f,
               const double* ubc, const double* gneu, const double* alpha, double inv_dt,
               double* b, hipStream_t s) {
  const int nele = (int)c->nele;
  const int64_t nthreads = (int64_t)nele * c->nloc;
  const dim3 blk(256), grd((unsigned)((nthreads + 255) / 256));
  if (c->D == 3)
    hipLaunchKernelGGL(k_rhs<3>, grd, blk, 0, s, nele, refm, detJ, c->g64.cf, c->g64.wf,
                       c->g64.pen, c->fcode, f, ubc, gneu, alpha, inv_dt, b);
  else
    hipLaunchKernelGGL(k_rhs<2>, grd, blk, 0, s, nele, refm, detJ, c->g64.cf, c->g64.wf,
                       c->g64.pen, c->fcode, f, ubc, gneu, alpha, inv_dt, b);
  DGMI_HIP(hipGetLastError());
  return 0;
}

// Column m of the 3-sweep damped point-Jacobi operator on one block, times omega:
// c = 0; 3x c += w (e_m - B c)/diag(B); S[:, m] = w c.
template <typename T, int NL>
__global__ __launch_bounds__(256) void k_build_S_jacobi(int nele, const double* __restrict__ B,
                                                        double omega, T* __restrict__ S) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (int64_t)nele * NL) return;
  const int e = (int)(tid / NL);
  const int m = (int)(tid - (int64_t)e * NL);
  const double* Be = B + (size_t)e * NL * NL;
  double c[NL];
#pragma unroll
  for (int n = 0; n < NL; ++n) c[n] = 0.0;
  for (int it = 0; it < 3; ++it) {
    double cn[NL];
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      double bc = 0.0;
#pragma unroll
      for (int k = 0; k < NL; ++k) bc += Be[n * NL + k] * c[k];
      const double rhs = (n == m ? 1.0 : 0.0) - bc;
      cn[n] = c[n] + omega * rhs / Be[n * NL + n];
    }
#pragma unroll
    for (int n = 0; n < NL; ++n) c[n] = cn[n];
  }
#pragma unroll
  for (int n = 0; n < NL; ++n) S[s_index<T, NL>(n, m, e, nele)] = (T)(omega * c[n]);
}

// omega * B^-1 for every element (the exact block smoother of
// z05/volume_mf_um_diff_varprop.py:384-392) in FP64, written straight into the packed
// symmetric smoother copies: P64 (FP64) and S32 (FP32) entry (a, b), a <= b, =
// (S[a][b] + S[b][a]) / 2 with S = omega B^-1; S16 (bfloat16) when set; H/Hs (FP16 pairs
// over the element's largest magnitude, which Hs keeps) when set.
// A wave holds 64 / NL elements, row r of element g in the registers of lane g NL + r:
// in-place Gauss-Jordan with partial pivoting.  The pivot of column c is the row, among
// those not yet used, with the largest |a[r][c]| to 15 mantissa bits (lowest row among
// those); the pivot keys and then the pivot row go through LDS to the element's lanes,
// and rows are never swapped, so row p_c ends up holding row c of B^-1 with columns
// permuted: stored[p_c][j] = B^-1[c][p_j].  The blocks come in through LDS with coalesced
// loads, the rows go back there, and the work-group forms the packed entries and writes
// EB consecutive elements of every packed vector together.  A singular block sets *fail
// to 1 + its element index.
// Orders LDS writes and reads among the lanes of one wave (its LDS operations execute in
// order; the fences keep the compiler from moving accesses across)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int SDIR_WAVES = 4;
template <int NL> struct SdirDims {
  static constexpr int G = 64 / NL;                // elements per wave
  static constexpr int EB = G * SDIR_WAVES;        // elements per work-group
  static constexpr int NT = 64 * SDIR_WAVES;       // threads
  static constexpr int NS = sym_size<NL>(), NP = (NS + 1) / 2;
  static constexpr int PPT = (NS * EB + NT - 1) / NT;  // packed entries per thread
  static constexpr int RP = NT / EB;                // threads per element in the max
};
template <int NL>
__global__ __launch_bounds__(64 * SDIR_WAVES) void k_build_S_direct(
    int nele, const double* __restrict__ B, double omega, double* __restrict__ P64,
    float* __restrict__ S32, uint16_t* __restrict__ S16, uint32_t* __restrict__ H,
    float* __restrict__ Hs, int* __restrict__ fail) {
  using Q = SdirDims<NL>;
  constexpr int G = Q::G, EB = Q::EB, NT = Q::NT, NS = Q::NS, NP = Q::NP;
  constexpr int PPT = Q::PPT, RP = Q::RP;
  static_assert(NS * EB <= EB * NL * NL, "packed entries fit in the row storage");
  __shared__ __align__(16) double L[EB * NL * NL];  // blocks/rows [le][r][j]; packed [i][le]
  constexpr int NK = (NL + 3) / 4 * 4;                // key rows padded to 16 bytes
  __shared__ __align__(16) uint32_t ck[EB][NK];        // pivot keys of column c
  __shared__ __align__(16) double pv[EB][NL];          // the pivot row, scaled
  __shared__ int pr[EB][NL], qi[EB][NL];
  __shared__ uint8_t ab[NS][2];  // (a, b) of packed entry i
  __shared__ double red[RP][EB];
  __shared__ double mx[EB];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = lane / NL, r = lane - (lane / NL) * NL;
  const bool row = g < G;  // lanes past G * NL hold no row
  const int le = w * G + (row ? g : 0);  // row-less lanes mirror the wave's first element
  const int e0 = blockIdx.x * EB;
  const bool act = row && e0 + le < nele;
  // the work-group's blocks are contiguous in B: coalesced 16-byte loads into LDS, then
  // each lane takes its row from there
  const int ne = min(EB, nele - e0);
  {
    const double2* src = reinterpret_cast<const double2*>(B + (size_t)e0 * NL * NL);
    double2* dst = reinterpret_cast<double2*>(L);
    for (int q = threadIdx.x; q < ne * NL * NL / 2; q += NT) dst[q] = src[q];
  }
  __syncthreads();
  double a[NL];
  {
    const double2* src = reinterpret_cast<const double2*>(L + (le * NL + r) * NL);
#pragma unroll
    for (int m = 0; m < NL / 2; ++m) {
      const double2 t = src[m];
      a[2 * m] = act ? t.x : 0.0;
      a[2 * m + 1] = act ? t.y : 0.0;
    }
  }
  if (row && r < NK - NL) ck[le][NL + r] = 0u;  // key padding
  for (int i = threadIdx.x; i < NS; i += NT) {
    int ra = 0, rem = i;
    while (rem >= NL - ra) {
      rem -= NL - ra;
      ++ra;
    }
    ab[i][0] = (uint8_t)ra;
    ab[i][1] = (uint8_t)(ra + rem);
  }
  bool used = !act;
  bool singular = false;
#pragma unroll
  for (int c = 0; c < NL; ++c) {
    // pivot key: the upper word of |a[r][c]| (exponent and 20 mantissa bits) with its low
    // 5 bits replaced by 31 - r, so the largest key is the largest magnitude to 15
    // mantissa bits, the lowest row among those; 0 for used rows and NaN.  The keys and
    // the pivot row stay within the element's wave.
    const double v = fabs(a[c]);
    const uint32_t hi = (uint32_t)((uint64_t)__double_as_longlong(v) >> 32);
    if (row) ck[le][r] = (used || !(v >= 0.0)) ? 0u : ((hi & ~31u) | (uint32_t)(31 - r));
    wave_lds_sync();
    uint32_t kmax = 0;
#pragma unroll
    for (int k = 0; k < NK / 4; ++k) {
      const uint4 t = reinterpret_cast<const uint4*>(ck[le])[k];
      kmax = max(max(kmax, max(t.x, t.y)), max(t.z, t.w));
    }
    const bool ok = (kmax >> 5) != 0;  // a nonzero candidate exists
    const int p = ok ? 31 - (int)(kmax & 31u) : 0;
    singular = singular || !ok;
    const bool me = act && r == p;
    // the pivot lane publishes its row scaled by d = 1 / a[p][c], with d in column c
    if (me) {
      const double d = 1.0 / a[c];
      double2* dst = reinterpret_cast<double2*>(pv[le]);
#pragma unroll
      for (int j = 0; j < NL / 2; ++j)
        dst[j] = make_double2(2 * j == c ? d : a[2 * j] * d, 2 * j + 1 == c ? d : a[2 * j + 1] * d);
    }
    if (row && r == 0) pr[le][c] = p;
    wave_lds_sync();
    const double f = a[c];
    // pivot row: a[j] = pv[j]; the others: a[j] - f pv[j] (sc = 0, fx = -1 give pv[j])
    const double sc = me ? 0.0 : 1.0, fx = me ? -1.0 : f;
    double pr2[NL];
#pragma unroll
    for (int j = 0; j < NL / 2; ++j) {
      const double2 t = reinterpret_cast<const double2*>(pv[le])[j];
      pr2[2 * j] = t.x;
      pr2[2 * j + 1] = t.y;
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      if (j == c) continue;
      a[j] = fma(-fx, pr2[j], a[j] * sc);
    }
    const double d = ok ? pr2[c] : 0.0;
    a[c] = me ? d : -f * d;
    used = used || me;
    wave_lds_sync();  // every lane has read ck and pv before the next column rewrites them
  }
  if (row) {
    double2* dst = reinterpret_cast<double2*>(L + (le * NL + r) * NL);
#pragma unroll
    for (int j = 0; j < NL / 2; ++j) dst[j] = make_double2(a[2 * j], a[2 * j + 1]);
    qi[le][r] = r;  // a singular block may repeat a pivot row: keep indices in range
  }
  if (act && r == 0 && singular) atomicCAS(fail, 0, 1 + e0 + le);
  __syncthreads();
  if (row) qi[le][pr[le][r]] = r;
  __syncthreads();
  // packed entries (i, element j), B^-1[a][b] = stored[p_a][q_b]
  double pk[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int idx = threadIdx.x + k * NT;
    pk[k] = 0.0;
    if (idx < NS * EB) {
      const int i = idx / EB, j = idx - (idx / EB) * EB;
      const int ra = ab[i][0], rb = ab[i][1];
      const double xab = L[(j * NL + pr[j][ra]) * NL + qi[j][rb]];
      const double xba = L[(j * NL + pr[j][rb]) * NL + qi[j][ra]];
      pk[k] = 0.5 * (omega * xab + omega * xba);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int idx = threadIdx.x + k * NT;
    if (idx < NS * EB) L[idx] = pk[k];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < NS * EB; idx += NT) {
    const int i = idx / EB, j = idx - (idx / EB) * EB;
    if (j >= ne) continue;
    const double v = L[idx];
    const size_t o = (size_t)i * nele + e0 + j;
    P64[o] = v;
    S32[o] = (float)v;
    if (S16) S16[o] = f32_to_bf16((float)v);
  }
  if (H) {
    // largest magnitude per element: RP partial maxima each, then one thread per element
    if ((int)threadIdx.x < RP * EB) {
      const int j = threadIdx.x % EB, part = threadIdx.x / EB;
      double m = 0.0;
      for (int i = part; i < NS; i += RP) m = fmax(m, fabs(L[i * EB + j]));
      red[part][j] = m;
    }
    __syncthreads();
    if ((int)threadIdx.x < EB) {
      double m = 0.0;
#pragma unroll
      for (int q = 0; q < RP; ++q) m = fmax(m, red[q][threadIdx.x]);
      mx[threadIdx.x] = m;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < NP * EB; idx += NT) {
      const int q = idx / EB, j = idx - (idx / EB) * EB;
      if (j >= ne) continue;
      const float mf = (float)mx[j];
      const double inv = mx[j] > 0.0 ? 1.0 / (double)mf : 0.0;
      const int i0 = 2 * q, i1 = 2 * q + 1;
      const _Float16 lo = (_Float16)(float)(L[i0 * EB + j] * inv);
      const _Float16 hi = i1 < NS ? (_Float16)(float)(L[i1 * EB + j] * inv) : (_Float16)0.0f;
      H[(size_t)q * nele + e0 + j] = (uint32_t)__builtin_bit_cast(uint16_t, lo) |
                                     ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
    }
    if ((int)threadIdx.x < ne) Hs[e0 + threadIdx.x] = (float)mx[threadIdx.x];
  }
}

// Packed symmetric copies of the full FP64 smoother: P[(a,b)] = (S[a][b] + S[b][a]) / 2,
// a <= b, in FP64, FP32 and (when S16 is set) bfloat16
template <int NL>
__global__ __launch_bounds__(256) void k_pack_S_sym(int nele, const double* __restrict__ S64,
                                                    double* __restrict__ P64,
                                                    float* __restrict__ S32,
                                                    uint16_t* __restrict__ S16) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele) return;
  for (int a = 0; a < NL; ++a)
    for (int b = a; b < NL; ++b) {
      const double v = 0.5 * (S64[s_index<double, NL>(a, b, e, nele)] +
                              S64[s_index<double, NL>(b, a, e, nele)]);
      P64[(size_t)sym_index<NL>(a, b) * nele + e] = v;
      S32[(size_t)sym_index<NL>(a, b) * nele + e] = (float)v;
      if (S16) S16[(size_t)sym_index<NL>(a, b) * nele + e] = f32_to_bf16((float)v);
    }
}

int launch_build_S_direct(dgmi_ctx* c, const double* bdiag, double omega, bool half, int* fail,
                          hipStream_t s) {
  const int nele = (int)c->nele;
  const int eb = c->D == 3 ? SdirDims<20>::EB : SdirDims<10>::EB;
  const dim3 blk(64 * SDIR_WAVES), grd((nele + eb - 1) / eb);
  uint32_t* H = half ? c->g32.Sh : nullptr;
  float* Hs = half ? c->g32.Shs : nullptr;
  if (c->D == 3)
    hipLaunchKernelGGL(k_build_S_direct<20>, grd, blk, 0, s, nele, bdiag, omega, c->g64.S,
                       c->g32.S, c->g32.S16, H, Hs, fail);
  else
    hipLaunchKernelGGL(k_build_S_direct<10>, grd, blk, 0, s, nele, bdiag, omega, c->g64.S,
                       c->g32.S, c->g32.S16, H, Hs, fail);
  DGMI_HIP(hipGetLastError());
  return 0;
}

// FP16 copy of the packed smoother: per element the largest magnitude m (kept in FP32) and
// the entries divided by m, rounded to FP16 (relative error 2^-11 of m), two per word
template <int NL>
__global__ __launch_bounds__(256) void k_pack_S_half(int nele, const double* __restrict__ P64,
                                                     uint32_t* __restrict__ H,
                                                     float* __restrict__ scale) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele) return;
  constexpr int NS = sym_size<NL>(), NP = (NS + 1) / 2;
  double m = 0.0;
  for (int i = 0; i < NS; ++i) m = fmax(m, fabs(P64[(size_t)i * nele + e]));
  const float mf = (float)m;
  const double inv = m > 0.0 ? 1.0 / (double)mf : 0.0;
  scale[e] = mf;
  for (int p = 0; p < NP; ++p) {
    const int i0 = 2 * p, i1 = 2 * p + 1;
    const _Float16 lo = (_Float16)(float)(P64[(size_t)i0 * nele + e] * inv);
    const _Float16 hi = i1 < NS ? (_Float16)(float)(P64[(size_t)i1 * nele + e] * inv) : (_Float16)0.0f;
    H[(size_t)p * nele + e] = (uint32_t)__builtin_bit_cast(uint16_t, lo) |
                              ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
  }
}
int launch_pack_S_half(dgmi_ctx* c, hipStream_t s) {
  const int nele = (int)c->nele;
  const dim3 blk(256), grd((nele + 255) / 256);
  if (c->D == 3)
    hipLaunchKernelGGL(k_pack_S_half<20>, grd, blk, 0, s, nele, (const double*)c->g64.S,
                       c->g32.Sh, c->g32.Shs);
  else
    hipLaunchKernelGGL(k_pack_S_half<10>, grd, blk, 0, s, nele, (const double*)c->g64.S,
                       c->g32.Sh, c->g32.Shs);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int launch_pack_S_sym(dgmi_ctx* c, const double* S, hipStream_t s) {
  const int nele = (int)c->nele;
  const dim3 blk(256), grd((nele + 255) / 256);
  if (c->D == 3)
    hipLaunchKernelGGL(k_pack_S_sym<20>, grd, blk, 0, s, nele, S, c->g64.S, c->g32.S, c->g32.S16);
  else
    hipLaunchKernelGGL(k_pack_S_sym<10>, grd, blk, 0, s, nele, S, c->g64.S, c->g32.S, c->g32.S16);
  DGMI_HIP(hipGetLastError());
  return 0;
}

__global__ void k_extract_diag(int64_t n, int nl, const double* __restrict__ B,
                               double* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t e = i / nl;
  const int k = (int)(i - e * nl);
  d[i] = B[(e * nl + k) * nl + k];
}

template <typename T>
__global__ void k_convert(const double* __restrict__ in, T* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[i];
}

// ---------------------------------------------------------------------------------------
// Trace buffers of the vector family: 0 = g.trace, 1 = g.trace2.  tin >= 0: x's traces are
// already in buffer tin (no trace pass); tin < 0: they are computed into buffer 0.  Mode 5
// with tout >= 0 writes the new x's traces into buffer tout (tout != the buffer read).
// Face-trace buffers of the two-pass kernels, allocated at their first use (the
// single-pass kernels need none): FP64 one, FP32 two (sweeps alternate them)
template <typename T>
bool ensure_traces(dgmi_ctx* c, DevGeom<T>& g) {
  using V2 = typename Vec2<T>::type;
  const size_t bytes = sizeof(V2) * c->nface * c->sngi * c->nele;
  if (!g.trace) g.trace = static_cast<V2*>(c->alloc(bytes));
  if (std::is_same<T, float>::value && !g.trace2) g.trace2 = static_cast<V2*>(c->alloc(bytes));
  if (!g.trace || (std::is_same<T, float>::value && !g.trace2)) {
    set_error(alloc_failure("face traces"));
    return false;
  }
  return true;
}
template bool ensure_traces<double>(dgmi_ctx*, DevGeom<double>&);
template bool ensure_traces<float>(dgmi_ctx*, DevGeom<float>&);

// the vector-kernel applies in one pass (every interior face needs its frame)
bool single_pass_apply(const dgmi_ctx* c) { return c->frames_ok && c->single_pass >= 1; }

template <typename T, int D>
static int apply_impl(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, T* r, int mode,
                      hipStream_t s, int tin = -1, int tout = -1) {
  const int nele = (int)c->nele;
  const dim3 blk(64 * (D + 1)), grd((nele + 63) / 64);
  const int32_t* act = c->mask_on ? c->act64 : nullptr;
  using V2 = typename Vec2<T>::type;
  V2* bufs[2] = {g.trace, g.trace2};
  if ((mode <= 1 || mode == 4) && tin < 0 && single_pass_apply(c)) {
    // one pass: traces formed from the rows of x (own and neighbours')
    const T* vt = g.tab + g.off_vtab;
    const T* wg = g.tab + g.off_wgt;
    const T* sw = g.tab + g.off_sw;
    if constexpr (std::is_same<T, float>::value) {
      if (c->staged) {
#define DGMI_ST(MODE)                                                                         \
  hipLaunchKernelGGL((k_staged<D, MODE>), grd, blk, 0, s, nele, x, b, r, vt, wg, g.G, sw, g.cf, \
                     g.wf, g.pen, c->nbface, c->fcode, g.S, act, (float)c->omega,              \
                     g.tab + g.off_fc, c->fnodes, g.cnb, c->nperm)
        if (mode == 0) DGMI_ST(0);
        else if (mode == 1) DGMI_ST(1);
        else DGMI_ST(4);
#undef DGMI_ST
        DGMI_HIP(hipGetLastError());
        return 0;
      }
    }
#define DGMI_RES1(MODE)                                                                       \
  hipLaunchKernelGGL((k_residual_w<T, D, MODE, false, T, true>), grd, blk, 0, s, nele, x, b, r, \
                     vt, wg, sw, c->aln, g.G, g.cf, g.wf, g.pen, c->nbface, c->fcode,          \
                     (const V2*)nullptr, g.S, act, (T)c->omega, (V2*)nullptr,                  \
                     g.tab + g.off_fc, c->fnodes, g.cnb, c->nperm)
    if (mode == 0) DGMI_RES1(0);
    else if (mode == 1) DGMI_RES1(1);
    else DGMI_RES1(4);
#undef DGMI_RES1
    DGMI_HIP(hipGetLastError());
    return 0;
  }
  if (!ensure_traces(c, g)) return -3;
  bufs[0] = g.trace;
  bufs[1] = g.trace2;
  if (mode == 6) {
    tin = tout == 0 ? 1 : 0;  // a sweep from zero reads no traces
  } else if (tin < 0) {
    hipLaunchKernelGGL((k_trace_w<T, D>), grd, blk, 0, s, nele, x, g.tab + g.off_fc, c->fnodes,
                       g.cf, g.trace, act);
    tin = 0;
  }
  if (!(tin <= 1 && tout <= 1 && tout != tin && (tout < 0 || mode >= 5) && bufs[tin] &&
        (tout < 0 || bufs[tout]))) {
    set_error("bad trace buffers");
    return -2;
  }
  const V2* trace = bufs[tin];
  V2* to = tout >= 0 ? bufs[tout] : nullptr;
  const T* vt = g.tab + g.off_vtab;
  const T* wg = g.tab + g.off_wgt;
  const T* sw = g.tab + g.off_sw;
#define DGMI_RES(MODE, EMIT)                                                                  \
  hipLaunchKernelGGL((k_residual_w<T, D, MODE, EMIT>), grd, blk, 0, s, nele, x, b, r, vt, wg,  \
                     sw, c->aln, g.G, g.cf, g.wf, g.pen, c->nbface, c->fcode, trace, g.S,     \
                     act, (T)c->omega, to, g.tab + g.off_fc, c->fnodes, g.cnb, c->nperm)
  switch (mode) {
    case 0: DGMI_RES(0, false); break;
    case 1: DGMI_RES(1, false); break;
    case 2: DGMI_RES(2, false); break;
    case 3: DGMI_RES(3, false); break;
    case 5:
      if (to) DGMI_RES(5, true);
      else DGMI_RES(5, false);
      break;
    case 6:
      if (to) DGMI_RES(6, true);
      else DGMI_RES(6, false);
      break;
    default: DGMI_RES(4, false); break;
  }
#undef DGMI_RES
  DGMI_HIP(hipGetLastError());
  return 0;
}

// The sweeps of the bf16 inner cycles: as modes 5 and 6 with traces out, reading the
// bfloat16 copy of the smoother (SS = uint16_t)
template <int D, typename SS>
static int sweep_s16(dgmi_ctx* c, float* x, const float* b, int mode, int tin, int tout,
                     hipStream_t s) {
  DevGeom<float>& g = c->g32;
  const int nele = (int)c->nele;
  const dim3 blk(64 * (D + 1)), grd((nele + 63) / 64);
  const int32_t* act = c->mask_on ? c->act64 : nullptr;
  using V2 = typename Vec2<float>::type;
  V2* bufs[2] = {g.trace, g.trace2};
  const SS* S16 = (const SS*)g.S16;
  if (!ensure_traces(c, g)) return -3;
  bufs[0] = g.trace;
  bufs[1] = g.trace2;
  if (mode == 6) tin = tout == 0 ? 1 : 0;
  if (!(S16 && (mode == 5 || mode == 6) && tin >= 0 && tin <= 1 && tout >= 0 && tout <= 1 &&
        tin != tout && bufs[tin] && bufs[tout])) {
    set_error("bad trace buffers or no bf16 smoother");
    return -2;
  }
  const float* vt = g.tab + g.off_vtab;
  const float* wg = g.tab + g.off_wgt;
  const float* sw = g.tab + g.off_sw;
#define DGMI_RES16(MODE)                                                                      \
  hipLaunchKernelGGL((k_residual_w<float, D, MODE, true, SS>), grd, blk, 0, s, nele, x, b,     \
                     nullptr, vt, wg, sw, c->aln, g.G, g.cf, g.wf, g.pen, c->nbface,           \
                     c->fcode, (const V2*)bufs[tin], S16, act, (float)c->omega, bufs[tout],    \
                     g.tab + g.off_fc, c->fnodes, g.cnb, c->nperm)
  if (mode == 5) DGMI_RES16(5);
  else DGMI_RES16(6);
#undef DGMI_RES16
  DGMI_HIP(hipGetLastError());
  return 0;
}
template <typename SS>
static int sweep16(dgmi_ctx* c, float* x, const float* b, int mode, int tin, int tout,
                   hipStream_t s) {
  return c->D == 3 ? sweep_s16<3, SS>(c, x, b, mode, tin, tout, s)
                   : sweep_s16<2, SS>(c, x, b, mode, tin, tout, s);
}

// FP32 vectors go to the matrix-core kernels unless the context selects the vector ones
template <typename T>
static int apply_dispatch(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, T* r, int mode,
                          hipStream_t s) {
  if constexpr (std::is_same<T, float>::value)
    if (c->f32_kernel != 0) return launch_apply_mf(c, c->f32_kernel == 2, x, b, r, mode, s);
  if (c->D == 3) return apply_impl<T, 3>(c, g, x, b, r, mode, s);
  return apply_impl<T, 2>(c, g, x, b, r, mode, s);
}

template <typename T>
int launch_apply(dgmi_ctx* c, DevGeom<T>& g, const T* x, const T* b, T* r, hipStream_t s) {
  T* xm = const_cast<T*>(x);  // not written in modes 0/1
  return apply_dispatch<T>(c, g, xm, b, r, b ? 1 : 0, s);
}

template <typename T>
int launch_apply_pos(dgmi_ctx* c, DevGeom<T>& g, const T* x, T* r, hipStream_t s) {
  return apply_dispatch<T>(c, g, const_cast<T*>(x), nullptr, r, 4, s);
}

// FP32 sweeps of the exact block smoother (S = omega A_ee^-1) use the neighbour-only form
// (mode 5) unless the context turns it off; FP64 keeps the reference's x += S (b - A x).
template <typename T>
static bool smooth_offd(const dgmi_ctx* c) {
  return std::is_same<T, float>::value && c->smooth_offdiag &&
         c->smoother_variant == DGMI_SMOOTH_DIRECT;
}

template <typename T>
int launch_smooth(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, T* r_out, hipStream_t s) {
  const bool offd = !r_out && smooth_offd<T>(c);
  return apply_dispatch<T>(c, g, x, b, r_out, offd ? 5 : (r_out ? 3 : 2), s);
}

// The preconditioner's sweeps with traces carried between them (FP32 vector kernels in
// the neighbour form only; trace_io() tells whether the context runs that way).
bool trace_io(const dgmi_ctx* c) {
  return c->f32_kernel == 0 && smooth_offd<float>(c);
}
int launch_smooth_io(dgmi_ctx* c, float* x, const float* b, int tin, int tout, hipStream_t s,
                     bool s16) {
  if (s16) return sweep16<uint16_t>(c, x, b, 5, tin, tout, s);
  if (c->D == 3) return apply_impl<float, 3>(c, c->g32, x, b, nullptr, 5, s, tin, tout);
  return apply_impl<float, 2>(c, c->g32, x, b, nullptr, 5, s, tin, tout);
}
int launch_smooth_zero_io(dgmi_ctx* c, float* x, const float* b, int tout, hipStream_t s,
                          bool s16) {
  if (s16) return sweep16<uint16_t>(c, x, b, 6, -1, tout, s);
  if (c->D == 3) return apply_impl<float, 3>(c, c->g32, x, b, nullptr, 6, s, -1, tout);
  return apply_impl<float, 2>(c, c->g32, x, b, nullptr, 6, s, -1, tout);
}
bool bf16_smoother(const dgmi_ctx* c) {
  return smooth_offd<float>(c) && c->g32.S16;
}
int launch_apply_io(dgmi_ctx* c, const float* x, const float* b, float* r, int mode, int tin,
                    hipStream_t s) {
  float* xm = const_cast<float*>(x);
  if (c->D == 3) return apply_impl<float, 3>(c, c->g32, xm, b, r, mode, s, tin, -1);
  return apply_impl<float, 2>(c, c->g32, xm, b, r, mode, s, tin, -1);
}

// FP32 neighbour-form sweeps in one pass: x_new = (1 - omega) x + S (b - A_off x), the
// neighbours' face traces formed from their rows of x; x_new must be another vector than x
bool single_pass_sweeps(const dgmi_ctx* c) {
  return c->single_pass >= 2 && c->frames_ok && c->f32_kernel == 0 && smooth_offd<float>(c);
}
template <int D>
static int sweep_nbx(dgmi_ctx* c, const float* x, const float* b, float* xnew, hipStream_t s) {
  DevGeom<float>& g = c->g32;
  const int nele = (int)c->nele;
  const dim3 blk(64 * (D + 1)), grd((nele + 63) / 64);
  const int32_t* act = c->mask_on ? c->act64 : nullptr;
  using V2 = typename Vec2<float>::type;
  if (c->staged && c->smoother_half && g.Sh)
    hipLaunchKernelGGL((k_staged<D, 5, true>), grd, blk, 0, s, nele, x, b, xnew,
                       g.tab + g.off_vtab, g.tab + g.off_wgt, g.G, g.tab + g.off_sw, g.cf, g.wf,
                       g.pen, c->nbface, c->fcode, g.S, act, (float)c->omega, g.tab + g.off_fc,
                       c->fnodes, g.cnb, c->nperm, (const uint32_t*)g.Sh, (const float*)g.Shs);
  else if (c->staged)
    hipLaunchKernelGGL((k_staged<D, 5>), grd, blk, 0, s, nele, x, b, xnew, g.tab + g.off_vtab,
                       g.tab + g.off_wgt, g.G, g.tab + g.off_sw, g.cf, g.wf, g.pen, c->nbface,
                       c->fcode, g.S, act, (float)c->omega, g.tab + g.off_fc, c->fnodes, g.cnb,
                       c->nperm, (const uint32_t*)nullptr, (const float*)nullptr);
  else
    hipLaunchKernelGGL((k_residual_w<float, D, 5, false, float, true>), grd, blk, 0, s, nele,
                       const_cast<float*>(x), b, xnew, g.tab + g.off_vtab, g.tab + g.off_wgt,
                       g.tab + g.off_sw, c->aln, g.G, g.cf, g.wf, g.pen, c->nbface, c->fcode,
                       (const V2*)nullptr, g.S, act, (float)c->omega, (V2*)nullptr,
                       g.tab + g.off_fc, c->fnodes, g.cnb, c->nperm);
  DGMI_HIP(hipGetLastError());
  return 0;
}
int launch_sweep_nbx(dgmi_ctx* c, const float* x, const float* b, float* xnew, hipStream_t s) {
  if (x == xnew) {
    set_error("a single-pass sweep writes another vector than it reads");
    return -2;
  }
  return c->D == 3 ? sweep_nbx<3>(c, x, b, xnew, s) : sweep_nbx<2>(c, x, b, xnew, s);
}



int launch_residual_dd(dgmi_ctx* c, const double* x, const double* b, double* r, hipStream_t s) {
  const int nele = (int)c->nele;
  const DevGeom<double>& g = c->g64;
  if (c->frames_ok) {
    const dim3 blk(128), grd((nele + 127) / 128);
#define DGMI_RDD2(D)                                                                           \
  hipLaunchKernelGGL(k_residual_dd_nodes<D>, grd, blk, 0, s, nele, x, b, r, g.tab + g.off_vtab, \
                     g.tab + g.off_wgt, g.G, g.cf, g.cnb, g.wf, g.pen, c->nbface, c->fcode,    \
                     g.tab + g.off_fc, c->nperm)
    if (c->D == 3) DGMI_RDD2(3);
    else DGMI_RDD2(2);
#undef DGMI_RDD2
    DGMI_HIP(hipGetLastError());
    return 0;
  }
  const dim3 blk(64), grd((nele + 63) / 64);
#define DGMI_RDD(D)                                                                            \
  hipLaunchKernelGGL(k_residual_dd<D>, grd, blk, 0, s, nele, x, b, r, g.tab + g.off_vtab,       \
                     g.tab + g.off_wgt, g.tab + g.off_sw, c->aln, g.G, g.cf, g.wf, g.pen,       \
                     c->nbface, c->fcode, g.tab + g.off_fc, c->fnodes)
  if (c->D == 3) DGMI_RDD(3);
  else DGMI_RDD(2);
#undef DGMI_RDD
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_smooth_zero(dgmi_ctx* c, DevGeom<T>& g, T* x, const T* b, hipStream_t s) {
  const int nele = (int)c->nele;
  if constexpr (std::is_same<T, float>::value) {
    if (c->staged) {
      const int32_t* act = c->mask_on ? c->act64 : nullptr;
      const dim3 grd((nele + 63) / 64);
#define DGMI_ZERO(D, H)                                                                       \
  hipLaunchKernelGGL((k_staged<D, 6, H>), grd, dim3(64 * (D + 1)), 0, s, nele, x, b, x,        \
                     g.tab + g.off_vtab, g.tab + g.off_wgt, g.G, g.tab + g.off_sw, g.cf, g.wf, \
                     g.pen, c->nbface, c->fcode, g.S, act, (float)c->omega, g.tab + g.off_fc, \
                     c->fnodes, g.cnb, c->nperm, (const uint32_t*)g.Sh, (const float*)g.Shs)
      const bool half = c->smoother_half && g.Sh;
      if (c->D == 3) {
        if (half) DGMI_ZERO(3, true);
        else DGMI_ZERO(3, false);
      } else {
        if (half) DGMI_ZERO(2, true);
        else DGMI_ZERO(2, false);
      }
#undef DGMI_ZERO
      DGMI_HIP(hipGetLastError());
      return 0;
    }
  }
  const dim3 blk(256), grd((nele + 255) / 256);
  if (c->D == 3)
    hipLaunchKernelGGL((k_smooth_zero<T, 20>), grd, blk, 0, s, nele, x, b, g.S,
                       c->mask_on ? (const int32_t*)c->act64 : nullptr);
  else
    hipLaunchKernelGGL((k_smooth_zero<T, 10>), grd, blk, 0, s, nele, x, b, g.S,
                       c->mask_on ? (const int32_t*)c->act64 : nullptr);
  DGMI_HIP(hipGetLastError());
  return 0;
}

// Per block of 256 elements: partial[b] = sum_e |detJ_e| sum_n cn[n] d[e,n]^2 and
// partial[nb + b] = max |d|, d = x_num - x_ana, cn[n] = sum_g w_g n[n,g]^2 (the
// reference's l2 squares every basis term separately).  FP64 sums in a fixed tree.
template <typename T, int NL>
__global__ __launch_bounds__(256) void k_l2_error(int nele, const T* __restrict__ xn,
                                                  const T* __restrict__ xa,
                                                  const double* __restrict__ cn,
                                                  const double* __restrict__ detJ,
                                                  double* __restrict__ partial) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0, m = 0.0;
  if (e < nele) {
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      const double d = (double)xn[(size_t)e * NL + n] - (double)xa[(size_t)e * NL + n];
      s += cn[n] * d * d;
      m = fmax(m, fabs(d));
    }
    s *= fabs(detJ[e]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    m = fmax(m, __shfl_xor(m, off, 64));
  }
  __shared__ double ss[4], sm[4];
  if ((threadIdx.x & 63) == 0) {
    ss[threadIdx.x >> 6] = s;
    sm[threadIdx.x >> 6] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = (ss[0] + ss[1]) + (ss[2] + ss[3]);
    partial[gridDim.x + blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
  }
}

template <typename T>
int launch_l2_error(dgmi_ctx* c, const T* xn, const T* xa, const double* cn, double* partial,
                    int nb, hipStream_t s) {
  if (c->D == 3)
    hipLaunchKernelGGL((k_l2_error<T, 20>), dim3(nb), dim3(256), 0, s, (int)c->nele, xn, xa, cn,
                       c->detJ, partial);
  else
    hipLaunchKernelGGL((k_l2_error<T, 10>), dim3(nb), dim3(256), 0, s, (int)c->nele, xn, xa, cn,
                       c->detJ, partial);
  DGMI_HIP(hipGetLastError());
  return 0;
}
template int launch_l2_error<double>(dgmi_ctx*, const double*, const double*, const double*,
                                     double*, int, hipStream_t);
template int launch_l2_error<float>(dgmi_ctx*, const float*, const float*, const double*,
                                    double*, int, hipStream_t);

int launch_bdiag(dgmi_ctx* c, const double* tab, double* bdiag, hipStream_t s) {
  const int nele = (int)c->nele;
  if (c->D == 3) {
    using B = BdiagDims<3>;
    hipLaunchKernelGGL(k_bdiag<3>, dim3((nele + B::EB - 1) / B::EB), dim3(B::THREADS), 0, s,
                       nele, tab, c->g64.G, c->g64.cf, c->g64.wf, c->g64.pen, c->fcode, bdiag);
  } else {
    using B = BdiagDims<2>;
    hipLaunchKernelGGL(k_bdiag<2>, dim3((nele + B::EB - 1) / B::EB), dim3(B::THREADS), 0, s,
                       nele, tab, c->g64.G, c->g64.cf, c->g64.wf, c->g64.pen, c->fcode, bdiag);
  }
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_build_S_jacobi(dgmi_ctx* c, const double* bdiag, double omega, T* S, hipStream_t s) {
  const int nele = (int)c->nele;
  const int64_t nthreads = (int64_t)nele * c->nloc;
  const dim3 blk(256), grd((unsigned)((nthreads + 255) / 256));
  if (c->D == 3)
    hipLaunchKernelGGL((k_build_S_jacobi<T, 20>), grd, blk, 0, s, nele, bdiag, omega, S);
  else
    hipLaunchKernelGGL((k_build_S_jacobi<T, 10>), grd, blk, 0, s, nele, bdiag, omega, S);
  DGMI_HIP(hipGetLastError());
  return 0;
}

int launch_extract_diag(dgmi_ctx* c, const double* B, double* d, hipStream_t s) {
  const int64_t n = c->nele * c->nloc;
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_extract_diag, grd, blk, 0, s, n, c->nloc, B, d);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_convert(const double* in, T* out, int64_t n, hipStream_t s) {
  const dim3 blk(256), grd((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL((k_convert<T>), grd, blk, 0, s, in, out, n);
  DGMI_HIP(hipGetLastError());
  return 0;
}

template int launch_apply<double>(dgmi_ctx*, DevGeom<double>&, const double*, const double*,
                                  double*, hipStream_t);
template int launch_apply<float>(dgmi_ctx*, DevGeom<float>&, const float*, const float*, float*,
                                 hipStream_t);
template int launch_apply_pos<double>(dgmi_ctx*, DevGeom<double>&, const double*, double*,
                                      hipStream_t);
template int launch_apply_pos<float>(dgmi_ctx*, DevGeom<float>&, const float*, float*, hipStream_t);
template int launch_smooth<double>(dgmi_ctx*, DevGeom<double>&, double*, const double*, double*,
                                   hipStream_t);
template int launch_smooth<float>(dgmi_ctx*, DevGeom<float>&, float*, const float*, float*,
                                  hipStream_t);
template int launch_smooth_zero<double>(dgmi_ctx*, DevGeom<double>&, double*, const double*,
                                        hipStream_t);
template int launch_smooth_zero<float>(dgmi_ctx*, DevGeom<float>&, float*, const float*,
                                       hipStream_t);
template int launch_build_S_jacobi<double>(dgmi_ctx*, const double*, double, double*, hipStream_t);
template int launch_convert<double>(const double*, double*, int64_t, hipStream_t);
template int launch_convert<float>(const double*, float*, int64_t, hipStream_t);

}  // namespace dgmi
