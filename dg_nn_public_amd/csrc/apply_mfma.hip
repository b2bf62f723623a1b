// SIPG operator apply on the matrix cores (gfx950 MFMA), FP32 and BF16-input variants.
//
// The element tables of z02/shape_function.py (volume derivatives nlx, face values sn and
// face derivatives snlx) are the same for every element, so the apply is a pair of dense
// contractions with elements packed along N:
//
//   Y = T Xe         T: (R x nloc) stacked tables, Xe: (nloc x 16 elements)
//   pointwise        metric (volume points) and SIPG flux (face points), per element
//   r = T^T Z        the same table, transposed, lifts Z back to the test functions
//
// T has one group of four rows per quadrature point:
//   volume point g:      (dphi/dxi_0, dphi/dxi_1, dphi/dxi_2 | 0, 0)
//   face point (f, g):   (phi, dphi/dxi_0, dphi/dxi_1, dphi/dxi_2 | 0)
// so a 16-row MFMA output tile holds four points and lane l of the wave receives the whole
// group (4*(l/16) + i, i = 0..3) of point 4t + l/16 for element l%16: the metric and flux
// math stays in registers, and the accumulator registers are fed straight back as the B
// operand of the lift (the K order of T^T is permuted to match, once, on the host).
//
// Variants:  FP32  v_mfma_f32_16x16x4_f32   (A, B, C fp32)
//            BF16  v_mfma_f32_16x16x32_bf16 (A = table, B = x or Z rounded to bf16; C fp32)
// Geometry, traces, residual, smoother: fp32.  Same two-launch structure as the vector
// kernels: trace pass (own face values), then the residual/smoothing pass.
#include <cstring>

#include "dgmi_internal.h"
#include "rowvec.h"

namespace dgmi {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int D> struct Mq;
template <> struct Mq<3> { static constexpr int NL = 20, NGI = 24, SG = 12, NF = 4, NG = 6; };
template <> struct Mq<2> { static constexpr int NL = 10, NGI = 12, SG = 4, NF = 3, NG = 3; };

template <int D> struct MfShape {
  using M = Mq<D>;
  static constexpr int NT = (M::NGI + M::NF * M::SG) / 4;  // 16-row tiles of T
  static constexpr int NTV = M::NGI / 4;                    // volume tiles
  static constexpr int FT = M::SG / 4;                      // tiles per face
  static constexpr int KS = (M::NL + 3) / 4;                // fp32 k-steps over nloc
  static constexpr int MT = (M::NL + 15) / 16;              // output tiles over nloc
  static_assert(M::NGI % 4 == 0 && M::SG % 4 == 0, "points must fill whole tiles");
  static_assert(NT % 2 == 0, "bf16 lift pairs tiles");
  static_assert(M::NL <= 32, "bf16 contraction covers nloc in one k-step");
};

// ---- shared tables staged in LDS -----------------------------------------------------
// FP32: T itself, row-major [R][nloc] (+16 zero words so padded reads stay inside).  Both
// operand patterns are bank-conflict free for nloc = 20 and 10:
//   Y = T Xe   lane l reads T[16t + l%16][4s + l/16]      (word stride 20: distinct banks)
//   r = T^T Z  lane l reads T[16t + 4(l/16) + i][16mt + l%16]
// BF16: the two operand tables in MFMA lane order (16 bytes per lane, conflict free).
template <typename TIN, int D> struct MfLds;

// Copy 16-byte units [FIRST, LAST) from global memory to LDS with NTH threads.
template <int NTH, int FIRST, int LAST, typename V>
__device__ __forceinline__ void stage16(V* __restrict__ dst, const V* __restrict__ src, int tid) {
  for (int i = FIRST + tid; i < LAST; i += NTH) dst[i] = src[i];
}

template <int D> struct MfLds<float, D> {
  using M = Mq<D>;
  using S = MfShape<D>;
  static constexpr int R = 16 * S::NT, WORDS = R * M::NL + 16;
  static_assert(WORDS % 4 == 0 && (16 * S::NTV * M::NL) % 4 == 0, "16-byte staging");
  __align__(16) float T[WORDS];
  int aln[D * M::SG];
  // FACES: only the face rows (the trace pass uses no volume rows)
  template <int NTH, bool FACES>
  __device__ __forceinline__ void load(const float* __restrict__ tT, const int32_t* __restrict__ al,
                                       int tid) {
    constexpr int F4 = FACES ? 16 * S::NTV * M::NL / 4 : 0;
    stage16<NTH, F4, WORDS / 4>(reinterpret_cast<float4*>(T),
                                reinterpret_cast<const float4*>(tT), tid);
    for (int i = tid; i < D * M::SG; i += NTH) aln[i] = al[i];
  }
  struct XOp { float v[S::KS]; };
  __device__ __forceinline__ static XOp load_x(const float* __restrict__ x, int e, int q) {
    XOp o;
#pragma unroll
    for (int s = 0; s < S::KS; ++s) {
      const int n = 4 * s + q;
      o.v[s] = n < M::NL ? x[(size_t)e * M::NL + n] : 0.0f;
    }
    return o;
  }
  // Y tile t = T[16t:16t+16, :] Xe
  __device__ __forceinline__ v4f tile(int t, int j, int q, const XOp& xo) const {
    v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < S::KS; ++s) {
      const int n = 4 * s + q;
      const float a = (M::NL % 4 == 0 || n < M::NL) ? T[(16 * t + j) * M::NL + n] : 0.0f;
      c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xo.v[s], c, 0, 0, 0);
    }
    return c;
  }
};

template <int D> struct MfLds<__bf16, D> {
  using M = Mq<D>;
  using S = MfShape<D>;
  static constexpr int N1 = S::NT * 64, N2 = (S::NT / 2) * S::MT * 64;
  bf16x8 t1[N1];
  bf16x8 t2[N2];
  int aln[D * M::SG];
  // FACES: only the face tiles of t1 (the trace pass lifts nothing)
  template <int NTH, bool FACES>
  __device__ __forceinline__ void load(const bf16x8* __restrict__ b1, const bf16x8* __restrict__ b2,
                                       const int32_t* __restrict__ al, int tid) {
    using U = uint4;
    stage16<NTH, FACES ? S::NTV * 64 : 0, N1>(reinterpret_cast<U*>(t1),
                                              reinterpret_cast<const U*>(b1), tid);
    if constexpr (!FACES)
      stage16<NTH, 0, N2>(reinterpret_cast<U*>(t2), reinterpret_cast<const U*>(b2), tid);
    for (int i = tid; i < D * M::SG; i += NTH) aln[i] = al[i];
  }
  struct XOp { bf16x8 v; };
  __device__ __forceinline__ static XOp load_x(const float* __restrict__ x, int e, int q) {
    XOp o;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int n = 8 * q + j;
      o.v[j] = (__bf16)(n < M::NL ? x[(size_t)e * M::NL + n] : 0.0f);
    }
    return o;
  }
  __device__ __forceinline__ v4f tile(int t, int lane, const XOp& xo) const {
    v4f c = {0.f, 0.f, 0.f, 0.f};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(t1[t * 64 + lane], xo.v, c, 0, 0, 0);
  }
};

// Lift accumulator: acc[mt] += (T^T)[mt rows, tile t columns] Z_t.
template <typename TIN, int D> struct Lift;

template <int D> struct Lift<float, D> {
  using M = Mq<D>;
  using S = MfShape<D>;
  v4f acc[S::MT];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int mt = 0; mt < S::MT; ++mt) acc[mt] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  __device__ __forceinline__ void add(const MfLds<float, D>& L, int t, int j, int q, const v4f& z) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int mt = 0; mt < S::MT; ++mt) {
        const int n = 16 * mt + j;
        // columns n >= nloc read the next row (or the zero pad) and are masked
        const float a = L.T[(16 * t + 4 * q + i) * M::NL + n];
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(n < M::NL ? a : 0.0f, z[i], acc[mt], 0, 0, 0);
      }
  }
};

template <int D> struct Lift<__bf16, D> {
  using S = MfShape<D>;
  v4f acc[S::MT];
  v4f zeven;
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int mt = 0; mt < S::MT; ++mt) acc[mt] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  // tiles are lifted in pairs (2tp, 2tp+1): K = 32 = two tiles of four groups
  __device__ __forceinline__ void add(const MfLds<__bf16, D>& L, int t, int lane, const v4f& z) {
    if ((t & 1) == 0) {
      zeven = z;
      return;
    }
    bf16x8 b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      b[i] = (__bf16)zeven[i];
      b[4 + i] = (__bf16)z[i];
    }
    const int tp = t >> 1;
#pragma unroll
    for (int mt = 0; mt < S::MT; ++mt)
      acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(L.t2[(tp * S::MT + mt) * 64 + lane], b,
                                                        acc[mt], 0, 0, 0);
  }
};

// table pointers of one variant
struct MfArgs {
  const float* tT;   // fp32 [R][nloc] + pad
  const void* b1;    // bf16 lane-ordered operands
  const void* b2;
  const int32_t* aln;
};

template <typename TIN, int D, int NTH, bool FACES>
__device__ __forceinline__ void mf_load_tables(MfLds<TIN, D>& L, const MfArgs& a) {
  if constexpr (std::is_same<TIN, float>::value)
    L.template load<NTH, FACES>(a.tT, a.aln, threadIdx.x);
  else
    L.template load<NTH, FACES>((const bf16x8*)a.b1, (const bf16x8*)a.b2, a.aln, threadIdx.x);
}

template <typename TIN, int D>
__device__ __forceinline__ v4f mf_tile(const MfLds<TIN, D>& L, int t, int lane,
                                       const typename MfLds<TIN, D>::XOp& xo) {
  if constexpr (std::is_same<TIN, float>::value)
    return L.tile(t, lane & 15, lane >> 4, xo);
  else
    return L.tile(t, lane, xo);
}

template <typename TIN, int D>
__device__ __forceinline__ void mf_lift(Lift<TIN, D>& lf, const MfLds<TIN, D>& L, int t, int lane,
                                        const v4f& z) {
  if constexpr (std::is_same<TIN, float>::value)
    lf.add(L, t, lane & 15, lane >> 4, z);
  else
    lf.add(L, t, lane, z);
}

// Does the work-group touch an element of an unfinished solve?  act holds one flag per 64
// elements; a work-group covers 16*WPB elements.
template <int WPB>
__device__ __forceinline__ bool mf_block_active(const int32_t* __restrict__ act, int nele, int grp) {
  const int e0 = grp * 16 * WPB;
  const int e1 = min(e0 + 16 * WPB, nele) - 1;
  bool on = false;
  for (int b = e0 >> 6; b <= (e1 >> 6); ++b) on |= act[b] != 0;
  return on;
}

// ---------------------------------------------------------------------------------------
// Trace pass: u and du/dn at every own face point -> trace[(e*nface + f)*sngi + g]
// (element-major: a neighbour's face is one or two cache lines for the gathers).
// Block: WPB waves (tile shape chosen by the autotuner), 16 elements per wave.
template <typename TIN, int D, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_trace_mf(int nele, const float* __restrict__ x,
                                                  MfArgs ta, const float* __restrict__ cf,
                                                  float2* __restrict__ trace,
                                                  const int32_t* __restrict__ act) {
  using M = Mq<D>;
  using S = MfShape<D>;
  __shared__ MfLds<TIN, D> L;
  const int grp = act ? (int)blockIdx.x : xcd_group(blockIdx.x, gridDim.x);
  if (act && !mf_block_active<WPB>(act, nele, grp)) return;  // before any barrier
  const int lane = threadIdx.x & 63, q = lane >> 4;
  const int wave = grp * WPB + (threadIdx.x >> 6);
  const int el = wave * 16 + (lane & 15);
  const bool valid = el < nele;
  const int ec = valid ? el : nele - 1;
  const auto xo = MfLds<TIN, D>::load_x(x, ec, q);
  float c[M::NF][D];
#pragma unroll
  for (int f = 0; f < M::NF; ++f)
#pragma unroll
    for (int d = 0; d < D; ++d) c[f][d] = cf[(f * D + d) * nele + ec];
  mf_load_tables<TIN, D, 64 * WPB, true>(L, ta);
  __syncthreads();
#pragma unroll
  for (int f = 0; f < M::NF; ++f) {
#pragma unroll
    for (int tt = 0; tt < S::FT; ++tt) {
      const v4f y = mf_tile<TIN, D>(L, S::NTV + f * S::FT + tt, lane, xo);
      float dn = 0.f;
#pragma unroll
      for (int d = 0; d < D; ++d) dn += c[f][d] * y[1 + d];
      const int g = 4 * tt + q;
      if (valid) trace[(el * M::NF + f) * M::SG + g] = make_float2(y[0], dn);
    }
  }
}

// Residual pass.  MODE 0: r = -A x; 1: r = b - A x; 2: x += S (b - A x); 3: as 2 and
// store r; 4: r = +A x; 5: x = (1 - omega) x + S (b - A_off x) (exact block smoother,
// neighbour terms only: no volume or own-trace contractions).  b and r may alias.  Every per-element operand and the
// neighbour traces are requested in the prologue, so they are in flight during the
// volume contractions.
template <typename TIN, int D, int MODE, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_residual_mf(
    int nele, float* __restrict__ x, const float* b, float* r, MfArgs ta,
    const float* __restrict__ wv, const float* __restrict__ G, const float* __restrict__ cf,
    const float* __restrict__ wf, const float* __restrict__ pen,
    const int32_t* __restrict__ nbface, const int32_t* __restrict__ fcode,
    const float2* __restrict__ trace, const float* __restrict__ Sm,
    const int32_t* __restrict__ act, float omega) {
  using M = Mq<D>;
  using S = MfShape<D>;
  constexpr int NL = M::NL, NF = M::NF, FT = S::FT;
  constexpr bool OFFD = MODE == 5;
  // mode 5 lifts no volume tile; the bf16 lift pairs tiles, so an odd volume count keeps
  // zero volume tiles for the pairing
  constexpr bool SKIP_VOL = OFFD && (std::is_same<TIN, float>::value || S::NTV % 2 == 0);
  __shared__ MfLds<TIN, D> L;
  const int grp = act ? (int)blockIdx.x : xcd_group(blockIdx.x, gridDim.x);
  if (act && !mf_block_active<WPB>(act, nele, grp)) return;  // before any barrier
  __shared__ float rx[(MODE == 2 || MODE == 3 || OFFD) ? WPB : 1][16][NL + 1];
  const int lane = threadIdx.x & 63, q = lane >> 4, wl = threadIdx.x >> 6;
  const int wave = grp * WPB + wl;
  const int el = wave * 16 + (lane & 15);
  const bool valid = el < nele;
  const int ec = valid ? el : nele - 1;
  // ---- prologue: element operands
  typename MfLds<TIN, D>::XOp xo{};
  float gm[M::NG] = {};
  if constexpr (!OFFD) {
    xo = MfLds<TIN, D>::load_x(x, ec, q);
#pragma unroll
    for (int k = 0; k < M::NG; ++k) gm[k] = G[k * nele + ec];
  }
  int code[NF], nbz[NF];
  float c[NF][D], wfe[NF], pe[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    code[f] = fcode[ec * NF + f];
    nbz[f] = nbface[ec * NF + f];
#pragma unroll
    for (int d = 0; d < D; ++d) c[f][d] = cf[(f * D + d) * nele + ec];
    wfe[f] = wf[f * nele + ec];
    pe[f] = pen[f * nele + ec];
  }
  mf_load_tables<TIN, D, 64 * WPB, false>(L, ta);
  __syncthreads();
  // neighbour traces at the aligned points of this lane's face points
  float2 tn[NF][FT];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int type = code[f] & 15, a = (code[f] >> 4) & 15;
#pragma unroll
    for (int tt = 0; tt < FT; ++tt) {
      tn[f][tt] = make_float2(0.f, 0.f);
      if (type == 0) tn[f][tt] = trace[nbz[f] * M::SG + L.aln[a * M::SG + 4 * tt + q]];
    }
  }
  Lift<TIN, D> lift;
  lift.init();
  // ---- volume points
#pragma unroll
  for (int t = 0; t < (SKIP_VOL ? 0 : S::NTV); ++t) {
    v4f z = {0.f, 0.f, 0.f, 0.f};
    if constexpr (OFFD) {
      mf_lift<TIN, D>(lift, L, t, lane, z);
      continue;
    }
    const v4f y = mf_tile<TIN, D>(L, t, lane, xo);
    const float wq = wv[4 * t + q];
    if constexpr (D == 3) {
      z[0] = wq * (gm[0] * y[0] + gm[3] * y[1] + gm[4] * y[2]);
      z[1] = wq * (gm[3] * y[0] + gm[1] * y[1] + gm[5] * y[2]);
      z[2] = wq * (gm[4] * y[0] + gm[5] * y[1] + gm[2] * y[2]);
    } else {
      z[0] = wq * (gm[0] * y[0] + gm[2] * y[1]);
      z[1] = wq * (gm[2] * y[0] + gm[1] * y[1]);
      z[2] = 0.f;
    }
    z[3] = 0.f;
    mf_lift<TIN, D>(lift, L, t, lane, z);
  }
  // ---- faces
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int type = code[f] & 15;
#pragma unroll
    for (int tt = 0; tt < FT; ++tt) {
      const int t = S::NTV + f * FT + tt;
      const float ws = wfe[f] * wv[M::NGI + 4 * tt + q];
      float qsn = 0.f, qdn = 0.f;
      v4f y = {0.f, 0.f, 0.f, 0.f};
      float dn = 0.f;
      if constexpr (!OFFD) {
        y = mf_tile<TIN, D>(L, t, lane, xo);
#pragma unroll
        for (int d = 0; d < D; ++d) dn += c[f][d] * y[1 + d];
      }
      if (OFFD) {
        if (type == 0) {
          const float2 nb = tn[f][tt];
          qsn = ws * (0.5f * nb.y - pe[f] * nb.x);
          qdn = ws * (0.5f * nb.x);
        }
      } else if (type == 0) {
        const float2 nb = tn[f][tt];
        const float jump = y[0] - nb.x;
        qsn = ws * (-0.5f * (dn - nb.y) + pe[f] * jump);
        qdn = ws * (-0.5f * jump);
      } else if (type == 1) {
        qsn = ws * (pe[f] * y[0] - dn);
        qdn = -ws * y[0];
      }
      v4f z;
      z[0] = qsn;
#pragma unroll
      for (int d = 0; d < D; ++d) z[1 + d] = qdn * c[f][d];
#pragma unroll
      for (int d = D; d < 3; ++d) z[1 + d] = 0.f;
      mf_lift<TIN, D>(lift, L, t, lane, z);
    }
  }
  // ---- lane holds (A x)[16 mt + 4q + i] of element el
  if constexpr (MODE <= 1 || MODE == 4) {
    if (valid) {
#pragma unroll
      for (int mt = 0; mt < S::MT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int n = 16 * mt + 4 * q + i;
          if (n < NL) {
            const size_t k = (size_t)el * NL + n;
            float v = lift.acc[mt][i];
            if constexpr (MODE == 0) v = -v;
            else if constexpr (MODE == 1) v = b[k] - v;
            r[k] = v;
          }
        }
    }
  } else {
    const int le = lane & 15;
#pragma unroll
    for (int mt = 0; mt < S::MT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int n = 16 * mt + 4 * q + i;
        if (n < NL) {
          const float rv = b[(size_t)ec * NL + n] - lift.acc[mt][i];
          rx[wl][le][n] = rv;
          if constexpr (MODE == 3)
            if (valid) r[(size_t)el * NL + n] = rv;
        }
      }
    __syncthreads();
    // the four lanes of an element take NL/4 consecutive rows each (the whole residual
    // row is in rx), so every lane issues the same number of smoother loads
    if (valid) {
      constexpr int RPL = (NL + 3) / 4;
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int n = RPL * q + i;
        if (n < NL) {
          float srow[NL];
          load_S_row_sym<NL>(Sm, n, el, nele, srow);
          const size_t k = (size_t)el * NL + n;
          float s = OFFD ? (1.0f - omega) * x[k] : x[k];
#pragma unroll
          for (int m = 0; m < NL; ++m) s += srow[m] * rx[wl][le][m];
          x[k] = s;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Host: operand tables in lane order.
static uint16_t to_bf16_bits(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;  // round to nearest even (tables are finite)
  return (uint16_t)(u >> 16);
}

int build_mfma_tables(dgmi_ctx* c) {
  const ElementTables& el = c->el;
  const int D = el.D, NL = el.nloc, NGI = el.ngi, SG = el.sngi, NF = el.nface;
  const int R = 4 * (NGI + NF * SG), NT = R / 16;
  const int MT = (NL + 15) / 16;
  // T[row][n]
  std::vector<double> T((size_t)R * NL, 0.0);
  for (int g = 0; g < NGI; ++g)
    for (int d = 0; d < D; ++d)
      for (int n = 0; n < NL; ++n) T[(size_t)(4 * g + d) * NL + n] = el.nlx[(d * NL + n) * NGI + g];
  for (int f = 0; f < NF; ++f)
    for (int g = 0; g < SG; ++g) {
      const int row = 4 * (NGI + f * SG + g);
      for (int n = 0; n < NL; ++n) {
        T[(size_t)row * NL + n] = el.sn[(f * NL + n) * SG + g];
        for (int d = 0; d < D; ++d)
          T[(size_t)(row + 1 + d) * NL + n] = el.snlx[((f * D + d) * NL + n) * SG + g];
      }
    }
  auto Tv = [&](int row, int n) -> double { return n < NL ? T[(size_t)row * NL + n] : 0.0; };
  std::vector<uint16_t> b1((size_t)NT * 64 * 8), b2((size_t)(NT / 2) * MT * 64 * 8);
  for (int t = 0; t < NT; ++t)
    for (int l = 0; l < 64; ++l)
      for (int j = 0; j < 8; ++j)
        b1[((size_t)t * 64 + l) * 8 + j] = to_bf16_bits((float)Tv(16 * t + l % 16, 8 * (l / 16) + j));
  for (int tp = 0; tp < NT / 2; ++tp)
    for (int mt = 0; mt < MT; ++mt)
      for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 8; ++j) {
          const int row = j < 4 ? 16 * (2 * tp) + 4 * (l / 16) + j : 16 * (2 * tp + 1) + 4 * (l / 16) + (j - 4);
          b2[(((size_t)tp * MT + mt) * 64 + l) * 8 + j] = to_bf16_bits((float)Tv(row, 16 * mt + l % 16));
        }
  std::vector<float> tT((size_t)R * NL + 16, 0.0f);
  for (size_t i = 0; i < (size_t)R * NL; ++i) tT[i] = (float)T[i];
  std::vector<float> wv(NGI + SG);
  for (int g = 0; g < NGI; ++g) wv[g] = (float)el.weight[g];
  for (int g = 0; g < SG; ++g) wv[NGI + g] = (float)el.sweight[g];
  auto up = [&](const void* h, size_t bytes) -> void* {
    void* d = c->alloc(bytes);
    if (!d) return nullptr;
    if (h2d(c, d, h, bytes, c->ustream)) return nullptr;
    return d;
  };
  c->mf.b1 = up(b1.data(), b1.size() * 2);
  c->mf.b2 = up(b2.data(), b2.size() * 2);
  c->mf.wv = (float*)up(wv.data(), wv.size() * 4);
  c->mf.tT = (float*)up(tT.data(), tT.size() * 4);
  if (!c->mf.b1 || !c->mf.b2 || !c->mf.wv || !c->mf.tT) {
    set_error(alloc_failure());
    return -3;
  }
  return 0;
}

template <typename TIN, int D, int WPB>
static int apply_mf_impl(dgmi_ctx* c, float* x, const float* b, float* r, int mode,
                         hipStream_t s) {
  const int nele = (int)c->nele;
  DevGeom<float>& g = c->g32;
  const MfArgs ta{c->mf.tT, c->mf.b1, c->mf.b2, c->aln};
  const dim3 blk(64 * WPB), grd((nele + 16 * WPB - 1) / (16 * WPB));
  const int32_t* act = c->mask_on ? c->act64 : nullptr;
  if (!ensure_traces(c, g)) return -3;
  hipLaunchKernelGGL((k_trace_mf<TIN, D, WPB>), grd, blk, 0, s, nele, x, ta, g.cf,
                     (float2*)g.trace, act);
#define DGMI_RES_MF(MODE)                                                                      \
  hipLaunchKernelGGL((k_residual_mf<TIN, D, MODE, WPB>), grd, blk, 0, s, nele, x, b, r, ta,    \
                     c->mf.wv, g.G, g.cf, g.wf, g.pen, c->nbface, c->fcode,                     \
                     (const float2*)g.trace, g.S, act, (float)c->omega)
  switch (mode) {
    case 0: DGMI_RES_MF(0); break;
    case 1: DGMI_RES_MF(1); break;
    case 2: DGMI_RES_MF(2); break;
    case 3: DGMI_RES_MF(3); break;
    case 5: DGMI_RES_MF(5); break;
    default: DGMI_RES_MF(4); break;
  }
#undef DGMI_RES_MF
  DGMI_HIP(hipGetLastError());
  return 0;
}

template <typename TIN, int D>
static int apply_mf_wpb(dgmi_ctx* c, float* x, const float* b, float* r, int mode,
                        hipStream_t s) {
  const int wpb = std::is_same<TIN, float>::value ? c->mf_wpb : c->bf_wpb;
  if (wpb == 8) return apply_mf_impl<TIN, D, 8>(c, x, b, r, mode, s);
  if (wpb == 2) return apply_mf_impl<TIN, D, 2>(c, x, b, r, mode, s);
  return apply_mf_impl<TIN, D, 4>(c, x, b, r, mode, s);
}

// mode as k_residual_mf; bf16 selects the BF16-input variant
int launch_apply_mf(dgmi_ctx* c, bool bf16, float* x, const float* b, float* r, int mode,
                    hipStream_t s) {
  if (c->nele == 0) return 0;
  if (bf16) {
    if (c->D == 3) return apply_mf_wpb<__bf16, 3>(c, x, b, r, mode, s);
    return apply_mf_wpb<__bf16, 2>(c, x, b, r, mode, s);
  }
  if (c->D == 3) return apply_mf_wpb<float, 3>(c, x, b, r, mode, s);
  return apply_mf_wpb<float, 2>(c, x, b, r, mode, s);
}

}  // namespace dgmi
