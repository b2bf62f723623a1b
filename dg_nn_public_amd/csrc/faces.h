// Face traces and face lifts of the vector kernels through a quadratic interpolant.
//
// On face f of a P3 element only the face's own nodes (10 of 20 in 3D, 4 of 10 in 2D)
// have nonzero basis functions, so the trace u and the lift of the flux term onto u's
// test functions involve those nodes only.  The normal derivative c . grad(u)
// (c = J^-T n_f, constant on an affine element) restricted to the face is a quadratic
// polynomial, fixed by its values at the face's quadratic nodes (3 vertices and 3 edge
// midpoints in 3D, 2 end points and the midpoint in 2D):
//   trace:  dn(g) = sum_k L_k(g) c . grad(u)(p_k)                 (instead of per point)
//   lift:   sum_g q(g) c . grad(phi_n)(g) = sum_k (sum_g L_k(g) q(g)) c . grad(phi_n)(p_k)
// Per face in 3D this takes 6 x 63 + 12 x 16 = 570 FMAs for the traces and the same for
// the lift, against 12 x 83 and 12 x 80 point by point.  The tables are built and checked
// against the tabulated face gradients on the host (exact to rounding for P3).
#pragma once
#include <hip/hip_runtime.h>

#include "rowvec.h"

namespace dgmi {

// Layout of the face-interpolation block (DevGeom::off_fc), in the table's precision:
//   PHI [f][g][NFN]     phi of the face's nodes at its points
//   LK  [g][NK]         quadratic face Lagrange basis at the face points
//   GK  [f][k][D][NL]   d phi / d xi of every basis function at the face's quadratic nodes
//   MF  [f][i][j]       sum_g sw_g phi_i(g) phi_j(g) over the face's nodes (face mass)
//   ML  [f][i][k]       sum_g sw_g phi_i(g) L_k(g)
//   VD  [d][i][NL]      d phi / d xi_d at the element's P2 nodes (NP2 of them)
//   VM  [i][j]          sum_g w_g psi_i(g) psi_j(g), psi the P2 Lagrange basis
// and the int32 node lists fnodes[f][NFN].
template <int D> struct FaceDims {
  static constexpr int NL = D == 3 ? 20 : 10, SG = D == 3 ? 12 : 4, NF = D + 1;
  static constexpr int NFN = D == 3 ? 10 : 4, NK = D == 3 ? 6 : 3, NP2 = D == 3 ? 10 : 6;
  static constexpr int PHI = 0, LK = NF * SG * NFN, GK = LK + SG * NK;
  static constexpr int MF = GK + NF * NK * D * NL, ML = MF + NF * NFN * NFN;
  static constexpr int VD = ML + NF * NFN * NK, VM = VD + D * NP2 * NL;
};

// The nodes of each face (ascending local node numbers; the host checks them against the
// element tables' fnodes), for kernels that index register rows by them at compile time.
template <int D> struct FaceNodes;
template <> struct FaceNodes<3> {
  static constexpr int v[4][10] = {{1, 2, 3, 8, 9, 12, 13, 14, 15, 16},
                                   {0, 2, 3, 4, 5, 10, 11, 14, 15, 18},
                                   {0, 1, 3, 6, 7, 10, 11, 12, 13, 19},
                                   {0, 1, 2, 4, 5, 6, 7, 8, 9, 17}};
};
template <> struct FaceNodes<2> {
  static constexpr int v[3][4] = {{0, 1, 3, 4}, {1, 2, 5, 6}, {0, 2, 7, 8}};
};

// The P2 node (vertices, then edge midpoints (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) in 3D,
// (0,1) (0,2) (1,2) in 2D) at quadratic node k of face f: the face-gradient table GK[f][k]
// equals the volume table VD at that node (the host checks it), so the lifts of all faces
// can be summed at the P2 nodes and applied once.
template <int D> struct FaceP2;
template <> struct FaceP2<3> {
  static constexpr int p[4][6] = {{2, 1, 3, 7, 8, 9}, {0, 2, 3, 5, 9, 6}, {1, 0, 3, 4, 6, 8},
                                  {0, 1, 2, 4, 7, 5}};
};
template <> struct FaceP2<2> {
  static constexpr int p[3][3] = {{0, 1, 3}, {1, 2, 5}, {2, 0, 4}};
};

// The nodes off face f (ascending), and per node the faces it lies on (bit f)
template <int D> struct FaceOff {
  static constexpr int NL = D == 3 ? 20 : 10, NF = D + 1, NFN = D == 3 ? 10 : 4;
  static constexpr int NO = NL - NFN;
  struct Tab {
    int off[NF][NO];
    int mask[NL];
  };
  static constexpr Tab make() {
    Tab t{};
    for (int n = 0; n < NL; ++n) t.mask[n] = 0;
    for (int f = 0; f < NF; ++f) {
      bool on[NL] = {};
      for (int i = 0; i < NFN; ++i) {
        on[FaceNodes<D>::v[f][i]] = true;
        t.mask[FaceNodes<D>::v[f][i]] |= 1 << f;
      }
      int j = 0;
      for (int n = 0; n < NL; ++n)
        if (!on[n]) t.off[f][j++] = n;
    }
    return t;
  }
  static constexpr Tab tab = make();
};

// c . grad(u) at the NK quadratic nodes of face f (c = J^-T n_f), from the row xv
template <typename T, int D, typename TT>
__device__ __forceinline__ void face_dk(int f, const TT* __restrict__ fc, const T* xv,
                                        const T* c, T* dk) {
  using F = FaceDims<D>;
  const TT* gk = fc + F::GK + f * F::NK * D * F::NL;
#pragma unroll 2
  for (int k = 0; k < F::NK; ++k) {
    T gd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) gd[d] = 0;
#pragma unroll
    for (int n = 0; n < F::NL; ++n)
#pragma unroll
      for (int d = 0; d < D; ++d) gd[d] += gk[(k * D + d) * F::NL + n] * xv[n];
    T s = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) s += c[d] * gd[d];
    dk[k] = s;
  }
}

// Traces (u, du/dn) of face f at its SG points: xv = the element's row, xf = the values of
// the face's nodes (fnodes order), c = J^-T n_f; store(g, u, dn) writes point g.
// The tables (fc) may be held in another type TT than the arithmetic T.
template <typename T, int D, typename TT, typename Store>
__device__ __forceinline__ void face_traces(int f, const TT* __restrict__ fc, const T* xv,
                                            const T* xf, const T* c, Store store) {
  using F = FaceDims<D>;
  T dk[F::NK];
  face_dk<T, D>(f, fc, xv, c, dk);
  const TT* ph = fc + F::PHI + f * F::SG * F::NFN;
  const TT* lk = fc + F::LK;
#pragma unroll
  for (int g = 0; g < F::SG; ++g) {
    T u = 0, dn = 0;
#pragma unroll
    for (int i = 0; i < F::NFN; ++i) u += ph[g * F::NFN + i] * xf[i];
#pragma unroll
    for (int k = 0; k < F::NK; ++k) dn += lk[g * F::NK + k] * dk[k];
    store(g, u, dn);
  }
}

// Accumulators of one face's lift: accf[i] += phi_i(g) qs(g) for the face's nodes, and
// Q[k] += L_k(g) qd(g) for the normal-derivative part.
template <typename T, int D> struct FaceLift {
  using F = FaceDims<D>;
  T accf[F::NFN];
  T Q[F::NK];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < F::NFN; ++i) accf[i] = 0;
#pragma unroll
    for (int k = 0; k < F::NK; ++k) Q[k] = 0;
  }
  template <typename TT>
  __device__ __forceinline__ void add(int f, int g, const TT* __restrict__ fc, T qs, T qd) {
    const TT* ph = fc + F::PHI + (f * F::SG + g) * F::NFN;
    const TT* lk = fc + F::LK + g * F::NK;
#pragma unroll
    for (int i = 0; i < F::NFN; ++i) accf[i] += ph[i] * qs;
#pragma unroll
    for (int k = 0; k < F::NK; ++k) Q[k] += lk[k] * qd;
  }
  // The same sums in node form, for fluxes that are linear in the face values:
  //   qs(g) = sw_g (sv v(g) + sd dv(g)),  qd(g) = sw_g sq v(g),
  // with v given by its values at the face's nodes (v(g) = sum_j phi_j(g) v_j) and dv by
  // its values at the quadratic nodes (dv(g) = sum_k L_k(g) dv_k):
  //   accf = sv MF v + sd ML dv,  Q = sq ML^T v.
  template <typename TT>
  __device__ __forceinline__ void set_nodes(int f, const TT* __restrict__ fc, const T* v,
                                            const T* dv, T sv, T sd, T sq) {
    const TT* mf = fc + F::MF + f * F::NFN * F::NFN;
    const TT* ml = fc + F::ML + f * F::NFN * F::NK;
#pragma unroll
    for (int k = 0; k < F::NK; ++k) Q[k] = 0;
#pragma unroll
    for (int i = 0; i < F::NFN; ++i) {
      T a = 0, b = 0;
#pragma unroll
      for (int j = 0; j < F::NFN; ++j) a += mf[i * F::NFN + j] * v[j];
#pragma unroll
      for (int k = 0; k < F::NK; ++k) {
        b += ml[i * F::NK + k] * dv[k];
        Q[k] += ml[i * F::NK + k] * v[i];
      }
      accf[i] = sv * a + sd * b;
    }
#pragma unroll
    for (int k = 0; k < F::NK; ++k) Q[k] *= sq;
  }
  // acc[n] += sum_{k,d} c_d Q_k d phi_n / d xi_d (p_k), over all nodes
  template <typename TT>
  __device__ __forceinline__ void finish(int f, const TT* __restrict__ fc, const T* c,
                                         T* acc) const {
    const TT* gk = fc + F::GK + f * F::NK * D * F::NL;
    T W[F::NK][D];
#pragma unroll
    for (int k = 0; k < F::NK; ++k)
#pragma unroll
      for (int d = 0; d < D; ++d) W[k][d] = c[d] * Q[k];
#pragma unroll
    for (int k = 0; k < F::NK; ++k)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const TT* t = gk + (k * D + d) * F::NL;
#pragma unroll
        for (int n = 0; n < F::NL; ++n) acc[n] += t[n] * W[k][d];
      }
  }
};

}  // namespace dgmi
