// Row-vector helpers shared by the apply kernels: 16- or 8-byte loads/stores of one
// element's DOF row and of rows of the per-element smoother matrices.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace dgmi {

// Work-groups are dispatched to the 8 XCDs round-robin by id (group b runs on XCD b % 8),
// and each XCD has its own L2.  The element kernels read their neighbours' face traces,
// and neighbours are close in the element (SFC) order, so each XCD is given one
// contiguous range of element groups: group b takes logical group
// (first group of XCD b % 8) + b / 8.  A bijection on [0, n) for any n.  Kernels that skip
// the groups of finished solves (packed contexts) keep the identity: there the remaining
// work would gather on the XCDs that hold the slow solves.
__device__ __forceinline__ int xcd_group(int b, int n) {
  constexpr int NX = 8;
  const int x = b % NX, q = n / NX, r = n % NX;
  return x * q + (x < r ? x : r) + b / NX;
}

// Vector loads/stores of one element's row: 16-byte accesses when the row length
// allows (3D P3 in both precisions, 2D in FP64), 8-byte otherwise (2D P3 in FP32).
// DOF vectors are required to be 16-byte aligned.
template <typename T, int NL> struct RowVec {
  static constexpr int W = (NL % (16 / sizeof(T)) == 0) ? 16 / sizeof(T) : 8 / sizeof(T);
  using V = typename std::conditional<
      sizeof(T) * W == 16, typename std::conditional<sizeof(T) == 8, double2, float4>::type,
      typename std::conditional<sizeof(T) == 8, double, float2>::type>::type;
  static_assert(NL % W == 0, "row length not a multiple of the vector width");
};

template <typename T, int NL>
__device__ __forceinline__ void load_row(const T* __restrict__ p, T* v) {
  using RV = RowVec<T, NL>;
  constexpr int W = RV::W;
  const typename RV::V* q = reinterpret_cast<const typename RV::V*>(p);
#pragma unroll
  for (int i = 0; i < NL / W; ++i) {
    typename RV::V t = q[i];
    const T* tt = reinterpret_cast<const T*>(&t);
#pragma unroll
    for (int k = 0; k < W; ++k) v[i * W + k] = tt[k];
  }
}
template <typename T, int NL>
__device__ __forceinline__ void store_row(T* __restrict__ p, const T* v) {
  using RV = RowVec<T, NL>;
  constexpr int W = RV::W;
  typename RV::V* q = reinterpret_cast<typename RV::V*>(p);
#pragma unroll
  for (int i = 0; i < NL / W; ++i) {
    typename RV::V t;
    T* tt = reinterpret_cast<T*>(&t);
#pragma unroll
    for (int k = 0; k < W; ++k) tt[k] = v[i * W + k];
    q[i] = t;
  }
}

// Smoother matrices are stored as structure-of-vectors: entry S[n][m] of element e lives in
// vector ((n*(NL/W) + m/W)*nele + e), component m%W, with W the row vector width above, so
// one wave-wide load moves 64 x 16 bytes.
template <typename T, int NL>
__host__ __device__ inline size_t s_index(int n, int m, int e, int nele) {
  constexpr int W = RowVec<T, NL>::W;
  return ((size_t)(n * (NL / W) + m / W) * nele + e) * W + (m % W);
}

// row n of S for element e
template <typename T, int NL>
__device__ __forceinline__ void load_S_row(const T* __restrict__ S, int n, int e, int nele, T* v) {
  using RV = RowVec<T, NL>;
  constexpr int W = RV::W;
  const typename RV::V* q = reinterpret_cast<const typename RV::V*>(S);
#pragma unroll
  for (int i = 0; i < NL / W; ++i) {
    typename RV::V t = q[(size_t)(n * (NL / W) + i) * nele + e];
    const T* tt = reinterpret_cast<const T*>(&t);
#pragma unroll
    for (int k = 0; k < W; ++k) v[i * W + k] = tt[k];
  }
}

// Smoother matrices are stored packed: both smoothers are symmetric, so entry
// (a, b), a <= b, of element e at sym_index(a, b)*nele + e carries S[a][b] and S[b][a]
// (NL(NL+1)/2 vectors of nele values instead of NL*NL).
template <int NL>
__host__ __device__ constexpr int sym_index(int a, int b) {
  return a * NL - a * (a - 1) / 2 + (b - a);
}
template <int NL>
constexpr int sym_size() {
  return NL * (NL + 1) / 2;
}

template <int NL, typename T>
__device__ __forceinline__ void load_S_row_sym(const T* __restrict__ S, int n, int e, int nele,
                                               T* v) {
#pragma unroll
  for (int m = 0; m < NL; ++m) {
    const int a = m < n ? m : n, b = m < n ? n : m;
    v[m] = S[(size_t)sym_index<NL>(a, b) * nele + e];
  }
}

// bfloat16 storage (the smoother copy of the bf16 inner cycles): round to nearest even on
// the way in, exact widening on the way out.  Inputs are finite.
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint32_t h) {
  const uint32_t u = h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// face trace (u, du/dn) of the vector kernels
__device__ __forceinline__ float2 trace_value(const float2& t) { return t; }
__device__ __forceinline__ double2 trace_value(const double2& t) { return t; }
template <typename TR, typename T>
__device__ __forceinline__ TR make_trace(T u, T dn) {
  TR t;
  t.x = u;
  t.y = dn;
  return t;
}

// row n of a packed symmetric smoother stored in bfloat16
template <int NL>
__device__ __forceinline__ void load_S_row_sym16(const uint16_t* __restrict__ S, int n, int e,
                                                 int nele, float* v) {
#pragma unroll
  for (int m = 0; m < NL; ++m) {
    const int a = m < n ? m : n, b = m < n ? n : m;
    v[m] = bf16_to_f32(S[(size_t)sym_index<NL>(a, b) * nele + e]);
  }
}

// row n of the packed symmetric smoother of element e (FP64, FP32 or bfloat16 storage)
template <typename T, int NL, typename SS = T>
__device__ __forceinline__ void load_smoother_row(const SS* __restrict__ S, int n, int e, int nele,
                                                  T* v) {
  if constexpr (std::is_same<SS, uint16_t>::value) load_S_row_sym16<NL>(S, n, e, nele, v);
  else load_S_row_sym<NL, T>(S, n, e, nele, v);
}

}  // namespace dgmi
