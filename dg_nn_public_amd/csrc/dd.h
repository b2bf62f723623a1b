// Double-double arithmetic for the compensated residual.
//
// A value is the unevaluated sum hi + lo of two doubles with |lo| <= ulp(hi)/2, about 106
// significant bits.  Built from error-free transformations:
//   TwoSum      s + e = a + b exactly (Knuth, TAOCP vol. 2, sec. 4.2.2, theorem B)
//   FastTwoSum  the same when |a| >= |b| (Dekker, Numer. Math. 18, 1971)
//   TwoProd     p + e = a * b exactly, e from one fused multiply-add
// and the accurate ("IEEE") double-double addition of Hida, Li & Bailey, "Algorithms for
// quad-double precision floating point arithmetic", ARITH-15 (2001).  Contraction of the
// plain operations into fused multiply-adds would break the transformations, so every
// function turns it off; the intended fused operations are written with fma().
#pragma once
#include <hip/hip_runtime.h>

namespace dgmi {

struct dd {
  double hi, lo;
  dd() = default;
  __host__ __device__ constexpr dd(double v) : hi(v), lo(0.0) {}
  __host__ __device__ constexpr dd(double h, double l) : hi(h), lo(l) {}
};

__device__ __forceinline__ dd two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  const double bb = s - a;
  return dd(s, (a - (s - bb)) + (b - bb));
}

__device__ __forceinline__ dd fast_two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  return dd(s, b - (s - a));
}

__device__ __forceinline__ dd operator+(dd a, dd b) {
#pragma clang fp contract(off)
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ dd operator-(dd a) { return dd(-a.hi, -a.lo); }
__device__ __forceinline__ dd operator-(dd a, dd b) { return a + (-b); }

__device__ __forceinline__ dd operator*(dd a, dd b) {
#pragma clang fp contract(off)
  const double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e = fma(a.hi, b.lo, e);
  e = fma(a.lo, b.hi, e);
  return fast_two_sum(p, e);
}

__device__ __forceinline__ dd operator*(double a, dd b) {
#pragma clang fp contract(off)
  const double p = a * b.hi;
  double e = fma(a, b.hi, -p);
  e = fma(a, b.lo, e);
  return fast_two_sum(p, e);
}

__device__ __forceinline__ dd operator*(dd a, double b) { return b * a; }

// a * b exactly (TwoProd)
__device__ __forceinline__ dd two_prod(double a, double b) {
#pragma clang fp contract(off)
  const double p = a * b;
  return dd(p, fma(a, b, -p));
}
__device__ __forceinline__ dd& operator+=(dd& a, dd b) { return a = a + b; }

}  // namespace dgmi
