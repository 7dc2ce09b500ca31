// dd.hpp -- double-double arithmetic (Dekker / Knuth error-free transforms).
// Used for the prefix sums that give every quadtree node its mass-weighted
// position sum: with ~106-bit partial sums the difference of two prefixes is
// accurate to FP64 round-off even deep in a million-point array.
// Translation units that include this must not contract a*b+c into FMA
// (-ffp-contract=off); two_prod below calls fma() explicitly.
#pragma once

#include <hip/hip_runtime.h>

namespace gvh {

struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__device__ __forceinline__ dd quick_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}

__device__ __forceinline__ dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  double e = s.lo + a.lo + b.lo;
  return quick_two_sum(s.hi, e);
}

__device__ __forceinline__ dd dd_neg(dd a) { return {-a.hi, -a.lo}; }

__device__ __forceinline__ dd two_prod(double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}

__device__ __forceinline__ dd dd_shfl_xor(dd v, int mask) {
  return {__shfl_xor(v.hi, mask, 64), __shfl_xor(v.lo, mask, 64)};
}

} // namespace gvh
