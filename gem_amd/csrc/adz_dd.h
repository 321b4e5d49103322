// Double-double accumulation shared by the fixer's global sums (gfx950).
//
// Each lane adds its terms into a (hi, lo) pair with TwoSum, so hi + lo carries the
// running sum to about twice FP64 precision; partial pairs are merged in a fixed tree and
// the final hi + lo rounds to the exactly rounded sum except in astronomically rare ties.
// The result therefore does not depend on how a domain is split among blocks.
#pragma once
#include "adz_common.h"

namespace adz {

struct DD {
  double hi, lo;
};
__device__ __forceinline__ void two_sum(double a, double b, double* s, double* e) {
  double x = a + b;
  double bp = x - a;
  *e = (a - (x - bp)) + (b - bp);
  *s = x;
}
__device__ __forceinline__ void dd_add(DD* a, double t) {
  double s, e;
  two_sum(a->hi, t, &s, &e);
  a->hi = s;
  a->lo += e;
}
__device__ __forceinline__ DD dd_merge(DD a, DD b) {
  double s, e;
  two_sum(a.hi, b.hi, &s, &e);
  e += a.lo + b.lo;
  DD r;
  two_sum(s, e, &r.hi, &r.lo);
  return r;
}

// One term of a tracer mass (mass_tr_X.F90:86-93, mass_tr_PM.F90:93-94,119-120): the
// tracer times the air mass is a `real` product when both are `real`, then times the
// FP64 geomh_mask_8.
template <class T>
__device__ __forceinline__ double mass_term(T x, T air, double mask) {
  T prod = x * air;
  return (double)prod * mask;
}

}  // namespace adz
