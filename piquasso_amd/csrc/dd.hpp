// Double-double ("dd") arithmetic for the compensated accumulation of the
// Glynn sum.  A dd value is an unevaluated sum hi + lo with |lo| <= ulp(hi)/2.
//
// Algorithms (all error-free transformations, exact in IEEE FP64 with FMA):
//   TwoSum       Knuth, TAOCP vol. 2, sec. 4.2.2 (6 flops)
//   FastTwoSum   Dekker 1971 (requires |a| >= |b|, 3 flops)
//   TwoProd      a*b = p + e with e = fma(a, b, -p) (Ogita, Rump, Oishi,
//                "Accurate sum and dot product", SIAM J. Sci. Comput. 26, 2005)
// Only sums and products by exact integers are needed here: the kernels add
// FP64 run sums into dd accumulators, scale them by integer weights and
// reduce dd values across lanes, waves and work units.
#pragma once

#include <hip/hip_runtime.h>

#define PQP_HD __host__ __device__ __forceinline__

namespace pqp {

struct dd {
  double hi, lo;
};

struct cdd {
  dd re, im;
};

PQP_HD dd dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  return dd{s, (a - (s - bb)) + (b - bb)};
}

PQP_HD dd dd_fast_two_sum(double a, double b) {
  double s = a + b;
  return dd{s, b - (s - a)};
}

// dd + double
PQP_HD dd dd_add_d(dd x, double d) {
  dd s = dd_two_sum(x.hi, d);
  s.lo += x.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}

// dd + dd (the accurate variant: both high and low parts through TwoSum)
PQP_HD dd dd_add(dd x, dd y) {
  dd s = dd_two_sum(x.hi, y.hi);
  dd t = dd_two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = dd_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}

// dd * double (TwoProd on the high part, low part folded in)
PQP_HD dd dd_mul_d(dd x, double d) {
  double p = x.hi * d;
  double e = __builtin_fma(x.hi, d, -p);
  e = __builtin_fma(x.lo, d, e);
  return dd_fast_two_sum(p, e);
}

PQP_HD cdd cdd_add(cdd x, cdd y) { return cdd{dd_add(x.re, y.re), dd_add(x.im, y.im)}; }

PQP_HD cdd cdd_scale(cdd x, double d) { return cdd{dd_mul_d(x.re, d), dd_mul_d(x.im, d)}; }

PQP_HD cdd cdd_zero() { return cdd{{0.0, 0.0}, {0.0, 0.0}}; }

}  // namespace pqp
