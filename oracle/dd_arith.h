/*
 * ORACLE — test infrastructure, not product code.
 *
 * Double-double arithmetic shared by the oracle's extended-precision sums:
 * TwoSum (Knuth, TAOCP vol. 2, 4.2.2), FastTwoSum and FMA-based TwoProd
 * (Dekker 1971; Ogita, Rump, Oishi, SIAM J. Sci. Comput. 26, 2005).
 * Relative accuracy of each operation is about 2^-104.
 */
#ifndef PQP_ORACLE_DD_ARITH_H
#define PQP_ORACLE_DD_ARITH_H

#include <math.h>

typedef struct {
  double hi, lo;
} dd_t;

static inline dd_t dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  dd_t r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}

static inline dd_t dd_fast_two_sum(double a, double b) {
  double s = a + b;
  dd_t r = {s, b - (s - a)};
  return r;
}

static inline dd_t dd_add(dd_t x, dd_t y) {
  dd_t s = dd_two_sum(x.hi, y.hi);
  dd_t t = dd_two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = dd_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}

static inline dd_t dd_add_d(dd_t x, double d) {
  dd_t s = dd_two_sum(x.hi, d);
  s.lo += x.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}

static inline dd_t dd_neg(dd_t x) {
  dd_t r = {-x.hi, -x.lo};
  return r;
}

static inline dd_t dd_mul(dd_t x, dd_t y) {
  double p = x.hi * y.hi;
  double e = fma(x.hi, y.hi, -p);
  e += x.hi * y.lo + x.lo * y.hi;
  return dd_fast_two_sum(p, e);
}

static inline dd_t dd_mul_d(dd_t x, double d) {
  double p = x.hi * d;
  double e = fma(x.hi, d, -p);
  e += x.lo * d;
  return dd_fast_two_sum(p, e);
}

static inline dd_t dd_from_prod(double a, double b) {
  double p = a * b;
  dd_t r = {p, fma(a, b, -p)};
  return r;
}

typedef struct {
  dd_t re, im;
} cdd_t;

static inline cdd_t cdd_mul(cdd_t x, cdd_t y) {
  cdd_t r;
  r.re = dd_add(dd_mul(x.re, y.re), dd_neg(dd_mul(x.im, y.im)));
  r.im = dd_add(dd_mul(x.re, y.im), dd_mul(x.im, y.re));
  return r;
}

#endif /* PQP_ORACLE_DD_ARITH_H */
