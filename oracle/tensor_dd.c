/*
 * ORACLE — test infrastructure, not product code.
 *
 * Tensor permanent of n uniformly partially distinguishable bosons, as
 * piquasso evaluates it for detection probabilities
 * (piquasso/_simulators/passive/probabilities.py:156-237; Tichy,
 * arXiv:1410.7687, Sec. III):
 *
 *   T(A, x) = sum over non-empty column subsets s, t of
 *             (-1)^(|s| + |t|) prod_i ( x R_s[i] conj(R_t[i]) + (1 - x) Q_{s&t}[i] )
 *
 * with R_s[i] = sum_{c in s} A[i, c] and Q_u[i] = sum_{c in u} |A[i, c]|^2
 * (the reference's sign (-1)^(n-|s|) (-1)^(n-|t|) is the same number).
 * Everything is evaluated in double-double: row sums, factors, products and
 * the sum, so the result carries ~1e-28 relative error times the
 * cancellation of the alternating sum.  Also returns sum |term| (double) for
 * forward-error tolerances of plain-FP64 evaluations.
 *
 * Pairs are split over OpenMP threads by s; every s's partial is kept and
 * the partials are added in order of s, so the result does not depend on the
 * thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "dd_arith.h"
#include "oracle.h"

static cdd_t cdd_add(cdd_t x, cdd_t y) {
  cdd_t r = {dd_add(x.re, y.re), dd_add(x.im, y.im)};
  return r;
}

int pqo_tensor_perm_dd(const double *a, int n, double x, int n_jobs, double *out,
                       double *abs_sum) {
  if (n < 0 || n > 16 || !out || (n > 0 && !a)) return -2;
  if (n == 0) {
    out[0] = 1.0;
    out[1] = out[2] = out[3] = 0.0;
    if (abs_sum) *abs_sum = 1.0;
    return 0;
  }
  const int64_t N = (int64_t)1 << n;
  cdd_t *R = (cdd_t *)calloc((size_t)(N * n), sizeof(cdd_t));
  dd_t *Q = (dd_t *)calloc((size_t)(N * n), sizeof(dd_t));
  cdd_t *part = (cdd_t *)calloc((size_t)N, sizeof(cdd_t));
  double *apart = (double *)calloc((size_t)N, sizeof(double));
  if (!R || !Q || !part || !apart) {
    free(R);
    free(Q);
    free(part);
    free(apart);
    return -2;
  }
  /* subset sums: R_s = R_{s without its lowest column} + column */
  for (int64_t s = 1; s < N; ++s) {
    const int c = __builtin_ctzll((unsigned long long)s);
    const int64_t p = s & (s - 1);
    for (int i = 0; i < n; ++i) {
      const double re = a[2 * (i * n + c)], im = a[2 * (i * n + c) + 1];
      R[s * n + i].re = dd_add_d(R[p * n + i].re, re);
      R[s * n + i].im = dd_add_d(R[p * n + i].im, im);
      dd_t sq = dd_add(dd_from_prod(re, re), dd_from_prod(im, im));
      Q[s * n + i] = dd_add(Q[p * n + i], sq);
    }
  }
  const dd_t xd = {x, 0.0};
  const dd_t yd = dd_two_sum(1.0, -x);
  if (n_jobs <= 0) {
#ifdef _OPENMP
    n_jobs = omp_get_num_procs();
#else
    n_jobs = 1;
#endif
  }
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_jobs)
  for (int64_t s = 1; s < N; ++s) {
    cdd_t acc = {{0.0, 0.0}, {0.0, 0.0}};
    double aacc = 0.0;
    const int ps = __builtin_popcountll((unsigned long long)s);
    for (int64_t t = 1; t < N; ++t) {
      const int64_t u = s & t;
      cdd_t prod = {{1.0, 0.0}, {0.0, 0.0}};
      for (int i = 0; i < n; ++i) {
        const cdd_t rs = R[s * n + i];
        const cdd_t rt = R[t * n + i];
        /* x R_s conj(R_t) + (1 - x) Q_u */
        cdd_t f;
        f.re = dd_add(dd_mul(rs.re, rt.re), dd_mul(rs.im, rt.im));
        f.im = dd_add(dd_mul(rs.im, rt.re), dd_neg(dd_mul(rs.re, rt.im)));
        f.re = dd_add(dd_mul(xd, f.re), dd_mul(yd, Q[u * n + i]));
        f.im = dd_mul(xd, f.im);
        prod = cdd_mul(prod, f);
      }
      if ((ps + __builtin_popcountll((unsigned long long)t)) & 1) {
        prod.re = dd_neg(prod.re);
        prod.im = dd_neg(prod.im);
      }
      acc = cdd_add(acc, prod);
      aacc += hypot(prod.re.hi, prod.im.hi);
    }
    part[s] = acc;
    apart[s] = aacc;
  }
  cdd_t sum = {{0.0, 0.0}, {0.0, 0.0}};
  double asum = 0.0;
  for (int64_t s = 1; s < N; ++s) {
    sum = cdd_add(sum, part[s]);
    asum += apart[s];
  }
  out[0] = sum.re.hi;
  out[1] = sum.re.lo;
  out[2] = sum.im.hi;
  out[3] = sum.im.lo;
  if (abs_sum) *abs_sum = asum;
  free(R);
  free(Q);
  free(part);
  free(apart);
  return 0;
}
