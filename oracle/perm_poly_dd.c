/*
 * ORACLE — test infrastructure, not product code.
 *
 * Coefficients of the permanent of a matrix polynomial,
 *
 *   c_e = [x^e] perm(B_0 + sum_m x_m B_m),     B_0, B_1..B_M complex n x n,
 *
 * by Ryser's formula over the non-empty column subsets S with the
 * polynomial of every subset truncated to the box 0 <= a_m <= max_p e_p[m]:
 * the subset loop of piquasso's
 * get_lossy_partially_distinguishable_detection_probabilities
 * (piquasso/_simulators/passive/probabilities.py:341-382, with
 * multiply_by_linear_truncated of piquasso/_math/polynomial.py:18-38):
 *
 *   poly_S = prod_i ( c_i(S) + sum_m x_m l_{m,i}(S) ),
 *   c_i(S) = sum_{j in S} B_0[i, j],  l_{m,i}(S) = sum_{j in S} B_m[i, j],
 *   c_e = sum_S (-1)^(n - |S|) poly_S[e].
 *
 * Everything is evaluated in double-double (row sums, products, the subset
 * sum).  Alongside, the same recursion over |c_i(S)| and |l_{m,i}(S)| in
 * plain double gives sum_S |poly|_S[e], which bounds every partial product a
 * plain-FP64 evaluation forms: its forward error is a small multiple of
 * (n + 2) eps times that sum.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "dd_arith.h"
#include "oracle.h"

static cdd_t cdd_add2(cdd_t x, cdd_t y) {
  cdd_t r = {dd_add(x.re, y.re), dd_add(x.im, y.im)};
  return r;
}

static cdd_t cdd_from(double re, double im) {
  cdd_t r = {{re, 0.0}, {im, 0.0}};
  return r;
}

static double cdd_abs(cdd_t x) { return hypot(x.re.hi, x.im.hi); }

int pqo_perm_poly_dd(const double *b0, const double *bm, int n, int n_vars, const int32_t *exps,
                     int n_out, double *out, double *abs_out) {
  if (n < 1 || n > 20 || n_vars < 0 || n_out < 0 || !b0 || (n_vars > 0 && !bm) ||
      (n_out > 0 && (!exps || !out || !abs_out)))
    return -2;
  const int M = n_vars;
  int64_t size = 1;
  int64_t *stride = calloc((size_t)(M > 0 ? M : 1), sizeof(int64_t));
  int32_t *target = calloc((size_t)(M > 0 ? M : 1), sizeof(int32_t));
  if (!stride || !target) {
    free(stride);
    free(target);
    return -3;
  }
  for (int p = 0; p < n_out; ++p)
    for (int m = 0; m < M; ++m) {
      if (exps[(int64_t)p * M + m] < 0) {
        free(stride);
        free(target);
        return -2;
      }
      if (exps[(int64_t)p * M + m] > target[m]) target[m] = exps[(int64_t)p * M + m];
    }
  for (int m = M - 1; m >= 0; --m) {
    stride[m] = size;
    size *= (int64_t)target[m] + 1;
    if (size > (int64_t)1 << 26) {
      free(stride);
      free(target);
      return -3;
    }
  }
  cdd_t *poly = malloc((size_t)size * sizeof(cdd_t));
  cdd_t *next = malloc((size_t)size * sizeof(cdd_t));
  cdd_t *coef = malloc((size_t)size * sizeof(cdd_t));
  double *apoly = malloc((size_t)size * sizeof(double));
  double *anext = malloc((size_t)size * sizeof(double));
  double *acoef = malloc((size_t)size * sizeof(double));
  cdd_t *rs = malloc((size_t)n * (size_t)(M + 1) * sizeof(cdd_t));
  if (!poly || !next || !coef || !apoly || !anext || !acoef || !rs) {
    free(poly), free(next), free(coef), free(apoly), free(anext), free(acoef), free(rs);
    free(stride), free(target);
    return -3;
  }
  for (int64_t k = 0; k < size; ++k) {
    coef[k] = cdd_from(0.0, 0.0);
    acoef[k] = 0.0;
  }
  for (uint32_t s = 1; s < (1u << n); ++s) {
    /* row sums: rs[i][m] = l_{m,i}(S) for m < M, rs[i][M] = c_i(S) */
    for (int i = 0; i < n; ++i)
      for (int m = 0; m <= M; ++m) {
        const double *B = (m == M) ? b0 : bm + (size_t)m * n * n * 2;
        cdd_t acc = cdd_from(0.0, 0.0);
        for (int j = 0; j < n; ++j)
          if (s >> j & 1u) {
            acc.re = dd_add_d(acc.re, B[2 * (i * n + j)]);
            acc.im = dd_add_d(acc.im, B[2 * (i * n + j) + 1]);
          }
        rs[i * (M + 1) + m] = acc;
      }
    for (int64_t k = 0; k < size; ++k) {
      poly[k] = cdd_from(0.0, 0.0);
      apoly[k] = 0.0;
    }
    poly[0] = cdd_from(1.0, 0.0);
    apoly[0] = 1.0;
    for (int i = 0; i < n; ++i) {
      const cdd_t c = rs[i * (M + 1) + M];
      const double ac = cdd_abs(c);
      for (int64_t k = 0; k < size; ++k) {
        next[k] = cdd_mul(c, poly[k]);
        anext[k] = ac * apoly[k];
      }
      for (int m = 0; m < M; ++m) {
        const cdd_t l = rs[i * (M + 1) + m];
        const double al = cdd_abs(l);
        for (int64_t k = 0; k < size; ++k) {
          if ((k / stride[m]) % ((int64_t)target[m] + 1) == 0) continue;
          next[k] = cdd_add2(next[k], cdd_mul(l, poly[k - stride[m]]));
          anext[k] += al * apoly[k - stride[m]];
        }
      }
      cdd_t *t = poly;
      poly = next;
      next = t;
      double *at = apoly;
      apoly = anext;
      anext = at;
    }
    const int odd = (n - __builtin_popcount(s)) & 1;
    for (int64_t k = 0; k < size; ++k) {
      cdd_t v = poly[k];
      if (odd) {
        v.re = dd_neg(v.re);
        v.im = dd_neg(v.im);
      }
      coef[k] = cdd_add2(coef[k], v);
      acoef[k] += apoly[k];
    }
  }
  for (int p = 0; p < n_out; ++p) {
    int64_t k = 0;
    for (int m = 0; m < M; ++m) k += (int64_t)exps[(int64_t)p * M + m] * stride[m];
    out[2 * p] = coef[k].re.hi + coef[k].re.lo;
    out[2 * p + 1] = coef[k].im.hi + coef[k].im.lo;
    abs_out[p] = acoef[k];
  }
  free(poly), free(next), free(coef), free(apoly), free(anext), free(acoef), free(rs);
  free(stride), free(target);
  return 0;
}
