/*
 * ORACLE — test infrastructure, not product code.
 *
 * Extended-precision (double-double) evaluation of the reference's Glynn /
 * BBFG permanent with multiplicities (formula of src/permanent.cpp:41-50,
 * Eq. (8) of arXiv:2309.07027), used as the arbiter for relative-error
 * claims; and the reference gradient grad_perm (src/permanent.cpp:271-300)
 * as explicit minors r_i c_j perm(A; r - e_i, c - e_j), each minor in
 * double-double.
 *
 * Double-double arithmetic: TwoSum (Knuth, TAOCP vol. 2, 4.2.2), FastTwoSum
 * and FMA-based TwoProd (Dekker 1971; Ogita, Rump, Oishi, SIAM J. Sci.
 * Comput. 26, 2005).  Relative accuracy of each operation is about 2^-104;
 * the row sums are updated exactly (2 b_ij is exact in double and is added
 * to a double-double), so the result carries ~1e-28 relative error times
 * the cancellation of the alternating sum (~1e7 at n = 32).
 *
 * Rows with zero multiplicity (radix-1 digits) and columns with zero
 * multiplicity are dropped before the sweep; that changes the summation
 * order but not the mathematical value, which is all this oracle vouches
 * for.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "dd_arith.h"
#include "glynn_common.h"
#include "oracle.h"

/* Glynn sum in double-double over an already prepended, compacted row set.
 * b: nr x mc interleaved complex, row 0 = the prepended row (multiplicity
 * 1, fixed sign +1); rho[1..nr-1] >= 1; c[0..mc-1] >= 1; n = sum c.
 * Returns the unscaled sum (the caller multiplies by 2^-(n-1)). */
static cdd_t dd_glynn_sum(const double *b, int nr, const int *rho, int mc,
                          const int *c, int n_jobs) {
  const int nd = nr > 0 ? nr - 1 : 0;
  int *lim = (int *)malloc((size_t)(nd ? nd : 1) * sizeof(int));
  int64_t T = 1;
  for (int i = 0; i < nd; ++i) {
    lim[i] = rho[i + 1] + 1;
    T *= lim[i];
  }
  if (n_jobs <= 0) n_jobs = pqo_num_procs();
  int64_t jobs = (int64_t)n_jobs * 8;
  if (jobs > T) jobs = T;
  cdd_t *res = (cdd_t *)calloc((size_t)jobs, sizeof(cdd_t));
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_jobs)
  for (int64_t job = 0; job < jobs; ++job) {
    int64_t first = T / jobs * job, last = T / jobs * (job + 1) - 1;
    if (job == jobs - 1) last = T - 1;
    int *gray = (int *)malloc((size_t)(nd ? nd : 1) * sizeof(int));
    int *ctr = (int *)malloc((size_t)(nd ? nd : 1) * sizeof(int));
    cdd_t *cs = (cdd_t *)malloc((size_t)mc * sizeof(cdd_t));
    pqo_gray_t g = {nd, lim, gray, ctr, 0, last};
    pqo_gray_init(&g, first);
    double weight = 1.0;
    int minus = 0;
    for (int j = 0; j < mc; ++j) {
      cs[j].re.hi = b[2 * j];
      cs[j].re.lo = 0.0;
      cs[j].im.hi = b[2 * j + 1];
      cs[j].im.lo = 0.0;
    }
    for (int i = 0; i < nd; ++i) {
      const double f = (double)(rho[i + 1] - 2 * gray[i]);
      const double *row = b + (size_t)(i + 1) * mc * 2u;
      for (int j = 0; j < mc; ++j) {
        cs[j].re = dd_add(cs[j].re, dd_from_prod(row[2 * j], f));
        cs[j].im = dd_add(cs[j].im, dd_from_prod(row[2 * j + 1], f));
      }
      minus += gray[i];
      weight *= (double)pqo_binom(rho[i + 1], gray[i]);
    }
    cdd_t acc = {{0, 0}, {0, 0}};
    int sign = (minus & 1) ? -1 : 1;
    for (int64_t it = first;; ++it) {
      cdd_t p = {{1.0, 0.0}, {0.0, 0.0}};
      for (int j = 0; j < mc; ++j)
        for (int q = 0; q < c[j]; ++q) p = cdd_mul(p, cs[j]);
      const double w = sign * weight;
      acc.re = dd_add(acc.re, dd_mul_d(p.re, w));
      acc.im = dd_add(acc.im, dd_mul_d(p.im, w));
      if (it == last) break;
      int ch = 0, pv = 0, val = 0;
      pqo_gray_next(&g, &ch, &pv, &val);
      sign = -sign;
      const double *row = b + (size_t)(ch + 1) * mc * 2u;
      const double d = (double)(2 * (pv - val)); /* +-2: exact scaling */
      for (int j = 0; j < mc; ++j) {
        cs[j].re = dd_add_d(cs[j].re, d * row[2 * j]);
        cs[j].im = dd_add_d(cs[j].im, d * row[2 * j + 1]);
      }
      const int rm = rho[ch + 1];
      weight = (double)pqo_binom(rm, val) / (double)pqo_binom(rm, pv) * weight;
      weight = nearbyint(weight); /* integer-valued; remove division rounding */
    }
    res[job] = acc;
    free(gray);
    free(ctr);
    free(cs);
  }
  cdd_t s = {{0, 0}, {0, 0}};
  for (int64_t job = 0; job < jobs; ++job) {
    s.re = dd_add(s.re, res[job].re);
    s.im = dd_add(s.im, res[job].im);
  }
  free(res);
  free(lim);
  return s;
}

/* Compact rows/columns with zero multiplicity, apply the reference prepend
 * rule on the compacted rows (same row chosen: the first minimal non-zero
 * one), then evaluate.  out = {re_hi, re_lo, im_hi, im_lo}. */
static int dd_perm_any(const double *a, int m_r, int m_c, const int *rows,
                       const int *cols, int n_jobs, int check_sums, double *out) {
  long sr = 0, sc = 0;
  for (int i = 0; i < m_r; ++i) {
    if (rows[i] < 0) return -2;
    sr += rows[i];
  }
  for (int j = 0; j < m_c; ++j) {
    if (cols[j] < 0) return -2;
    sc += cols[j];
  }
  if (check_sums && sr != sc) return -1;
  if (sr == 0) {
    out[0] = 1.0;
    out[1] = out[2] = out[3] = 0.0;
    return 0;
  }
  int *ri = (int *)malloc((size_t)m_r * sizeof(int));
  int *ci = (int *)malloc((size_t)m_c * sizeof(int));
  int nr = 0, nc = 0;
  for (int i = 0; i < m_r; ++i)
    if (rows[i] > 0) ri[nr++] = i;
  for (int j = 0; j < m_c; ++j)
    if (cols[j] > 0) ci[nc++] = j;
  double *ac = (double *)malloc((size_t)nr * nc * 2u * sizeof(double));
  int *rc = (int *)malloc((size_t)nr * sizeof(int));
  int *cc = (int *)malloc((size_t)nc * sizeof(int));
  for (int i = 0; i < nr; ++i) {
    rc[i] = rows[ri[i]];
    for (int j = 0; j < nc; ++j) {
      ac[2 * ((size_t)i * nc + j)] = a[2 * ((size_t)ri[i] * m_c + ci[j])];
      ac[2 * ((size_t)i * nc + j) + 1] = a[2 * ((size_t)ri[i] * m_c + ci[j]) + 1];
    }
  }
  for (int j = 0; j < nc; ++j) cc[j] = cols[ci[j]];
  pqo_rows_t R;
  int st = pqo_rows_build(ac, nr, nc, rc, &R);
  if (st == 0) {
    /* drop the prepended-from row if its multiplicity fell to zero */
    int keep = 0;
    for (int i = 0; i < R.n_rows; ++i) {
      if (i > 0 && R.rho[i] == 0) continue;
      if (keep != i)
        memmove(R.b + (size_t)keep * nc * 2u, R.b + (size_t)i * nc * 2u,
                (size_t)nc * 2u * sizeof(double));
      R.rho[keep] = R.rho[i];
      keep++;
    }
    cdd_t s = dd_glynn_sum(R.b, keep, R.rho, nc, cc, n_jobs);
    const int e = -(int)(sr - 1);
    out[0] = ldexp(s.re.hi, e);
    out[1] = ldexp(s.re.lo, e);
    out[2] = ldexp(s.im.hi, e);
    out[3] = ldexp(s.im.lo, e);
    pqo_rows_free(&R);
  }
  free(ri);
  free(ci);
  free(ac);
  free(rc);
  free(cc);
  return st;
}

static int dd_perm(const double *a, int m_r, int m_c, const int *rows,
                   const int *cols, int n_jobs, double *out) {
  return dd_perm_any(a, m_r, m_c, rows, cols, n_jobs, 1, out);
}

int pqo_perm_dd_c128(const double *a, int m_r, int m_c, const int *rows,
                     const int *cols, int n_jobs, double *out) {
  if (m_r < 0 || m_c < 0 || !out) return -2;
  return dd_perm(a, m_r, m_c, rows, cols, n_jobs, out);
}

/* permanent_laplace_cpp<double> (src/permanent_laplace.cpp:43-240) as
 * separate double-double sums: out[l] = perm(A; r, c - e_l) for c_l > 0 and,
 * for c_l = 0, the Glynn sum with the uncut column multiplicities, scaled by
 * 2^-(sum r - 1) (the reference's loop runs the product over all columns
 * there, :206-212, and divides by that power of two, :228).  Requires
 * sum c = sum r + 1 and sum r > 0; out is m_c interleaved complex values. */
int pqo_laplace_dd_c128(const double *a, int m_r, int m_c, const int *rows,
                        const int *cols, int n_jobs, double *out) {
  if (m_r <= 0 || m_c <= 0 || !out) return -2;
  long sr = 0, sc = 0;
  for (int i = 0; i < m_r; ++i) sr += rows[i];
  for (int j = 0; j < m_c; ++j) sc += cols[j];
  if (sr <= 0 || sc != sr + 1) return -1;
  int *c2 = (int *)malloc((size_t)m_c * sizeof(int));
  memcpy(c2, cols, (size_t)m_c * sizeof(int));
  int st = 0;
  for (int l = 0; l < m_c && st == 0; ++l) {
    double v[4];
    if (cols[l] > 0) {
      c2[l] -= 1;
      st = dd_perm(a, m_r, m_c, rows, c2, n_jobs, v);
      c2[l] += 1;
    } else {
      st = dd_perm_any(a, m_r, m_c, rows, cols, n_jobs, 0, v);
    }
    out[2 * l] = v[0] + v[1];
    out[2 * l + 1] = v[2] + v[3];
  }
  free(c2);
  return st;
}

int pqo_grad_dd_c128(const double *a, int m_r, int m_c, const int *rows,
                     const int *cols, int n_jobs, double *g) {
  if (m_r < 0 || m_c < 0 || !g) return -2;
  long sr = 0, sc = 0;
  for (int i = 0; i < m_r; ++i) sr += rows[i];
  for (int j = 0; j < m_c; ++j) sc += cols[j];
  if (sr != sc) return -1;
  int *r2 = (int *)malloc((size_t)(m_r ? m_r : 1) * sizeof(int));
  int *c2 = (int *)malloc((size_t)(m_c ? m_c : 1) * sizeof(int));
  memcpy(r2, rows, (size_t)m_r * sizeof(int));
  memcpy(c2, cols, (size_t)m_c * sizeof(int));
  int st = 0;
  for (int i = 0; i < m_r && st == 0; ++i)
    for (int j = 0; j < m_c && st == 0; ++j) {
      double *gij = g + 2 * ((size_t)i * m_c + j);
      if (rows[i] == 0 || cols[j] == 0) {
        gij[0] = gij[1] = 0.0;
        continue;
      }
      r2[i] -= 1;
      c2[j] -= 1;
      double p[4];
      st = dd_perm(a, m_r, m_c, r2, c2, n_jobs, p);
      r2[i] += 1;
      c2[j] += 1;
      const double f = (double)rows[i] * (double)cols[j];
      dd_t re = {p[0], p[1]}, im = {p[2], p[3]};
      re = dd_mul_d(re, f);
      im = dd_mul_d(im, f);
      gij[0] = re.hi + re.lo;
      gij[1] = im.hi + im.lo;
    }
  free(r2);
  free(c2);
  return st;
}
