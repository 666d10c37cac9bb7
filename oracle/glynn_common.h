/*
 * ORACLE — test infrastructure, not product code.
 *
 * Pieces shared by the plain restatement (glynn_port.c) and the
 * double-double oracle (glynn_dd.c): the min-row prepend and the reflected
 * mixed-radix Gray counter of the reference.
 */
#ifndef PQP_GLYNN_COMMON_H
#define PQP_GLYNN_COMMON_H

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/*
 * Row set after the reference's "prepend a copy of the first minimal
 * non-zero row" step (src/permanent.cpp:54-95):
 *   row 0      = A[i0, :] with multiplicity 1,
 *   row k + 1  = A[k, :]  with multiplicity rows[k] (minus one for k == i0).
 * `b` holds (m_r + 1) x m_c interleaved complex values, `rho` the m_r + 1
 * multiplicities.  When every row multiplicity is zero nothing is prepended
 * (then prepended == 0 and b / rho are copies of the input).
 */
typedef struct {
  int n_rows;     /* rows in b (m_r + 1 when prepended, else m_r) */
  int m_c;
  int prepended;
  double *b;      /* n_rows x m_c interleaved complex */
  int *rho;       /* n_rows multiplicities */
} pqo_rows_t;

/* The tie-break of src/permanent.cpp:57-64: the running minimum starts at 0
 * and is replaced by any element while it is 0, afterwards only by a smaller
 * non-zero element, so the first minimal non-zero row wins. */
static inline int pqo_min_row(const int *rows, int m_r, int *minelem_out) {
  int min_idx = 0, minelem = 0;
  for (int i = 0; i < m_r; ++i) {
    if (minelem == 0 || (rows[i] < minelem && rows[i] != 0)) {
      minelem = rows[i];
      min_idx = i;
    }
  }
  *minelem_out = minelem;
  return min_idx;
}

static inline int pqo_rows_build(const double *a, int m_r, int m_c,
                                 const int *rows, pqo_rows_t *out) {
  int minelem = 0;
  int i0 = pqo_min_row(rows, m_r, &minelem);
  int prep = (m_r > 0 && minelem != 0);
  out->n_rows = m_r + prep;
  out->m_c = m_c;
  out->prepended = prep;
  size_t nb = (size_t)out->n_rows * (size_t)m_c * 2u;
  out->b = (double *)malloc(nb ? nb * sizeof(double) : sizeof(double));
  out->rho = (int *)malloc((size_t)(out->n_rows ? out->n_rows : 1) * sizeof(int));
  if (!out->b || !out->rho) return -2;
  if (prep) {
    memcpy(out->b, a + (size_t)i0 * m_c * 2, (size_t)m_c * 2 * sizeof(double));
    memcpy(out->b + (size_t)m_c * 2, a, (size_t)m_r * m_c * 2 * sizeof(double));
    out->rho[0] = 1;
    for (int k = 0; k < m_r; ++k) out->rho[k + 1] = rows[k];
    out->rho[1 + i0] -= 1;
  } else {
    if (nb) memcpy(out->b, a, nb * sizeof(double));
    for (int k = 0; k < m_r; ++k) out->rho[k] = rows[k];
  }
  return 0;
}

static inline void pqo_rows_free(pqo_rows_t *r) {
  free(r->b);
  free(r->rho);
}

/*
 * Reflected mixed-radix Gray counter (src/n_aryGrayCodeCounter.hpp).
 * Digit 0 is the least significant.  `initialize` (:170-194): mixed-radix
 * digits of the offset, then top-down `code = parity ? lim-1-d : d`,
 * `parity ^= code & 1`.  `next` (:209-254): increment the counter with
 * carry, then scan top-down for the first digit whose Gray value changed.
 * The offset is kept as int64 (the reference narrows it to int at :179,
 * which is exact while the term count stays below 2^31).
 */
typedef struct {
  int nd;
  const int *lim;
  int *gray;
  int *ctr;
  int64_t offset, offset_max;
} pqo_gray_t;

static inline void pqo_gray_init(pqo_gray_t *g, int64_t offset) {
  int64_t t = offset;
  for (int i = 0; i < g->nd; ++i) {
    g->ctr[i] = (int)(t % g->lim[i]);
    t /= g->lim[i];
  }
  int parity = 0;
  for (int i = g->nd - 1; i >= 0; --i) {
    int code = parity ? g->lim[i] - 1 - g->ctr[i] : g->ctr[i];
    g->gray[i] = code;
    parity ^= (code & 1);
  }
  g->offset = offset;
}

/* Returns 1 when the counter is exhausted, else 0 and the changed digit. */
static inline int pqo_gray_next(pqo_gray_t *g, int *changed, int *prev,
                                int *value) {
  *changed = 0;
  if (g->offset >= g->offset_max) return 1;
  int idx = 0;
  for (;;) {
    if (g->ctr[idx] < g->lim[idx] - 1) {
      g->ctr[idx]++;
      break;
    }
    g->ctr[idx] = 0;
    idx++;
  }
  int parity = 0;
  for (int i = g->nd - 1; i >= 0; --i) {
    int v = parity ? g->lim[i] - 1 - g->ctr[i] : g->ctr[i];
    parity ^= (v & 1);
    if (v != g->gray[i]) {
      *prev = g->gray[i];
      *value = v;
      g->gray[i] = v;
      *changed = i;
      break;
    }
  }
  g->offset++;
  return 0;
}

/* Binomial coefficient, exact in int64 for the sizes used here
 * (same values as binomialCoeff<int>, src/utils.hpp:43-61). */
static inline int64_t pqo_binom(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

#endif /* PQP_GLYNN_COMMON_H */
