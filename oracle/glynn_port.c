/*
 * ORACLE — test infrastructure, not product code.
 *
 * Plain-precision CPU restatement of the reference permanent,
 * permanent_cpp<T> (src/permanent.cpp:49-264), written from its behaviour:
 *
 *   1. prepend a copy of the first minimal non-zero row, multiplicity 1,
 *      and decrement that row (:54-95);
 *   2. sum(rows) != sum(cols) -> error (:97-104); empty -> 1 (:106-108);
 *   3. the Gray digits are the remaining rows, radix rho_i + 1 (:130-142);
 *   4. split [0, T) into `n_jobs` equal chunks, the last takes the rest
 *      (:144-165); per chunk start the Gray counter at the chunk offset,
 *      build the row sums b_0j + sum_i (rho_i - 2 g_i) b_ij, the binomial
 *      weight prod C(rho_i, g_i) and the sign (:166-215);
 *   5. walk the chunk: each step moves one digit by +-1, adds
 *      2 (prev - value) b_row to every row sum, recomputes the product
 *      sign * prod_j colsum_j^{c_j} (all columns, zero multiplicity included,
 *      as :233-241) and updates the weight by exact integer ratio (:243-247);
 *   6. sum the chunk results in chunk order and divide by 2^(n-1) (:253-259).
 *
 * The dead single-row branch (:110-122) is not restated: after the prepend
 * there are always at least two rows.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "glynn_common.h"
#include "oracle.h"

int pqo_num_procs(void) {
#ifdef _OPENMP
  return omp_get_num_procs();
#else
  return 1;
#endif
}

#define PQO_DEFINE_PORT(NAME, T)                                               \
  int NAME(const T *a, int m_r, int m_c, const int *rows, const int *cols,    \
           int n_jobs, T *out) {                                              \
    if (m_r < 0 || m_c < 0 || !out) return -2;                                \
    pqo_rows_t R;                                                              \
    double *ad = (double *)malloc(                                             \
        ((size_t)m_r * m_c * 2u + 1u) * sizeof(double));                       \
    if (!ad) return -2;                                                        \
    for (size_t k = 0; k < (size_t)m_r * m_c * 2u; ++k) ad[k] = (double)a[k];  \
    if (pqo_rows_build(ad, m_r, m_c, rows, &R) != 0) {                         \
      free(ad);                                                                \
      return -2;                                                               \
    }                                                                          \
    free(ad);                                                                  \
    long sum_r = 0, sum_c = 0;                                                 \
    for (int i = 0; i < R.n_rows; ++i) sum_r += R.rho[i];                      \
    for (int j = 0; j < m_c; ++j) sum_c += cols[j];                            \
    if (sum_r != sum_c) {                                                      \
      pqo_rows_free(&R);                                                       \
      return -1;                                                               \
    }                                                                          \
    if (R.n_rows == 0 || m_c == 0 || sum_r == 0) {                             \
      out[0] = (T)1;                                                           \
      out[1] = (T)0;                                                           \
      pqo_rows_free(&R);                                                       \
      return 0;                                                                \
    }                                                                          \
    const int nd = R.n_rows - 1;                                               \
    int *lim = (int *)malloc((size_t)(nd ? nd : 1) * sizeof(int));             \
    int64_t idx_max = 1;                                                       \
    for (int i = 0; i < nd; ++i) {                                             \
      lim[i] = R.rho[i + 1] + 1;                                               \
      idx_max *= lim[i];                                                       \
    }                                                                          \
    if (n_jobs <= 0) n_jobs = 4 * pqo_num_procs();                             \
    int64_t conc = n_jobs < idx_max ? n_jobs : idx_max;                        \
    T *res = (T *)calloc((size_t)conc * 2u, sizeof(T));                        \
    T *bt = (T *)malloc((size_t)R.n_rows * m_c * 2u * sizeof(T));              \
    for (size_t k = 0; k < (size_t)R.n_rows * m_c * 2u; ++k)                   \
      bt[k] = (T)R.b[k];                                                       \
    _Pragma("omp parallel for schedule(static) num_threads(conc)")             \
    for (int64_t job = 0; job < conc; ++job) {                                 \
      int64_t batch = idx_max / conc;                                          \
      int64_t first = job * batch;                                             \
      int64_t last = (job + 1) * batch - 1;                                    \
      if (job == conc - 1) last = idx_max - 1;                                 \
      int *gray = (int *)malloc((size_t)(nd ? nd : 1) * sizeof(int));         \
      int *ctr = (int *)malloc((size_t)(nd ? nd : 1) * sizeof(int));          \
      T *cs = (T *)malloc((size_t)m_c * 2u * sizeof(T));                       \
      pqo_gray_t g = {nd, lim, gray, ctr, 0, last};                            \
      pqo_gray_init(&g, first);                                                \
      int64_t binom = 1;                                                       \
      memcpy(cs, bt, (size_t)m_c * 2u * sizeof(T));                            \
      int minus_all = 0;                                                       \
      for (int i = 0; i < nd; ++i) {                                           \
        int k = gray[i], rm = R.rho[i + 1];                                    \
        const T f = (T)(rm - 2 * k);                                           \
        const T *brow = bt + (size_t)(i + 1) * m_c * 2u;                       \
        for (int j = 0; j < m_c; ++j) {                                        \
          cs[2 * j] += brow[2 * j] * f;                                        \
          cs[2 * j + 1] += brow[2 * j + 1] * f;                                \
        }                                                                      \
        minus_all += k;                                                        \
        binom *= pqo_binom(rm, k);                                             \
      }                                                                        \
      int parity = (minus_all % 2 == 0) ? 1 : -1;                              \
      T pr = (T)parity, pi = (T)0;                                             \
      for (int j = 0; j < m_c; ++j)                                            \
        for (int q = 0; q < cols[j]; ++q) {                                    \
          T xr = cs[2 * j], xi = cs[2 * j + 1];                                \
          T nr = pr * xr - pi * xi, ni = pr * xi + pi * xr;                    \
          pr = nr;                                                             \
          pi = ni;                                                             \
        }                                                                      \
      T acc_r = pr * (T)binom, acc_i = pi * (T)binom;                          \
      for (int64_t it = first + 1; it < last + 1; ++it) {                      \
        int ch = 0, pv = 0, val = 0;                                           \
        if (pqo_gray_next(&g, &ch, &pv, &val)) break;                          \
        parity = -parity;                                                      \
        const T *b2 = bt + (size_t)(ch + 1) * m_c * 2u;                        \
        const T d = (T)(2 * (pv - val));                                       \
        pr = (T)parity;                                                        \
        pi = (T)0;                                                             \
        for (int j = 0; j < m_c; ++j) {                                        \
          cs[2 * j] += b2[2 * j] * d;                                          \
          cs[2 * j + 1] += b2[2 * j + 1] * d;                                  \
          for (int q = 0; q < cols[j]; ++q) {                                  \
            T xr = cs[2 * j], xi = cs[2 * j + 1];                              \
            T nr = pr * xr - pi * xi, ni = pr * xi + pi * xr;                  \
            pr = nr;                                                           \
            pi = ni;                                                           \
          }                                                                    \
        }                                                                      \
        int rm = R.rho[ch + 1];                                                \
        binom = val < pv ? binom * pv / (rm - val)                             \
                         : binom * (rm - pv) / val;                            \
        acc_r += pr * (T)binom;                                                \
        acc_i += pi * (T)binom;                                                \
      }                                                                        \
      res[2 * job] = acc_r;                                                    \
      res[2 * job + 1] = acc_i;                                                \
      free(gray);                                                              \
      free(ctr);                                                               \
      free(cs);                                                                \
    }                                                                          \
    T sr = 0, si = 0;                                                          \
    for (int64_t job = 0; job < conc; ++job) {                                 \
      sr += res[2 * job];                                                      \
      si += res[2 * job + 1];                                                  \
    }                                                                          \
    T scale = (T)ldexp(1.0, (int)sum_r - 1);                                   \
    out[0] = sr / scale;                                                       \
    out[1] = si / scale;                                                       \
    free(res);                                                                 \
    free(bt);                                                                  \
    free(lim);                                                                 \
    pqo_rows_free(&R);                                                         \
    return 0;                                                                  \
  }

PQO_DEFINE_PORT(pqo_perm_port_c128, double)
PQO_DEFINE_PORT(pqo_perm_port_c64, float)

int pqo_gray_states(const int *limits, int n_digits, int64_t offset,
                    int64_t count, int *codes) {
  if (n_digits < 0 || offset < 0 || count < 0) return -2;
  int64_t total = 1;
  for (int i = 0; i < n_digits; ++i) {
    if (limits[i] < 1) return -2;
    total *= limits[i];
  }
  if (offset + count > total) return -2;
  if (count == 0) return 0;
  int *gray = (int *)malloc((size_t)(n_digits ? n_digits : 1) * sizeof(int));
  int *ctr = (int *)malloc((size_t)(n_digits ? n_digits : 1) * sizeof(int));
  pqo_gray_t g = {n_digits, limits, gray, ctr, 0, total - 1};
  pqo_gray_init(&g, offset);
  memcpy(codes, gray, (size_t)n_digits * sizeof(int));
  for (int64_t s = 1; s < count; ++s) {
    int ch = 0, pv = 0, val = 0;
    pqo_gray_next(&g, &ch, &pv, &val);
    memcpy(codes + s * n_digits, gray, (size_t)n_digits * sizeof(int));
  }
  free(gray);
  free(ctr);
  return 0;
}
