/*
 * ORACLE — test infrastructure, not product code.
 *
 * Exact permanent of an integer matrix with row and column multiplicities:
 * the reference's Glynn / BBFG sum (src/permanent.cpp:49-264: the min-row
 * prepend, the reflected mixed-radix Gray order of
 * src/n_aryGrayCodeCounter.hpp:170-254, binomial weights) evaluated in
 * exact multi-word integer arithmetic instead of floating point.  Every term
 * sign * prod C(rho_i, g_i) * prod_j cs_j^{c_j} is formed as a 512-bit
 * magnitude (64-bit limbs, products by 64-bit factors through unsigned
 * __int128), summed into a 512-bit two's-complement accumulator per OpenMP
 * chunk, the chunks added, and the total shifted right by n - 1 (the sum is
 * 2^(n-1) times an integer, checked).  The arithmetic shares nothing with
 * the GPU's multi-modular evaluation, so it can arbitrate it.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "glynn_common.h"
#include "oracle.h"

#define PQO_LIMBS 8  /* 512 bits */

typedef unsigned __int128 u128;

/* acc (two's complement) += sign * mag (magnitude, PQO_LIMBS limbs) */
static void acc_add(uint64_t *acc, const uint64_t *mag, int negative) {
  if (!negative) {
    u128 carry = 0;
    for (int k = 0; k < PQO_LIMBS; ++k) {
      u128 s = (u128)acc[k] + mag[k] + carry;
      acc[k] = (uint64_t)s;
      carry = s >> 64;
    }
  } else {
    uint64_t borrow = 0;
    for (int k = 0; k < PQO_LIMBS; ++k) {
      uint64_t a = acc[k], m = mag[k];
      uint64_t d = a - m - borrow;
      borrow = (a < m) || (a - m < borrow);
      acc[k] = d;
    }
  }
}

/* mag *= x; returns 0, or -3 when the product leaves 512 bits */
static int mag_mul(uint64_t *mag, uint64_t x) {
  u128 carry = 0;
  for (int k = 0; k < PQO_LIMBS; ++k) {
    u128 p = (u128)mag[k] * x + carry;
    mag[k] = (uint64_t)p;
    carry = p >> 64;
  }
  /* keep one spare bit for the sign of the accumulator */
  return (carry != 0 || (mag[PQO_LIMBS - 1] >> 62) != 0) ? -3 : 0;
}

int pqo_perm_int_exact(const int64_t *a, int m_r, int m_c, const int *rows, const int *cols,
                       int n_jobs, uint64_t *limbs) {
  if (m_r < 0 || m_c < 0 || !limbs || (m_r > 0 && !rows) || (m_c > 0 && !cols)) return -2;
  for (int i = 0; i < m_r; ++i)
    if (rows[i] < 0) return -2;
  for (int j = 0; j < m_c; ++j)
    if (cols[j] < 0) return -2;
  long sum_r = 0, sum_c = 0;
  for (int i = 0; i < m_r; ++i) sum_r += rows[i];
  for (int j = 0; j < m_c; ++j) sum_c += cols[j];
  if (sum_r != sum_c) return -1;
  memset(limbs, 0, PQO_LIMBS * sizeof(uint64_t));
  if (sum_r == 0) {
    limbs[0] = 1;
    return 0;
  }
  /* the prepend on a complex copy (entries below 2^53 are exact) */
  double *ad = (double *)malloc(((size_t)m_r * m_c * 2u + 1u) * sizeof(double));
  if (!ad) return -2;
  for (size_t k = 0; k < (size_t)m_r * m_c; ++k) {
    ad[2 * k] = (double)a[k];
    ad[2 * k + 1] = 0.0;
  }
  pqo_rows_t R;
  if (pqo_rows_build(ad, m_r, m_c, rows, &R) != 0) {
    free(ad);
    return -2;
  }
  free(ad);
  /* Gray digits: rows 1.. with rho > 0; row 0 (rho 1) is fixed at +1 */
  int nd = 0;
  int *dig_row = (int *)malloc((size_t)(R.n_rows + 1) * sizeof(int));
  int *lim = (int *)malloc((size_t)(R.n_rows + 1) * sizeof(int));
  int64_t *b = (int64_t *)malloc(((size_t)R.n_rows * m_c + 1u) * sizeof(int64_t));
  if (!dig_row || !lim || !b) {
    free(dig_row);
    free(lim);
    free(b);
    pqo_rows_free(&R);
    return -2;
  }
  for (int i = 0; i < R.n_rows; ++i)
    for (int j = 0; j < m_c; ++j) b[(size_t)i * m_c + j] = (int64_t)R.b[((size_t)i * m_c + j) * 2];
  for (int i = 1; i < R.n_rows; ++i)
    if (R.rho[i] > 0) {
      dig_row[nd] = i;
      lim[nd] = R.rho[i] + 1;
      nd++;
    }
  int64_t T = 1;
  for (int d = 0; d < nd; ++d) T *= lim[d];
  const int n = (int)sum_r;
  if (n_jobs <= 0) n_jobs = pqo_num_procs();
  if (n_jobs > T) n_jobs = (int)T;
  uint64_t *parts = (uint64_t *)calloc((size_t)n_jobs * PQO_LIMBS, sizeof(uint64_t));
  int status = parts ? 0 : -2;
  if (status == 0) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int job = 0; job < n_jobs; ++job) {
      const int64_t lo = T / n_jobs * job + (job < T % n_jobs ? job : T % n_jobs);
      const int64_t hi = lo + T / n_jobs + (job < T % n_jobs ? 1 : 0);
      int *gray = (int *)malloc((size_t)(nd + 1) * sizeof(int));
      int *ctr = (int *)malloc((size_t)(nd + 1) * sizeof(int));
      int64_t *cs = (int64_t *)malloc(((size_t)m_c + 1u) * sizeof(int64_t));
      uint64_t *acc = parts + (size_t)job * PQO_LIMBS;
      if (!gray || !ctr || !cs) {
#pragma omp atomic write
        status = -2;
      } else {
        pqo_gray_t g = {nd, lim, gray, ctr, 0, hi - 1};
        pqo_gray_init(&g, lo);
        for (int j = 0; j < m_c; ++j) {
          int64_t s = b[j];
          for (int d = 0; d < nd; ++d)
            s += (int64_t)(R.rho[dig_row[d]] - 2 * gray[d]) * b[(size_t)dig_row[d] * m_c + j];
          cs[j] = s;
        }
        for (int64_t t = lo; t < hi; ++t) {
          uint64_t mag[PQO_LIMBS];
          memset(mag, 0, sizeof(mag));
          mag[0] = 1;
          int negative = 0;
          for (int d = 0; d < nd; ++d) {
            mag_mul(mag, (uint64_t)pqo_binom(R.rho[dig_row[d]], gray[d]));
            negative ^= gray[d] & 1;
          }
          int zero = 0, over = 0;
          for (int j = 0; j < m_c && !zero; ++j) {
            if (cols[j] == 0) continue;
            if (cs[j] == 0) {
              zero = 1;
              break;
            }
            const uint64_t x = (uint64_t)(cs[j] < 0 ? -cs[j] : cs[j]);
            for (int q = 0; q < cols[j]; ++q) over |= mag_mul(mag, x);
            if (cs[j] < 0 && (cols[j] & 1)) negative ^= 1;
          }
          if (over) {
#pragma omp atomic write
            status = -3;
            break;
          }
          if (!zero) acc_add(acc, mag, negative);
          if (t + 1 < hi) {
            int changed, prev, value;
            pqo_gray_next(&g, &changed, &prev, &value);
            const int64_t f = 2 * (int64_t)(prev - value);
            const int64_t *row = b + (size_t)dig_row[changed] * m_c;
            for (int j = 0; j < m_c; ++j) cs[j] += f * row[j];
          }
        }
      }
      free(gray);
      free(ctr);
      free(cs);
    }
  }
  if (status == 0) {
    for (int job = 0; job < n_jobs; ++job) {
      u128 carry = 0;
      for (int k = 0; k < PQO_LIMBS; ++k) {
        u128 s = (u128)limbs[k] + parts[(size_t)job * PQO_LIMBS + k] + carry;
        limbs[k] = (uint64_t)s;
        carry = s >> 64;
      }
    }
    /* exact division by 2^(n-1): arithmetic shift, the low bits must be 0 */
    for (int s = 0; s < n - 1; ++s) {
      if (limbs[0] & 1u) {
        status = -4;
        break;
      }
      const uint64_t sign = limbs[PQO_LIMBS - 1] >> 63;
      for (int k = 0; k < PQO_LIMBS - 1; ++k) limbs[k] = (limbs[k] >> 1) | (limbs[k + 1] << 63);
      limbs[PQO_LIMBS - 1] = (limbs[PQO_LIMBS - 1] >> 1) | (sign << 63);
    }
  }
  free(parts);
  free(dig_row);
  free(lim);
  free(b);
  pqo_rows_free(&R);
  return status;
}
