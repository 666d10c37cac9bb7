/*
 * ORACLE — test infrastructure, not product code.
 *
 * A CPU restatement of the reference's permanent path, used only by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker
 * (or the timed CPU baseline).  The product library under piquasso_amd/
 * never links, loads or calls anything declared here.
 *
 * Two families:
 *
 *   pqo_perm_port_*   plain-precision restatement of permanent_cpp<T>
 *                     (src/permanent.cpp:49-264): min-row prepend, Glynn /
 *                     BBFG sum with multiplicities over the reflected n-ary
 *                     Gray code (src/n_aryGrayCodeCounter.hpp:170-254), the
 *                     same static chunking into `n_jobs` OpenMP jobs
 *                     (src/permanent.cpp:144-165) and the same ordered
 *                     reduction (:253-259).
 *
 *   pqo_perm_dd_*     the same sum evaluated in double-double arithmetic
 *                     (TwoSum / FMA-TwoProd; Dekker 1971, Knuth TAOCP 4.2.2,
 *                     Ogita-Rump-Oishi 2005): the extended-precision arbiter
 *                     for relative-error claims.  pqo_grad_dd follows
 *                     grad_perm (src/permanent.cpp:271-300): explicit minors
 *                     r_i * c_j * perm(A; r - e_i, c - e_j).
 *
 * Conventions: `a` is row-major interleaved complex (re, im) with leading
 * dimension m_c; multiplicities are int32; results are (re, im) pairs.
 * Return 0 on success, -1 when sum(rows) != sum(cols) (the reference throws
 * there, src/permanent.cpp:100-104), -2 on invalid arguments.
 */
#ifndef PQP_ORACLE_H
#define PQP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Plain FP64 restatement; n_jobs <= 0 means 4 * number of processors,
 * as std::thread::hardware_concurrency() * 4 in src/permanent.cpp:145-147. */
int pqo_perm_port_c128(const double *a, int m_r, int m_c, const int *rows,
                       const int *cols, int n_jobs, double *out);

/* Plain FP32 restatement (permanent_cpp<float>, src/permanent.cpp:266-267). */
int pqo_perm_port_c64(const float *a, int m_r, int m_c, const int *rows,
                      const int *cols, int n_jobs, float *out);

/* Extended precision: result as double-double, out = {re_hi, re_lo, im_hi,
 * im_lo}.  n_jobs <= 0 picks the processor count. */
int pqo_perm_dd_c128(const double *a, int m_r, int m_c, const int *rows,
                     const int *cols, int n_jobs, double *out);

/* Gradient d perm / d A_ij (holomorphic), square or not: g is m_r x m_c
 * interleaved complex, each entry r_i * c_j * perm(A; r - e_i, c - e_j)
 * evaluated in double-double and rounded to double. */
int pqo_grad_dd_c128(const double *a, int m_r, int m_c, const int *rows,
                     const int *cols, int n_jobs, double *g);

/* Column minors in double-double (permanent_laplace_cpp<double>,
 * src/permanent_laplace.cpp:43-240): out[l] = perm(A; r, c - e_l) for
 * c_l > 0, the uncut Glynn sum for c_l = 0.  Needs sum c = sum r + 1 > 1. */
int pqo_laplace_dd_c128(const double *a, int m_r, int m_c, const int *rows,
                        const int *cols, int n_jobs, double *out);

/* Mixed-radix reflected Gray code, restating n_aryGrayCodeCounter
 * (src/n_aryGrayCodeCounter.hpp:170-254).  Writes the digit vectors of the
 * Gray states with indices [offset, offset + count) into `codes`
 * (count x n_digits, row-major).  Returns 0, or -2 on invalid input. */
int pqo_gray_states(const int *limits, int n_digits, int64_t offset,
                    int64_t count, int *codes);

/* Exact permanent of an integer matrix (row-major int64, leading dimension
 * m_c; entries and row sums below 2^53): the same Glynn sum in exact
 * multi-word integer arithmetic.  limbs: 8 x 64-bit, little-endian two's
 * complement (512 bits).  Returns 0, -1 (sum mismatch), -2 (invalid
 * arguments), -3 (a term leaves 511 bits), -4 (internal check). */
int pqo_perm_int_exact(const int64_t *a, int m_r, int m_c, const int *rows, const int *cols,
                       int n_jobs, uint64_t *limbs);

/* Number of processors OpenMP sees (used for the default chunking). */
int pqo_num_procs(void);

/* Tensor permanent of uniformly partially distinguishable bosons
 * (piquasso/_simulators/passive/probabilities.py:193-237) of the n x n
 * matrix a (interleaved complex, leading dimension n) at particle overlap
 * x, in double-double: out = {re_hi, re_lo, im_hi, im_lo}; abs_sum (may be
 * null) = sum over subset pairs of |term|.  n <= 16; n = 0 gives 1. */
int pqo_tensor_perm_dd(const double *a, int n, double x, int n_jobs, double *out,
                       double *abs_sum);

/* Coefficients of the permanent of a matrix polynomial,
 * out[p] = [x^e_p] perm(B_0 + sum_m x_m B_m), by Ryser's formula with the
 * polynomial of every column subset truncated to the box of the requested
 * exponents (piquasso/_simulators/passive/probabilities.py:341-382), all in
 * double-double.  b0: n x n, bm: n_vars matrices n x n one after another
 * (interleaved complex); exps: n_out x n_vars; out: (re, im) per exponent;
 * abs_out: sum over subsets of the absolute-value polynomial's coefficient
 * (a scale for forward-error bounds).  n <= 20; -2 on invalid arguments,
 * -3 when the box exceeds 2^26 coefficients or memory runs out. */
int pqo_perm_poly_dd(const double *b0, const double *bm, int n, int n_vars, const int32_t *exps,
                     int n_out, double *out, double *abs_out);

#ifdef __cplusplus
}
#endif

#endif /* PQP_ORACLE_H */
