/*
 * pqp_permanent.h — C ABI of the MI355X permanent engine (libpqp_permanent.so).
 *
 * The engine evaluates matrix permanents with row and column multiplicities,
 *
 *   perm(A; r, c) = permanent of the n x n matrix that repeats row i of A
 *                   r_i times and column j c_j times,  n = sum r = sum c,
 *
 * by Glynn's formula with multiplicities over a reflected mixed-radix Gray
 * code, the method of the reference's permanent_cpp<T>
 * (src/permanent.cpp:49-264, src/n_aryGrayCodeCounter.hpp), on AMD gfx950
 * GPUs.  Every sum and product runs on the GPU; the host only validates,
 * compacts and tabulates.  There is no CPU fallback: without a usable HIP
 * device every compute entry point returns PQP_ERR_DEVICE.  Setting the
 * environment variable PQP_FORCE_NO_DEVICE (to anything but "" or "0")
 * before the first call makes the library behave so on a machine that has
 * a device, without initialising the HIP runtime.
 *
 * Conventions
 *  - Matrices are row-major, interleaved complex (re, im), leading dimension
 *    `ld` in complex elements (ld >= m_c).  Host pointers; inputs are read
 *    only for the duration of the call and never written.
 *  - Multiplicities are int32, one per row (length m_r) and per column
 *    (length m_c); zero entries are allowed.
 *  - Return value: PQP_OK (0) or a negative PQP_ERR_* code; the message of
 *    the last failure on the calling thread is pqp_last_error().  Nothing
 *    throws across this boundary.
 *  - Thread-safe: calls may come from several threads (they serialise on
 *    the device they share).
 */
#ifndef PQP_PERMANENT_H
#define PQP_PERMANENT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PQP_OK 0
#define PQP_ERR_SUM_MISMATCH (-1) /* sum(rows) != sum(cols): the reference throws (src/permanent.cpp:100-104) */
#define PQP_ERR_INVALID (-2)      /* bad shape, negative multiplicity, null pointer */
#define PQP_ERR_RANGE (-3)        /* more than 64 columns / particles, or too many Gray terms */
#define PQP_ERR_DEVICE (-4)       /* no HIP device, or a HIP runtime error */
#define PQP_ERR_NOMEM (-5)        /* host or device allocation failed */
#define PQP_ERR_NOKERNEL (-6)     /* no kernel variant for this request */

/* Flags */
#define PQP_FLAG_PLAIN 0x1u  /* plain FP64 accumulation and one launch per permanent (the baseline) */
#define PQP_FLAG_FP32 0x2u   /* FP32 row sums and products, FP64 + double-double accumulation */
#define PQP_FLAG_STATIC 0x4u /* batches: one block per work unit instead of the persistent queue */
/* Columns with multiplicity c > 1 are, by default, kept as one column whose
 * row sum is raised to the power c; PQP_FLAG_EXPAND_COLS expands them into c
 * identical columns instead (branch-free product, wider kernel). */
#define PQP_FLAG_EXPAND_COLS 0x8u
#define PQP_FLAG_KEEP_COLS 0x10u
/* Batches: also evaluate d perm / d A for every permanent (pqp_batch_gradients). */
#define PQP_FLAG_GRADIENT 0x20u
/* The planner evaluates perm(A; r, c) as perm(A^T; c, r), with the columns
 * as Gray digits, when that sweep is clearly cheaper (fewer Gray terms, or
 * unit multiplicities on the product side); the two are equal.
 * PQP_FLAG_AS_GIVEN keeps the rows as the digits, as the reference does,
 * for results that round as the reference's orientation does. */
#define PQP_FLAG_AS_GIVEN 0x80u

/* Library version string. */
const char *pqp_version(void);

/* Message of the last failure on the calling thread ("" if none). */
const char *pqp_last_error(void);

/* Error bound of the last successful single-permanent call
 * (pqp_permanent_c128, _c64, _c128_noalloc) on the calling thread.  The
 * floating-point sweeps sum the magnitudes of their terms beside the terms;
 * the bound is 4 (n + 2) u sum_t |term_t| / 2^(n-1), u the unit roundoff of
 * the row sums and products (2^-53, or 2^-24 with FP32 products), plus the
 * final rounding of the result: the forward-error bound of a sum of
 * products whose n factors each carry a relative error of order u.  A row
 * sum that itself cancels can carry more, so this is an a-posteriori
 * estimate, not a proof; the GPU tests check it against the double-double
 * oracle.  For a result of the exact integer path it is the final rounding
 * alone.  NaN after a PQP_FLAG_PLAIN call (the plain kernel keeps no
 * magnitudes).  No counterpart in the reference, whose permanent_cpp
 * returns the value alone. */
double pqp_last_error_bound(void);

/* Device 0 properties: returns PQP_OK and fills the outputs, or
 * PQP_ERR_DEVICE when no HIP device is usable.  `name` may be null. */
int pqp_device_info(int *n_devices, int *cu_count, char *name, int name_len);

/* Test hook, no counterpart in the reference: with after >= 0 the library
 * lets `after` more device or pinned-host allocations through and fails the
 * next one with PQP_ERR_NOMEM (then the hook is off again); after < 0 turns
 * it off.  Process-wide; always returns PQP_OK.  Buffers are cached and
 * grow on demand, so a call allocates only what it does not find. */
int pqp_debug_fail_allocation(int64_t after);

/*
 * One permanent, complex128 in and out.
 * Replaces permanent_cpp<double> (src/permanent.cpp:49-264) as reached from
 * the pybind entry permanent(matrix, rows, cols)
 * (piquasso/_math/permanent.cpp:26-41,76-78): same arguments, same result,
 * empty or all-zero multiplicities give 1 (src/permanent.cpp:106-108), a
 * row/column sum mismatch is PQP_ERR_SUM_MISMATCH.
 *   out    : 2 doubles (re, im), the double-double result rounded to FP64
 *   out_dd : 4 doubles (re_hi, re_lo, im_hi, im_lo) or null
 *   flags  : 0, PQP_FLAG_PLAIN or PQP_FLAG_FP32
 * Integer matrices: without PQP_FLAG_PLAIN, a matrix whose selected entries
 * are integers (|a| <= 2^47, zero imaginary parts) and whose Glynn terms may
 * reach 2^47 (2^24 with PQP_FLAG_FP32) is evaluated by the exact
 * multi-modular path of pqp_permanent_int64: `out` is the exact permanent
 * rounded once to nearest, `out_dd` its two leading doubles.  Below that
 * bound the floating-point sweep is exact already: every product is an
 * exact integer, runs of 64 of them are summed in FP64 below 2^53, and the
 * double-double sums of integers below 2^100 are exact.  Inputs the exact
 * path refuses (PQP_ERR_RANGE) take the floating-point sweep.
 */
int pqp_permanent_c128(const double *matrix, int64_t m_r, int64_t m_c, int64_t ld,
                       const int32_t *rows, const int32_t *cols, uint32_t flags,
                       double *out, double *out_dd);

/*
 * pqp_permanent_c128 without allocating: no hipMalloc, hipFree or pinned
 * host allocation in the call.  No counterpart in the reference: this entry
 * is for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid it).  Same arguments, flags and result as
 * pqp_permanent_c128, from the same kernels and plan (bit-identical),
 * except that integer matrices are not routed to the exact path.  The
 * device memory the call needs (plan tables, per-wave partial sums, result
 * slots) comes from a fixed 8 MiB device array that each GPU holds for the
 * life of the process; that holds permanents of up to 2^39 Gray terms
 * (n = 40 at unit multiplicities), and one needing more returns
 * PQP_ERR_RANGE with the size in the message.  Calls serialise on one lock; the array keeps
 * the last call's contents, and nothing in it is assumed to start at zero.
 * Runs on the null stream with synchronous copies from and to host memory.
 */
int pqp_permanent_c128_noalloc(const double *matrix, int64_t m_r, int64_t m_c, int64_t ld,
                               const int32_t *rows, const int32_t *cols, uint32_t flags,
                               double *out, double *out_dd);

/*
 * One permanent, complex64 in and out.
 * Replaces permanent_cpp<float> (src/permanent.cpp:266-267), the overload
 * the pybind entry picks for complex64 input (piquasso/_math/permanent.cpp:73-75).
 * Row sums and products run in FP32, the accumulation in FP64 +
 * double-double (more accurate than the reference's all-FP32 sum).
 *   out    : 2 floats (re, im)
 *   out_dd : 4 doubles or null
 */
int pqp_permanent_c64(const float *matrix, int64_t m_r, int64_t m_c, int64_t ld,
                      const int32_t *rows, const int32_t *cols, uint32_t flags, float *out,
                      double *out_dd);

/*
 * The exact permanent of an integer matrix with multiplicities.
 * Replaces permanent_cpp<T> (src/permanent.cpp:49-264) for integer-valued
 * input such as the reference's integer case
 * (tests/_math/test_permanent.py:252-282): the same Glynn sum over the same
 * Gray order, taken modulo up to eight primes just below 2^49 on the GPU
 * (FP64 modular multiplies), the integer recovered on the host by the
 * Chinese remainder theorem.  Counts perfect matchings of 0/1 matrices
 * exactly at every size the forward path takes (n <= 64).
 *   matrix : m_r x m_c int64, row-major, leading dimension ld
 *   out    : 1 double, the permanent rounded once to nearest (ties to even)
 *   limbs  : n_limbs >= 8 uint64 words receiving the permanent as a 512-bit
 *            little-endian two's complement integer, or null
 * PQP_ERR_RANGE when a selected entry exceeds 2^47 in magnitude, a
 * column's sum_i r_i |a_ij| exceeds 2^48, prod_i C(r_i, r_i / 2) exceeds
 * 2^52, or the bound prod_i (sum_j c_j |a_ij|)^{r_i} exceeds 2^390.
 */
int pqp_permanent_int64(const int64_t *matrix, int64_t m_r, int64_t m_c, int64_t ld,
                        const int32_t *rows, const int32_t *cols, double *out, uint64_t *limbs,
                        int32_t n_limbs);

/*
 * Batches: many permanents of uneven size, one launch per column-width
 * bucket (the buckets run concurrently on forked streams), then one finalize
 * launch; the reference has no batched entry, and each item's value equals
 * pqp_permanent_c128 on that item.  Each bucket's launch is a persistent,
 * work-stealing kernel unless flags has PQP_FLAG_STATIC, which launches one
 * block per work unit in decreasing cost order instead (faster on MI355X
 * for the measured batches; the default of the Python layer's autotuned
 * configuration).  The batch owns its device buffers and stream; inputs are
 * copied in at creation.
 *   matrices : count pointers, complex128 (is_c64 = 0) or complex64 (1)
 *   m_r, m_c, ld : count entries each
 *   rows, cols   : count pointers to int32 multiplicities
 *   target_inner : inner Gray states per lane (0 = the planner's cost model)
 *   total_units  : wave work units to spread over the batch in proportion to
 *                  each permanent's cost (0 = 32 per compute unit; gradient
 *                  batches 8 per compute unit)
 */
typedef struct pqp_batch pqp_batch;

int pqp_batch_create(int64_t count, const void *const *matrices, int is_c64, const int64_t *m_r,
                     const int64_t *m_c, const int64_t *ld, const int32_t *const *rows,
                     const int32_t *const *cols, uint32_t flags, int64_t target_inner,
                     int64_t total_units, pqp_batch **out);

/* Run modes for pqp_batch_run / pqp_batch_time */
#define PQP_RUN_EAGER 0  /* the sweep kernel(s) + finalize, launched directly on the batch stream */
#define PQP_RUN_GRAPH 1  /* the same launches captured once in a hipGraph and replayed */
#define PQP_RUN_PLAIN 2  /* plain FP64 kernel, one launch (+ finalize) per permanent */
#define PQP_RUN_KERNEL 3 /* timing only: the sweep kernel(s) alone, finalize outside the events */

/* Enqueue one evaluation of the whole batch (asynchronous). */
int pqp_batch_run(pqp_batch *b, int mode);

/* Wait for the batch stream. */
int pqp_batch_sync(pqp_batch *b);

/* Copy the results of the last run: out = 2 doubles per item, out_dd = 4
 * doubles per item (may be null).  Synchronises first. */
int pqp_batch_results(pqp_batch *b, double *out, double *out_dd);

/* Error bounds of the last evaluation, one per item: pqp_last_error_bound's
 * estimate of |result k - perm k| from the term magnitudes the forward
 * sweeps sum; NaN for every item when the last evaluation kept none (mode
 * PQP_RUN_PLAIN without FP32 products, gradient batches of at most 32
 * particles). */
int pqp_batch_error_bounds(pqp_batch *b, double *bounds);

/* Time `reps` evaluations with HIP events recorded on the batch stream:
 * ms[k] = elapsed milliseconds of evaluation k. */
int pqp_batch_time(pqp_batch *b, int mode, int reps, float *ms);

/* Time `reps` >= 1 evaluations (mode PQP_RUN_GRAPH or PQP_RUN_EAGER) issued
 * back to back between one pair of HIP events on the batch stream, after
 * the stream has drained: *ms = elapsed milliseconds of all of them, every
 * kernel of an evaluation included (ms / reps is the period of one).  A
 * measurement entry with no counterpart in the reference (its benchmarks
 * time host wall clock, benchmarks/permanent_jax_ffi_cuda.py:228-259). */
int pqp_batch_time_span(pqp_batch *b, int mode, int reps, float *ms);

/* Planning statistics: algorithmic FLOP (sum of T * (2 m_c + 6 (n-1) + 4)),
 * Gray terms, wave work units, device arena size, and lane terms (sum over
 * units of 64 lanes x inner states: terms / lane_terms is the fraction of
 * lane-steps doing useful work).  Any pointer may be null. */
int pqp_batch_stats(const pqp_batch *b, double *flops, int64_t *terms, int64_t *units,
                    int64_t *arena_bytes, int64_t *lane_terms);

/* Gradient batches (PQP_FLAG_GRADIENT): grads[k] receives the m_r x m_c
 * complex matrix d perm_k / d A_k (row-major, interleaved) of the last run. */
int pqp_batch_gradients(pqp_batch *b, double *const *grads);

void pqp_batch_destroy(pqp_batch *b);

/*
 * Column minors, complex128: out[l] = perm(A; rows, cols - e_l) for every
 * column l with cols[l] > 0, where sum(cols) = sum(rows) + 1, all from one
 * Gray sweep per work unit (prefix / suffix products).  Replaces
 * permanent_laplace_cpp<double> (src/permanent_laplace.cpp:43-240) as reached
 * from the pybind entry permanent_laplace(matrix, rows, cols)
 * (piquasso/_math/permanent.cpp:43-58,82-84).  For a column with cols[l] = 0
 * the value is the uncut sum (the reference computes the same there); with
 * all row multiplicities 0 every entry is 1.  out: 2 * m_c doubles.  Up to
 * 32 columns of non-zero multiplicity one sweep gives every minor; wider
 * problems (up to 64 particles) run one forward sweep per column.
 */
int pqp_laplace_c128(const double *matrix, int64_t m_r, int64_t m_c, int64_t ld,
                     const int32_t *rows, const int32_t *cols, double *out);

/*
 * Column minors of `count` problems from one launch: outs[k] (2 * m_c[k]
 * doubles) receives what pqp_laplace_c128 writes for problem k.  The batched
 * form of permanent_laplace (piquasso/_math/permanent.cpp:43-58,82-84) for
 * callers that need many at once; the reference calls the single entry once
 * per shot and photon step of Clifford & Clifford sampling
 * (piquasso/_simulators/passive/sampling.py:723-749), and running the shots
 * in lockstep turns each step into one such batch.
 */
int pqp_laplace_batch_c128(int64_t count, const double *const *matrices, const int64_t *m_r,
                           const int64_t *m_c, const int64_t *ld, const int32_t *const *rows,
                           const int32_t *const *cols, double *const *outs);

/*
 * Tensor permanent of n uniformly partially distinguishable bosons with
 * particle overlap x, the quantity behind piquasso's detection probability
 * of such inputs (_calculate_tensor_permanent,
 * piquasso/_simulators/passive/probabilities.py:193-237, called from
 * get_partially_distinguishable_detection_probability, :156-190):
 *   T = sum over column subsets s, t of (-1)^(|s|+|t|)
 *       prod_i ( x R_s[i] conj(R_t[i]) + (1 - x) Q_{s&t}[i] ),
 *   R_s[i] = sum_{c in s} A[i, c],  Q_u[i] = sum_{c in u} |A[i, c]|^2,
 * for the n x n complex matrix A (row-major interleaved, leading dimension
 * ld >= n; the caller passes the multiplicity-expanded interferometer).
 * The kernels sum the same value over pairs of sign vectors (Glynn's form,
 * 4^(n-1) pairs, much less cancellation than the subsets).
 * T is real (the reference returns it in a complex dtype with a zero
 * imaginary part): *out = T, summed in double-double and rounded.  The
 * alternating sum cancels heavily; *abs_sum (may be null) receives the sum
 * of the absolute values of the summed terms, so that about
 * 8 (n + 2) 2^-53 abs_sum bounds the error of T.  n = 0 gives 1; n <= 20.
 */
int pqp_tensor_permanent_c128(const double *matrix, int64_t n, int64_t ld, double overlap,
                              double *out, double *abs_sum);

/* The same for `count` matrices in one launch per matrix order: out[k] and
 * abs_sum[k] (may be null). */
int pqp_tensor_permanent_batch_c128(int64_t count, const double *const *matrices, const int64_t *n,
                                    const int64_t *ld, const double *overlap, double *out,
                                    double *abs_sum);

/*
 * Coefficients of the permanent of a matrix polynomial,
 *   out[q] = [x^e_q] perm(B + sum_{m < n_vars} x_m C_m)   (2 doubles: re, im)
 * for n x n complex matrices B (b0) and C_0..C_{n_vars-1} (bm: n_vars
 * matrices one after another; row-major, interleaved complex) and the
 * exponent rows e_q (exponents: n_out x n_vars int32, row-major).  This is
 * the subset loop of get_lossy_partially_distinguishable_detection_probabilities
 * (piquasso/_simulators/passive/probabilities.py:341-382): Ryser's formula
 * over the 2^n - 1 non-empty column subsets, each subset's product of row
 * factors kept as a polynomial; here on the monomials that divide some e_q
 * (the reference keeps the box of the largest exponent per variable; the
 * kept coefficients are the same).  Rows with total degree above n give 0.
 * info (may be null, 6 entries): monomials, links, workgroups, 1 if the
 * polynomial was kept in LDS, algorithmic FP64 flops per subset, subsets.
 * n <= 24, at most 255 variables with a non-zero exponent, n x (those
 * variables + 1) <= 3072 and at most 2^23 monomials (PQP_ERR_RANGE beyond).
 */
int pqp_perm_poly_coefficients_c128(int64_t n, int64_t n_vars, const double *b0, const double *bm,
                                    int64_t n_out, const int32_t *exponents, double *out,
                                    int64_t *info);

/*
 * Planning only (no device involved): the monomial closure
 * pqp_perm_poly_coefficients_c128 would use for these exponent rows, for
 * inspection by tests of the planner.  info (6 entries): monomials, links,
 * variables in use, largest kept total degree, algorithmic FP64 flops per
 * subset, exponent rows kept (total degree <= n).  No counterpart in the
 * reference.
 */
int pqp_perm_poly_plan_info(int64_t n, int64_t n_vars, int64_t n_out, const int32_t *exponents,
                            int64_t *info);

/*
 * Gradients of `count` permanents: grads[k] (m_r[k] x m_c[k] complex,
 * row-major, interleaved) = d perm(A_k; r_k, c_k) / d (A_k)_ij
 * = r_i c_j perm(A_k; r_k - e_i, c_k - e_j) (0 where r_i or c_j is 0), the
 * result of grad_perm (src/permanent.cpp:271-300) and of the JAX extension's
 * backward (src/jax_perm/jax_perm_core.cpp:141-208), here without the
 * square-matrix restriction.  Up to 32 particles one Gray sweep per
 * permanent gives the permanent and every entry (perm_grad.hip); larger
 * permanents (up to 64) take one column-minor sweep per row of non-zero
 * multiplicity, or, beyond 32 distinct columns, one forward sweep per
 * (row, column) minor, all in one launch.  perms (2 * count doubles, may be
 * null) receives the permanents.
 */
int pqp_grad_batch_c128(int64_t count, const double *const *matrices, const int64_t *m_r,
                        const int64_t *m_c, const int64_t *ld, const int32_t *const *rows,
                        const int32_t *const *cols, double *perms, double *const *grads);

/*
 * Planning only (no device involved): the device arena the kernels would
 * read for one permanent, for inspection by tests of the planner.  Layout:
 * the ItemDesc record at offset 0, then the work units and tables at the
 * offsets it names (piquasso_amd/csrc/pqp_types.hpp).  With buf == null only
 * *needed_bytes is written.  flags & PQP_FLAG_GRADIENT plans the column-minor
 * sweep of pqp_laplace_c128 instead, flags & PQP_PLAN_GRADIENT_SWEEP the
 * single-sweep gradient item of pqp_grad_batch_c128.  No counterpart in the
 * reference.
 */
#define PQP_PLAN_GRADIENT_SWEEP 0x40u
int pqp_plan_image(const double *matrix, int64_t m_r, int64_t m_c, int64_t ld,
                   const int32_t *rows, const int32_t *cols, uint32_t flags,
                   int64_t target_inner, int64_t min_units, void *buf, int64_t buf_bytes,
                   int64_t *needed_bytes);

#ifdef __cplusplus
}
#endif

#endif /* PQP_PERMANENT_H */
