// Helpers of the C ABI shared between capi.cpp and the translation units
// that define further entry points (noalloc.hip): the thread-local error
// message, the device check, and the planning options of single-permanent
// calls, so that every single-permanent entry plans the same way.
#pragma once

#include <cstdint>
#include <string>

#include "plan.hpp"
#include "pqp_types.hpp"

namespace pqp {

// Record `msg` as the calling thread's pqp_last_error() and return `code`.
int capi_fail(int code, const std::string &msg);

// PQP_OK when a HIP device is usable, else PQP_ERR_DEVICE with the message.
int capi_require_device();

// Gray terms of one permanent, prod (rho_i + 1) after the prepend (the
// first minimal non-zero row loses one), as the planner counts them;
// `in.rows` must be valid (negative entries count as 0).
double gray_terms(const PermInput &in);

// Single permanents of at most this many Gray terms and at most 16 columns
// run in one launch of one block (perm_single_block_kernel), planned for at
// most kSingleBlockUnits waves, kSingleBlockUnitsSmall up to
// kSingleBlockSmallTerms terms.  Each lane's start state costs about as much
// as a few Gray steps and is latency-bound, so fewer, longer units finish
// sooner (kernel time, MI355X: n = 12 at 4 / 8 / 16 waves 7.9 / 8.3 / 10.2
// us; n = 14 14.2 / 13.6 / 15.4 us); larger ones spread over the device.
constexpr double kSingleBlockMaxTerms = 16384.0;
constexpr int kSingleBlockUnits = 8;
constexpr double kSingleBlockSmallTerms = 2048.0;
constexpr int kSingleBlockUnitsSmall = 4;

// Planning options of a single-permanent call of `in` with these
// PQP_FLAG_* flags: about 1.4 sqrt(T) waves, capped at 16 for the
// one-block form and per compute unit otherwise; one-wave permanents write
// their result from the sweep kernel.
PlanOptions single_plan_options(uint32_t flags, const PermInput &in);

// Whether the one-item plan runs as one block (perm_single_block_kernel):
// compensated (no PQP_FLAG_PLAIN), at least one unit, at most 16 columns
// and kSingleBlockThreads / 64 units; `bucket` receives its width bucket.
bool single_block_eligible(const Plan &plan, uint32_t flags, int *bucket);

// The one-block launch's arguments for the one-item plan: `a` with the
// rows of B copied into the arguments when they fit (else read from the
// arena), the arena's whole 64-byte lines, and the done slot and value.
void single_block_args(const Plan &plan, const LaunchArgs &a, uint64_t *done, uint64_t seq,
                       SingleArgs &args);

// Bound on |computed - perm| of a floating-point sweep from its term
// magnitudes m = sum_t |term_t| / 2^(n-1) (as the kernels accumulate them):
// 4 (n + 2) u m, u the unit roundoff of the row sums and products (2^-53,
// or 2^-24 with FP32 products), plus the rounding of the double-double
// value (re, im) to FP64.  NaN when m is NaN (the plain kernel keeps none).
double sweep_error_bound(int n, bool fp32, double m, double re, double im);

// The calling thread's pqp_last_error_bound().
void capi_set_last_bound(double bound);

}  // namespace pqp
