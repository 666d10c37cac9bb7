// Data layout shared by the host planner (plan.cpp) and the HIP kernels.
//
// A launch evaluates a batch of permanents.  The planner turns every
// permanent into a Glynn sweep over a reflected mixed-radix Gray code
// (the method of src/permanent.cpp:49-264, src/n_aryGrayCodeCounter.hpp):
//
//   perm = 2^-(n-1) * sum_g (-1)^{|g|} prod_i C(rho_i, g_i)
//                    * prod_j ( b_0j + sum_i (rho_i - 2 g_i) b_ij )^{c_j}
//
// and splits the digit set into "inner" digits, swept in the same order by
// all 64 lanes of a wavefront, and "top" digits, one top state per lane.
// All lanes of a wave therefore change the same row with the same sign at
// the same step, so row data, step descriptors and weights are wave-uniform
// (scalar loads), and only the m_c complex row sums live in per-lane VGPRs.
//
// Everything the kernels read lives in one device "arena":
//   [ItemDesc x items][WorkUnit x units][per-item tables][shared tables]
// Offsets below are byte offsets into the arena.
#pragma once

#include <cstdint>

namespace pqp {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kSingleBlockThreads = 1024;  // one-launch single permanents: 16 waves in one block
constexpr int kMaxCols = 64;       // widest row-sum vector a kernel keeps in VGPRs
constexpr int kMaxDigits = 63;     // Gray digits (rows) per permanent
constexpr int kDdFold = 64;        // FP64 run length folded into dd accumulators
constexpr int kPowGroupedMaxCols = 24;  // widest kernel with the grouped column powers

enum ItemFlags : int {
  kColsAllOnes = 1,   // every column multiplicity is 1: plain product
  kTopBinary = 2,     // every top digit has radix 2 (rho = 1): bit-level Gray
  kMinors = 4,        // column-minor item: perm(A; r, c - e_l) for every column l
  kPowGrouped = 8,    // column powers from prefix-product snapshots (masks at cpow_off)
  kDirectOut = 16,    // one work unit: the unit writes the scaled result itself (no finalize)
  kGrad = 32,         // gradient item: the permanent and d perm / d B for every row of B
};

constexpr int kGradMaxInner = 16;  // inner digits of a gradient item (per-row LDS accumulators)
constexpr int kIntMaxPrimes = 8;   // moduli of the exact integer permanent

// Exact integer permanents (perm_int.hip): the arena holds, at mod_off,
// kIntMaxPrimes primes, their reciprocals and 2^-(n-1) modulo each
// (doubles); products are formed in exact groups of `group` columns.
struct IntArgs {
  int64_t mod_off;
  int32_t n_primes;
  int32_t group;
};

// Doubles per work unit of a gradient item: the permanent's dd partial (4),
// then one FP64 value per (row of B, real component) with the components of
// a row laid out [re_0..re_{mcp-1}, im_0..im_{mcp-1}]; rounded up to whole
// 4-double partial slots.  (constexpr: usable in host and device code.)
constexpr int64_t grad_unit_stride(int rows_B, int mcp) {
  return (4 + static_cast<int64_t>(rows_B) * 2 * mcp + 3) / 4 * 4;
}

// One permanent; read by the kernels with scalar loads (the item index is
// wave-uniform).
struct alignas(16) ItemDesc {
  int64_t b_off;        // B: (1 + L + Dt) rows x mcp complex, row 0 = prepended row,
                        //    rows 1..L inner digits, rows L+1.. top digits (unscaled)
  int64_t b2_off;       // the inner rows of B (row 1 of B): the row added at an inner
                        //    step is (+-2) x row k of this block
  int64_t step_off;     // inner table, S_in + 1 entries of InnerEntry (shared by items
                        //    with the same inner multiplicity profile)
  int64_t coef_off;     // the same table (kept for the layout)
  int64_t alpha_off;    // double per inner digit (L): rho_k (inner digits start at g = 0)
  int64_t top_off;      // double x 4 per top digit: {1 / P_t, radix, rho, 0} with place
                        //    value P_t = prod_{t' < t} radix_t'
  int64_t cpow_off;     // uint64 x 3 snapshot masks (bit j of mask k: c_j - c_{j+1} > k),
                        //    then int32 per r = 2..cmax: number of columns with c_j >= r
  int64_t cmul_off;     // int32 per kernel column: its multiplicity c_j (minors kernel)
  int64_t S_in;         // inner states per lane
  int64_t U;            // top states (lanes of work)
  int64_t partial_base; // first partial slot of this permanent
  int32_t units;        // waves of work: ceil(U / 64)
  int32_t mcp;          // compile-time width of the kernel variant (>= mc)
  int32_t mc;           // columns with non-zero multiplicity (sorted by c_j, descending)
  int32_t n;            // sum of multiplicities
  int32_t L;            // inner digits
  int32_t Dt;           // top digits
  int32_t flags;        // ItemFlags
  int32_t cmax;         // largest column multiplicity
  int32_t out_index;    // result slot (minors: first of mc + 1 consecutive slots)
  int32_t pad0;
};
static_assert(sizeof(ItemDesc) == 128, "ItemDesc layout");

// Inner Gray state v: its coefficient (-1)^v prod_inner C(rho, g), and the
// update applied after it: row sums += sign2 * (inner row `row` of B), with
// sign2 = +2 when the digit moves down, -2 when it moves up (entries past
// the last step are padding: sign2 0, row 0, coefficient 0).
struct alignas(16) InnerEntry {
  double coef;
  float sign2;
  uint32_t row;
};
static_assert(sizeof(InnerEntry) == 16, "InnerEntry layout");

// A unit of work: one wavefront sweeps top states [64*local, 64*local + 64)
// of permanent `item`.
struct WorkUnit {
  int32_t item;
  int32_t local;
};

// Kernel parameters shared by every launch.
struct LaunchArgs {
  const char *arena;
  const ItemDesc *items;
  const WorkUnit *units;
  double *partials;     // 4 doubles (re.hi, re.lo, im.hi, im.lo) per unit
  double *out;          // 2 doubles per permanent
  double *out_dd;       // 4 doubles per permanent (may be null)
  int *queue_head;      // dynamic work-queue head (persistent launches)
  int32_t n_units;
  int32_t n_items;
  // Forward sweeps with double-double accumulation also sum the magnitudes
  // of their terms, sum_t |coef| |w_top| (|Re p| + |Im p|) >= sum_t |term|:
  // one double per unit (part_abs) and, scaled by 2^-(n-1), one per
  // permanent (out_abs).  Null where a launch does not keep them.
  double *part_abs = nullptr;
  double *out_abs = nullptr;
  // One-item plans only: the rows of B of item 0 at this address instead of
  // arena + b_off (the copy a one-block launch carries in its kernel
  // arguments); null everywhere else.
  const char *b_rows = nullptr;
  // One-block launches of items whose top digits are all binary: per unit,
  // the row sums of its first state less the lowest kLaneDigits top digits
  // (the prepended row, the inner rows at their start coefficients and the
  // top digits that are the same for all 64 lanes of the unit), m_cp values
  // of the item's real type per unit, computed on the host; null elsewhere.
  const char *unit_base = nullptr;
  // One-item static launches of a single permanent: each block counts itself
  // at last_counter after writing its partial; the block that counts last
  // sums the item's partials, writes the result, resets the counter and
  // publishes seq at done (mapped host memory).  Null elsewhere.
  int *last_counter = nullptr;
  uint64_t *done = nullptr;
  uint64_t seq = 0;
};

// The queue-head slot that holds a single permanent's block counter, and
// the widest kernel that finalizes in its last block.
constexpr int kLastCounterSlot = 63;
constexpr int kLastFinalizeMaxWidth = 16;

// Top digits that differ between the 64 lanes of a unit (lane l of unit w
// holds top state 64 w + l): the reflected binary digits 0..5.
constexpr int kLaneDigits = 6;

// Kernel arguments of the one-block single-permanent launch: the launch
// arguments, the done slot it publishes to, and up to kSingleArgBBytes of
// the rows of B (when they fit), which then need no copy to the device.
// The whole struct is the kernel's only argument, so it sits at offset 0
// of the kernel-argument segment (the first explicit argument's place).
constexpr int kSingleArgBBytes = 3840;
struct SingleArgs {
  LaunchArgs a;
  int32_t arena_lines;  // 64-byte lines of the arena the block touches first
  int32_t b_in_args;    // 1: rows of B in b[] (a.b_rows is set in the kernel)
  int32_t base_in_args; // 1: per-unit start rows in b[] at base_off (a.unit_base is set in the kernel)
  int32_t base_off;     // byte offset of those rows in b[]
  uint64_t *done;       // host-visible slot receiving seq
  uint64_t seq;
  alignas(64) char b[kSingleArgBBytes];
};

}  // namespace pqp
