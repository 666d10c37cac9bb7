// Host-side planning of permanent launches: validation, row/column
// compaction, the reference's min-row prepend, the inner/top split of the
// Gray digits, the wave-uniform step and weight tables, and the work-unit
// list.  No floating-point work on the matrix entries happens here: the
// tables only hold copies (and exact doublings) of matrix entries and exact
// integer weights; every sum and product of the permanent runs on the GPU.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "pqp_types.hpp"

namespace pqp {

// Status codes of the C ABI (include/pqp_permanent.h).
enum Status : int {
  kOk = 0,
  kErrSumMismatch = -1,  // sum(rows) != sum(cols)
  kErrInvalid = -2,      // shapes, negative multiplicities, null pointers
  kErrRange = -3,        // term count or widths beyond the supported range
  kErrDevice = -4,       // HIP runtime error / no device
  kErrNoMem = -5,        // allocation failure
  kErrNoKernel = -6,     // no kernel variant for this width / precision
};

// One permanent as handed over by the caller: row-major interleaved
// complex matrix with leading dimension `ld` (in complex elements).
struct PermInput {
  const void *a;        // double* (complex128) or float* (complex64)
  bool a_is_float;
  int64_t m_r, m_c, ld;
  const int32_t *rows;
  const int32_t *cols;
  // forward items only: the columns may sum to the rows' sum + col_excess
  // (1 for the uncut product sum_t s(t) prod_j cs_j^{c_j} that the
  // reference's column-minor loop gives a zero-multiplicity column)
  int32_t col_excess = 0;
  // set by the planner only: this is the transposed view of the caller's
  // input (m_r / m_c and rows / cols swapped; entry (i, j) is the caller's
  // (j, i), `ld` still the caller's row stride)
  bool transposed = false;
};

struct PlanOptions {
  bool fp32_products = false;   // FP32 row sums and products (tables in float)
  int64_t target_inner = 0;     // requested inner states per lane (0 = cost model)
  int64_t min_units = 0;        // split every permanent into at least this many waves
  bool direct_out = false;      // single-unit items write their result directly (kDirectOut)
  int64_t auto_units_cap = 0;   // > 0: at least ~1.4 sqrt(T) waves per permanent, at most this many
                                //      (measured optimum for one permanent on MI355X, n = 16..32)
  int expand_cols = 2;          // columns with multiplicity c > 1: 0 keep (power c),
                                // 1 expand into c columns, 2 planner's choice (keep:
                                // expansion widens the kernel and lowers occupancy)
  bool minors = false;          // column-minor items: sum(c) = sum(r) + 1, outputs
                                // perm(A; r, c - e_l) for every column l plus the
                                // uncut product (m_c + 1 values per item)
  int64_t total_units = 0;      // spread about this many waves over the batch, in
                                // proportion to each permanent's cost (0 = off)
  bool pad8 = false;            // kernel width padded to a multiple of 8 (exact integer path)
  bool grad = false;            // gradient items: the permanent and d perm / d B for every
                                // row of B from one sweep (perm_grad.hip); outputs
                                // 1 + rows_B * m_c values per item
  bool orient = true;           // forward items: plan perm(A^T; c, r) instead of
                                // perm(A; r, c) when its sweep is clearly cheaper
};

// Per-item bookkeeping kept on the host.
struct ItemInfo {
  int64_t terms = 0;            // T = prod (rho_i + 1)
  double flops = 0.0;           // T * (2 m_c + 6 (n - 1) + 4)
  double unit_cost = 0.0;       // modelled cost of one wave unit (LPT key)
  bool trivial_one = false;     // n == 0: permanent is 1 by definition
  int64_t out_offset = 0;       // first output slot (complex values)
  std::vector<int64_t> colsrc;  // source column of each kernel column
  std::vector<int64_t> rowsrc;  // source row of each row of B (prepended, inner, top)
  bool transposed = false;      // planned as perm(A^T; c, r): rowsrc are columns of A, colsrc rows
};

struct Plan {
  std::vector<char> arena;          // host image of the device arena
  int64_t items_off = 0;            // ItemDesc array offset in the arena
  int64_t units_off = 0;            // WorkUnit array offset
  std::vector<ItemDesc> items;
  std::vector<ItemInfo> info;
  std::vector<WorkUnit> units;      // LPT order within each width bucket
  std::vector<int> bucket_of_unit;  // kernel bucket id per unit
  // contiguous unit ranges per kernel bucket: {bucket id, first unit, count}
  struct Range { int bucket; int32_t first; int32_t count; double cost; };
  std::vector<Range> ranges;
  int64_t n_partials = 0;     // partial slots (4 doubles each)
  int32_t max_units = 0;      // largest number of work units of one item
  bool all_direct = false;    // every item writes its own result: no finalize launch
  int64_t n_out = 0;          // output slots (complex values)
  bool minors = false;        // column-minor items (perm_minors.hip kernels)
  bool grad = false;          // gradient items (perm_grad.hip kernels)
  bool fp32 = false;
  double total_flops = 0.0;
  int64_t total_terms = 0;
};

// Kernel width buckets.  A bucket's kernel handles every mcp in [lo, hi]
// with exact-width code; permanents wider than 32 columns are padded to the
// next multiple of 8 (padding columns hold 1 + 0i in the prepended row and
// 0 elsewhere, so they multiply the product by exactly 1).
struct Bucket { int id; int lo; int hi; };
const std::vector<Bucket> &buckets();
int bucket_for_width(int mc, bool fp32, int *mcp_out);

// Rewrite the rows of B in a forward plan's arena for new matrices of the
// same shapes and multiplicities as the plan's inputs (every other table
// depends on the multiplicities alone); the byte range of item k's rows is
// [items[k].b_off, + info[k].rowsrc.size() * mcp * 2 * (4 or 8)).
void plan_refresh_matrix(Plan &plan, const std::vector<PermInput> &inputs);

int plan_build(const std::vector<PermInput> &inputs, const PlanOptions &opt,
               Plan &plan, std::string &err);

// Reflected mixed-radix Gray sequence of `radix` (digit 0 least significant),
// starting from the all-zero state: for each of the prod(radix) - 1 steps the
// digit that moves and its new value.  Same order as the reference counter
// (src/n_aryGrayCodeCounter.hpp:209-254) started at offset 0.
void gray_steps(const std::vector<int> &radix, std::vector<int> &digit,
                std::vector<int> &value);

}  // namespace pqp
