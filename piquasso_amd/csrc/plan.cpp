// Host-side planning for the permanent kernels (see plan.hpp).
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace pqp {

namespace {

constexpr int64_t kMaxTerms = int64_t(1) << 62;
constexpr int64_t kMaxInner = int64_t(1) << 16;   // step/coef table length cap
constexpr int kXcds = 8;                           // MI355X: 8 XCDs, one L2 each
constexpr int64_t kGradMaxStates = 128;           // inner states per lane of a gradient unit
constexpr int kExpandAutoMax = 0;                 // auto mode: never (measured slower)

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Stable insertion sort for the planner's short per-item lists (at most 64
// columns or digits): std::stable_sort takes a temporary buffer from the
// heap on every call, which dominated planning batches of small items.
template <typename T, typename Less>
void small_stable_sort(std::vector<T> &v, Less less) {
  for (size_t i = 1; i < v.size(); ++i) {
    T x = v[i];
    size_t j = i;
    for (; j > 0 && less(x, v[j - 1]); --j) v[j] = v[j - 1];
    v[j] = x;
  }
}

int64_t binom64(int n, int k) {
  if (k < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Append `bytes` of zeroed space to the arena at 64-byte alignment: table
// starts on scalar-cache lines (no kernel load needs more; 256 bytes doubled
// the arena of a batch of small permanents).
int64_t arena_alloc(std::vector<char> &arena, int64_t bytes) {
  int64_t off = align_up(static_cast<int64_t>(arena.size()), 64);
  arena.resize(static_cast<size_t>(off + std::max<int64_t>(bytes, 1)), 0);
  return off;
}

template <typename T>
T *at(std::vector<char> &arena, int64_t off) {
  return reinterpret_cast<T *>(arena.data() + off);
}

// One digit of the Gray code: a matrix row (compacted columns) and its
// multiplicity rho >= 1 (radix rho + 1).
struct Digit {
  int64_t src_row;
  int rho;
};

// Shared inner tables, keyed by the inner multiplicity profile.
struct InnerTables {
  int64_t step_off = 0;
  int64_t coef_off = 0;
};

}  // namespace

const std::vector<Bucket> &buckets() {
  static const std::vector<Bucket> b = {
      {0, 1, 8},   {1, 9, 16},  {2, 17, 24}, {3, 25, 32},
      {4, 40, 40}, {5, 48, 48}, {6, 56, 56}, {7, 64, 64},
  };
  return b;
}

int bucket_for_width(int mc, bool /*fp32*/, int *mcp_out) {
  if (mc < 1 || mc > kMaxCols) return -1;
  int mcp = mc <= 32 ? mc : static_cast<int>(align_up(mc, 8));
  for (const Bucket &b : buckets())
    if (mcp >= b.lo && mcp <= b.hi) {
      *mcp_out = mcp;
      return b.id;
    }
  return -1;
}

void gray_steps(const std::vector<int> &radix, std::vector<int> &digit,
                std::vector<int> &value) {
  const int nd = static_cast<int>(radix.size());
  int64_t total = 1;
  for (int r : radix) total *= r;
  digit.assign(static_cast<size_t>(std::max<int64_t>(total - 1, 0)), 0);
  value.assign(digit.size(), 0);
  std::vector<int> ctr(nd, 0), gray(nd, 0);
  for (int64_t s = 0; s + 1 < total; ++s) {
    // increment the mixed-radix counter with carry
    int idx = 0;
    while (ctr[idx] == radix[idx] - 1) ctr[idx++] = 0;
    ctr[idx]++;
    // top-down: Gray digit = counter digit, reflected when the parity of the
    // higher Gray digits is odd; the first mismatch from the top is the move
    int parity = 0;
    for (int i = nd - 1; i >= 0; --i) {
      int v = parity ? radix[i] - 1 - ctr[i] : ctr[i];
      parity ^= (v & 1);
      if (v != gray[i]) {
        digit[s] = i;
        value[s] = v;
        gray[i] = v;
        break;
      }
    }
  }
}

namespace {

// The caller's row stride covers the view's rows (its columns when the
// view is transposed).
bool stride_ok(const PermInput &in) {
  return in.transposed ? (in.m_c == 0 || in.ld >= in.m_r) : (in.m_r == 0 || in.ld >= in.m_c);
}

// Modelled cost of the Glynn sweep with `rows` as Gray digits and `cols`
// as the product side: T = prod (rho_i + 1) after the prepend, times the
// work per term (row updates, column product with grouped powers or the
// count form, accumulation), the grouped form weighted by the lower issue
// rate of its per-column snapshot tests (about 76% VALU issue against 94%
// for unit multiplicities, measured on MI355X).
double sweep_cost(const int32_t *rows, int64_t m_r, const int32_t *cols, int64_t m_c) {
  int minelem = 0;
  int64_t i0 = -1;
  for (int64_t i = 0; i < m_r; ++i)
    if (rows[i] > 0 && (minelem == 0 || rows[i] < minelem)) {
      minelem = rows[i];
      i0 = i;
    }
  double T = 1.0;
  for (int64_t i = 0; i < m_r; ++i) {
    const int rho = rows[i] - (i == i0 ? 1 : 0);
    if (rho > 0) T *= rho + 1;
  }
  int mc = 0, cmax = 0, n = 0;
  for (int64_t j = 0; j < m_c; ++j)
    if (cols[j] > 0) {
      ++mc;
      cmax = std::max(cmax, static_cast<int>(cols[j]));
      n += cols[j];
    }
  if (cmax <= 1) return T * (2.0 * mc + 6.0 * std::max(mc - 1, 0) + 4.0);
  const bool grouped = mc <= kPowGroupedMaxCols && cmax <= 4;
  const double mult = grouped ? mc + cmax - 2 : n - 1;
  return T * (2.0 * mc + 6.0 * mult + 4.0) * 1.25;
}

// Items whose transposed sweep costs less than 0.9 of the given one are
// planned transposed: perm(A; r, c) = perm(A^T; c, r), the permanent of the
// expanded matrix being invariant under transposition.  Invalid items are
// left alone (the planner reports them).
void orient_items(const std::vector<PermInput> &in, std::vector<PermInput> &out, bool &any) {
  any = false;
  out = in;
  for (PermInput &v : out) {
    if (v.col_excess || v.m_r <= 0 || v.m_c <= 0 || !v.a || !v.rows || !v.cols || !stride_ok(v))
      continue;
    int64_t sr = 0, sc = 0;
    bool neg = false;
    for (int64_t i = 0; i < v.m_r; ++i) {
      neg = neg || v.rows[i] < 0;
      sr += v.rows[i];
    }
    for (int64_t j = 0; j < v.m_c; ++j) {
      neg = neg || v.cols[j] < 0;
      sc += v.cols[j];
    }
    if (neg || sr != sc || sr == 0 || sr > kMaxCols) continue;
    if (sweep_cost(v.cols, v.m_c, v.rows, v.m_r) < 0.9 * sweep_cost(v.rows, v.m_r, v.cols, v.m_c)) {
      std::swap(v.m_r, v.m_c);
      std::swap(v.rows, v.cols);
      v.transposed = true;
      any = true;
    }
  }
}

}  // namespace

int plan_build(const std::vector<PermInput> &caller_inputs, const PlanOptions &opt,
               Plan &plan, std::string &err) {
  std::vector<PermInput> oriented;
  bool any_transposed = false;
  if (opt.orient && !opt.minors && !opt.grad) orient_items(caller_inputs, oriented, any_transposed);
  const std::vector<PermInput> &inputs = any_transposed ? oriented : caller_inputs;
  plan = Plan();
  plan.fp32 = opt.fp32_products;
  const int64_t n_items = static_cast<int64_t>(inputs.size());
  const size_t rsz = opt.fp32_products ? sizeof(float) : sizeof(double);

  plan.items.resize(n_items);
  plan.info.resize(n_items);
  // typical item tables are a few KB: reserve once instead of regrowing
  plan.arena.reserve(static_cast<size_t>(n_items) * 8192 + 65536);
  plan.items_off = arena_alloc(plan.arena, n_items * int64_t(sizeof(ItemDesc)));

  std::map<std::vector<int>, InnerTables> inner_cache;
  int64_t out_cursor = 0;
  std::vector<int64_t> item_units(n_items, 0);
  std::vector<int> item_bucket(n_items, -1);
  int64_t partial_cursor = 0;

  // Pass 1: validate every item and estimate its cost, so that large items
  // can be split into a share of the requested total number of work units
  // proportional to their cost (load balance across uneven items).
  std::vector<double> item_cost(n_items, 0.0);
  std::vector<double> item_terms(n_items, 0.0);
  double total_cost = 0.0;
  for (int64_t k = 0; k < n_items; ++k) {
    const PermInput &in = inputs[k];
    if (in.m_r < 0 || in.m_c < 0 || (in.m_r * in.m_c > 0 && !in.a) ||
        (in.m_r > 0 && !in.rows) || (in.m_c > 0 && !in.cols) || !stride_ok(in)) {
      err = "item " + std::to_string(k) + ": invalid shape or null pointer";
      return kErrInvalid;
    }
    int64_t sr = 0, sc = 0, mc = 0;
    for (int64_t i = 0; i < in.m_r; ++i) sr += std::max(in.rows[i], 0);
    for (int64_t j = 0; j < in.m_c; ++j) {
      sc += std::max(in.cols[j], 0);
      mc += in.cols[j] > 0;
    }
    // minors: sum(c) = sum(r) + 1; forward items: sum(c) = sum(r) + col_excess
    const int64_t extra = opt.minors ? 1 : in.col_excess;
    if (sr == 0 || sr + extra != sc || sc > 64) continue;  // reported in pass 2
    // T = prod (rho + 1) with the prepend applied to the first minimal row
    int minelem = 0;
    int64_t i0 = 0;
    for (int64_t i = 0; i < in.m_r; ++i)
      if (minelem == 0 || (in.rows[i] < minelem && in.rows[i] != 0)) {
        minelem = in.rows[i];
        i0 = i;
      }
    double T = 1.0;
    for (int64_t i = 0; i < in.m_r; ++i) {
      const int rho = in.rows[i] - (i == i0 ? 1 : 0);
      if (rho > 0) T *= rho + 1;
    }
    item_cost[k] = opt.minors ? T * (16.0 * mc + 10.0)
                   : opt.grad ? T * (25.0 * mc + 10.0)
                              : T * (2.0 * mc + 4.0 * (sr - 1) + 8.0);
    item_terms[k] = T;
    total_cost += item_cost[k];
  }

  // Pass 2: tables.  The per-item work lists live outside the loop and are
  // cleared per item, so their storage is allocated once per plan.
  std::vector<int64_t> colsel, colsrc, src_rows;
  std::vector<int> colmult, order, inner, top, profile;
  std::vector<Digit> digits;
  std::vector<char> is_inner;
  for (int64_t k = 0; k < n_items; ++k) {
    const PermInput &in = inputs[k];
    ItemDesc d;
    std::memset(&d, 0, sizeof(d));
    d.out_index = static_cast<int32_t>(out_cursor);
    plan.info[k].out_offset = out_cursor;
    int64_t item_min_units = opt.min_units;
    if (opt.auto_units_cap > 0 && item_terms[k] > 1024) {
      // up to 1024 terms one wave does the whole sum and writes the result;
      // above, several waves and a finalize launch are faster (measured on
      // MI355X: n = 13, 4096 terms, 40 -> 33 us per call)
      const double want = std::ceil(1.4 * std::sqrt(static_cast<double>(item_terms[k])));
      item_min_units = std::max<int64_t>(
          item_min_units, std::min<int64_t>(opt.auto_units_cap, static_cast<int64_t>(want)));
    }
    if (opt.total_units > 0 && total_cost > 0.0) {
      const double share = item_cost[k] / total_cost * static_cast<double>(opt.total_units);
      item_min_units = std::max<int64_t>(item_min_units, static_cast<int64_t>(std::ceil(share)));
    }
    if (in.m_r < 0 || in.m_c < 0 || (in.m_r * in.m_c > 0 && !in.a) ||
        (in.m_r > 0 && !in.rows) || (in.m_c > 0 && !in.cols) || !stride_ok(in)) {
      err = "item " + std::to_string(k) + ": invalid shape or null pointer";
      return kErrInvalid;
    }
    int64_t sr = 0, sc = 0;
    for (int64_t i = 0; i < in.m_r; ++i) {
      if (in.rows[i] < 0) {
        err = "item " + std::to_string(k) + ": negative row multiplicity";
        return kErrInvalid;
      }
      sr += in.rows[i];
    }
    for (int64_t j = 0; j < in.m_c; ++j) {
      if (in.cols[j] < 0) {
        err = "item " + std::to_string(k) + ": negative column multiplicity";
        return kErrInvalid;
      }
      sc += in.cols[j];
    }
    if (!opt.minors && (in.col_excess < 0 || in.col_excess > 1 || (opt.grad && in.col_excess))) {
      err = "item " + std::to_string(k) + ": invalid column excess";
      return kErrInvalid;
    }
    if (!opt.minors && sr + in.col_excess != sc) {
      err = "Number of input and output states should be equal";
      return kErrSumMismatch;
    }
    if (opt.minors && sr + 1 != sc) {
      err = "column multiplicities must sum to one more than the row multiplicities";
      return kErrSumMismatch;
    }
    if (sr == 0) {
      // perm of an empty selection is 1 (src/permanent.cpp:106-108); a
      // column-minor item of an empty row set has the minor 1 in its one
      // column of multiplicity 1
      plan.info[k].trivial_one = true;
      d.units = 0;
      d.n = 0;
      d.partial_base = partial_cursor;
      if (opt.grad) d.flags = kGrad;
      if (opt.minors) {
        d.flags = kMinors;
        plan.info[k].colsrc.reserve(static_cast<size_t>(in.m_c));
        for (int64_t j = 0; j < in.m_c; ++j)
          if (in.cols[j] > 0) plan.info[k].colsrc.push_back(j);
        d.mc = static_cast<int32_t>(plan.info[k].colsrc.size());
        out_cursor += d.mc + 1;
      } else {
        out_cursor += 1;
      }
      plan.items[k] = d;
      continue;
    }
    if (sc > 64) {
      err = "item " + std::to_string(k) + ": more than 64 particles";
      return kErrRange;
    }
    const int n = static_cast<int>(sr);

    // --- compaction: columns with c_j > 0, sorted by multiplicity (desc) ---
    colsel.clear();
    for (int64_t j = 0; j < in.m_c; ++j)
      if (in.cols[j] > 0) colsel.push_back(j);
    small_stable_sort(colsel, [&](int64_t x, int64_t y) { return in.cols[x] > in.cols[y]; });
    const int mc_alg = static_cast<int>(colsel.size());  // algorithmic column count
    int cmax_in = 0;
    for (int64_t j : colsel) cmax_in = std::max(cmax_in, static_cast<int>(in.cols[j]));
    // Kernel columns: a column of multiplicity c either stays one column
    // whose row sum is raised to the power c (extra multiplies behind
    // wave-uniform branches), or is expanded into c identical columns of
    // multiplicity 1 (branch-free product, 2 (c - 1) more additions per term).
    bool expand = false;
    if (cmax_in > 1 && !opt.minors && !opt.grad) {
      if (opt.expand_cols == 1) expand = n <= kMaxCols;
      if (opt.expand_cols == 2) expand = n <= kExpandAutoMax;
    }
    // gradient items: the kernel works with unit column multiplicities; a
    // column of multiplicity c becomes c identical columns, whose gradient
    // entries add up to the column's (its c_j factor)
    if (cmax_in > 1 && opt.grad) expand = true;
    colsrc.clear();   // source column of each kernel column
    colmult.clear();  // multiplicity of each kernel column
    for (int64_t j : colsel) {
      if (expand) {
        for (int q = 0; q < in.cols[j]; ++q) {
          colsrc.push_back(j);
          colmult.push_back(1);
        }
      } else {
        colsrc.push_back(j);
        colmult.push_back(in.cols[j]);
      }
    }
    const int mc = static_cast<int>(colsrc.size());
    int mcp = 0;
    int bucket = bucket_for_width(mc, opt.fp32_products, &mcp);
    if (opt.pad8 && bucket >= 0) {
      mcp = static_cast<int>(align_up(mc, 8));
      bucket = mcp / 8 - 1;  // one range per padded width
    }
    if (opt.minors && mcp > 32) {
      err = "item " + std::to_string(k) + ": column minors support at most 32 columns";
      return kErrRange;
    }
    if (opt.grad && mcp > 32) {
      err = "item " + std::to_string(k) + ": single-sweep gradients support at most 32 particles";
      return kErrRange;
    }
    if (bucket < 0) {
      err = "item " + std::to_string(k) + ": " + std::to_string(mc) +
            " columns exceed the widest kernel (64)";
      return kErrRange;
    }

    // --- prepend rule of src/permanent.cpp:54-95: the first row with the
    //     minimal non-zero multiplicity becomes row 0 with multiplicity 1 ---
    int64_t i0 = 0;
    int minelem = 0;
    for (int64_t i = 0; i < in.m_r; ++i)
      if (minelem == 0 || (in.rows[i] < minelem && in.rows[i] != 0)) {
        minelem = in.rows[i];
        i0 = i;
      }
    digits.clear();
    for (int64_t i = 0; i < in.m_r; ++i) {
      int rho = in.rows[i] - (i == i0 ? 1 : 0);
      if (rho > 0) digits.push_back(Digit{i, rho});
    }
    const int D = static_cast<int>(digits.size());
    if (D > kMaxDigits) {
      err = "item " + std::to_string(k) + ": too many Gray digits";
      return kErrRange;
    }
    int64_t T = 1;
    for (const Digit &g : digits) {
      if (T > kMaxTerms / (g.rho + 1)) {
        err = "item " + std::to_string(k) + ": Gray term count exceeds 2^62";
        return kErrRange;
      }
      T *= (g.rho + 1);
    }

    // --- inner/top split: inner digits = smallest radices first ---
    order.resize(D);
    std::iota(order.begin(), order.end(), 0);
    small_stable_sort(order, [&](int x, int y) { return digits[x].rho < digits[y].rho; });
    int cmax = 0;
    for (int c : colmult) cmax = std::max(cmax, c);
    const double step_cost =
        opt.minors ? 16.0 * mc + 10.0 + (cmax > 1 ? 8.0 * (n + 1 - mc) : 0.0)
        : opt.grad ? 25.0 * mc + 10.0 + (cmax > 1 ? 8.0 * (n - mc) : 0.0)
                   : 2.0 * mc + 4.0 * (n - 1) + 2.0 + 6.0 + (cmax > 1 ? 2.0 * (n - mc) : 0.0);
    // per-unit overhead: start state, plus the wave reduction of every
    // column accumulator for column-minor items, and for gradient items the
    // lane reduction of each top row and the writes of every row
    const double unit_extra = opt.minors ? 160.0 * (mc + 1)
                              : opt.grad ? 12.0 * mc * D + 8.0 * (D + 1) + 200.0
                                         : 0.0;
    int best_L = 0;
    double best_total = -1.0;
    // gradient items: a work unit is one wave sweeping S_in states, and the
    // batch's time is at least its longest unit's, so S_in is capped
    // (target_inner, else kGradMaxStates) and the cost model chooses below it
    const int64_t s_cap = opt.grad ? (opt.target_inner > 0 ? opt.target_inner : kGradMaxStates)
                                   : kMaxInner;
    {
      int64_t S = 1;
      for (int L = 0; L <= D; ++L) {
        if (L > 0) S *= (digits[order[L - 1]].rho + 1);
        if (S > kMaxInner || (L > 0 && S > s_cap)) break;
        const int64_t U = T / S;
        const int64_t units = (U + kWave - 1) / kWave;
        const int Dt = D - L;
        const double init_cost = (1.0 + D) * 2.0 * mcp + Dt * 40.0 + 400.0 + unit_extra;
        const double unit_cost = init_cost + static_cast<double>(S) * step_cost;
        const double total = static_cast<double>(units) * unit_cost;
        bool ok = true;
        if (opt.target_inner > 0 && !opt.grad) {
          // largest inner size not above the request
          ok = (S <= opt.target_inner) || L == 0;
          if (ok) {
            best_L = L;
            best_total = total;
          }
          continue;
        }
        if (item_min_units > 0 && units < item_min_units && L > 0) ok = false;
        if (ok && (best_total < 0 || total < best_total)) {
          best_total = total;
          best_L = L;
        }
      }
    }
    const int L = best_L;
    const int Dt = D - L;
    inner.assign(order.begin(), order.begin() + L);
    top.clear();
    {
      is_inner.assign(D, 0);
      for (int x : inner) is_inner[x] = 1;
      for (int x = 0; x < D; ++x)
        if (!is_inner[x]) top.push_back(x);
    }
    int64_t S_in = 1;
    for (int x : inner) S_in *= digits[x].rho + 1;
    const int64_t U = T / S_in;
    if (U >= (int64_t(1) << 50)) {
      err = "item " + std::to_string(k) + ": too many top states";
      return kErrRange;
    }

    // --- tables ---
    const int rows_B = 1 + L + Dt;
    d.b_off = arena_alloc(plan.arena, int64_t(rows_B) * mcp * 2 * rsz);
    // inner rows of B: the sweep adds +-2 times one of them per step
    d.b2_off = d.b_off + int64_t(mcp) * 2 * rsz;
    auto put = [&](int64_t off, int64_t row, int col, double re, double im) {
      int64_t e = (row * mcp + col) * 2;
      if (opt.fp32_products) {
        at<float>(plan.arena, off)[e] = static_cast<float>(re);
        at<float>(plan.arena, off)[e + 1] = static_cast<float>(im);
      } else {
        at<double>(plan.arena, off)[e] = re;
        at<double>(plan.arena, off)[e + 1] = im;
      }
    };
    auto entry = [&](int64_t row, int64_t col, double &re, double &im) {
      const int64_t e = (in.transposed ? col * in.ld + row : row * in.ld + col) * 2;
      if (in.a_is_float) {
        const float *p = static_cast<const float *>(in.a);
        re = p[e];
        im = p[e + 1];
      } else {
        const double *p = static_cast<const double *>(in.a);
        re = p[e];
        im = p[e + 1];
      }
    };
    // B rows: prepended row, inner digit rows, top digit rows
    src_rows.clear();
    src_rows.push_back(i0);
    for (int x : inner) src_rows.push_back(digits[x].src_row);
    for (int x : top) src_rows.push_back(digits[x].src_row);
    for (int r = 0; r < rows_B; ++r) {
      for (int j = 0; j < mcp; ++j) {
        double re = 0.0, im = 0.0;
        if (j < mc)
          entry(src_rows[r], colsrc[j], re, im);
        else if (r == 0)
          re = 1.0;  // padding column: row sum stays exactly 1 + 0i
        put(d.b_off, r, j, re, im);
      }
    }
    // inner start coefficients alpha_k = rho_k (all inner g = 0); gradient
    // items also keep the coefficients of the last inner state: the counter
    // ends at all digits radix - 1, whose reflected Gray digits are, top
    // down, radix - 1 or 0 by the parity of the higher Gray digits
    d.alpha_off = arena_alloc(plan.arena, int64_t(std::max(L, 1)) * (opt.grad ? 16 : 8));
    for (int t = 0; t < L; ++t)
      at<double>(plan.arena, d.alpha_off)[t] = static_cast<double>(digits[inner[t]].rho);
    if (opt.grad) {
      int parity = 0;
      for (int t = L - 1; t >= 0; --t) {
        const int rho = digits[inner[t]].rho;
        const int g = parity ? 0 : rho;
        parity ^= (g & 1);
        at<double>(plan.arena, d.alpha_off)[L + t] = static_cast<double>(rho - 2 * g);
      }
    }
    // shared inner step / coefficient tables
    profile.clear();
    for (int x : inner) profile.push_back(digits[x].rho);
    auto it = inner_cache.find(profile);
    if (it == inner_cache.end()) {
      InnerTables tb;
      std::vector<int> radix;
      for (int r : profile) radix.push_back(r + 1);
      std::vector<int> sd, sv;
      gray_steps(radix, sd, sv);
      // Interleaved inner table, one 16-byte entry per inner state v:
      // {double coefficient, float sign2, uint32 inner row}: after state v
      // the row sums get sign2 x (inner row of B).  S_in + 1 entries: the
      // kernels read one entry ahead, and the entries past the last step are
      // padding (sign2 0, row 0, coefficient 0).
      tb.step_off = arena_alloc(plan.arena, (S_in + 1) * 16);
      tb.coef_off = tb.step_off;
      char *tab = plan.arena.data() + tb.step_off;
      const double one = 1.0;
      std::memcpy(tab, &one, 8);
      std::vector<int> g(L, 0);
      int64_t w = 1;
      for (size_t s = 0; s < sd.size(); ++s) {
        const int t = sd[s];
        const int prev = g[t];
        g[t] = sv[s];
        // moving g up subtracts 2 b_t from the row sums, moving it down adds
        // 2 b_t (doubling is exact)
        const float sign2 = sv[s] > prev ? -2.0f : 2.0f;
        const uint32_t row = static_cast<uint32_t>(t);
        std::memcpy(tab + 16 * s + 8, &sign2, 4);
        std::memcpy(tab + 16 * s + 12, &row, 4);
        w = w / binom64(profile[t], prev) * binom64(profile[t], g[t]);
        const double sgn = ((s + 1) & 1) ? -1.0 : 1.0;
        const double coef = sgn * static_cast<double>(w);
        std::memcpy(tab + 16 * (s + 1), &coef, 8);
      }
      it = inner_cache.emplace(profile, tb).first;
    }
    d.step_off = it->second.step_off;
    d.coef_off = it->second.coef_off;
    // top digits: {1 / place value, radix, rho, 0}
    d.top_off = arena_alloc(plan.arena, int64_t(std::max(Dt, 1)) * 32);
    bool top_binary = true;
    {
      int64_t place = 1;
      for (int t = 0; t < Dt; ++t) {
        const int rho = digits[top[t]].rho;
        double *td = at<double>(plan.arena, d.top_off) + 4 * t;
        td[0] = 1.0 / static_cast<double>(place);
        td[1] = static_cast<double>(rho + 1);
        td[2] = static_cast<double>(rho);
        td[3] = 0.0;
        place *= rho + 1;
        if (rho != 1) top_binary = false;
      }
    }
    // column multiplicity counts: cpow[r-2] = #{j : c_j >= r}, r = 2..cmax
    d.cmul_off = arena_alloc(plan.arena, int64_t(mcp) * 4);
    for (int j = 0; j < mc; ++j) at<int32_t>(plan.arena, d.cmul_off)[j] = colmult[j];
    // Power table at cpow_off: three uint64 snapshot masks, then the counts.
    // With d_j = c_j - c_{j+1} (c_mc = 1) the product prod_j cs_j^{c_j}
    // equals prod_j P_j^{d_j} * P_{mc-1}, P_j the running prefix product over
    // columns 0..j; bit j of mask k is set when d_j > k.  When some d_j > 3
    // the kernels use the counts (cnt_r = #{j : c_j >= r}, r = 2..cmax)
    // instead.
    bool grouped = cmax > 1 && mcp <= kPowGroupedMaxCols;
    uint64_t snap[3] = {0, 0, 0};
    for (int j = 0; j < mc; ++j) {
      const int next = j + 1 < mc ? colmult[j + 1] : 1;
      const int dj = colmult[j] - next;
      if (dj > 3) grouped = false;
      for (int k = 0; k < 3; ++k)
        if (dj > k) snap[k] |= uint64_t(1) << j;
    }
    d.cpow_off = arena_alloc(plan.arena, 24 + int64_t(std::max(cmax - 1, 1)) * 4);
    for (int k = 0; k < 3; ++k) at<uint64_t>(plan.arena, d.cpow_off)[k] = grouped ? snap[k] : 0;
    for (int r = 2; r <= cmax; ++r) {
      int cnt = 0;
      for (int c : colmult) cnt += c >= r;
      at<int32_t>(plan.arena, d.cpow_off + 24)[r - 2] = cnt;
    }

    d.S_in = S_in;
    d.U = U;
    d.units = static_cast<int32_t>((U + kWave - 1) / kWave);
    d.mcp = mcp;
    d.mc = mc;
    d.n = n;
    d.L = L;
    d.Dt = Dt;
    d.cmax = cmax;
    d.flags = (cmax == 1 ? kColsAllOnes : 0) | (top_binary ? kTopBinary : 0) |
              (opt.minors ? kMinors : 0) | (opt.grad ? kGrad : 0) | (grouped ? kPowGrouped : 0) |
              (opt.direct_out && !opt.minors && !opt.grad && d.units == 1 ? kDirectOut : 0);
    d.partial_base = partial_cursor;
    if (opt.grad) {
      partial_cursor += static_cast<int64_t>(d.units) * (grad_unit_stride(rows_B, mcp) / 4);
      out_cursor += 1 + static_cast<int64_t>(rows_B) * mc;
    } else {
      partial_cursor += static_cast<int64_t>(d.units) * (opt.minors ? mcp + 1 : 1);
      out_cursor += opt.minors ? mc + 1 : 1;
    }
    plan.info[k].colsrc.assign(colsrc.begin(), colsrc.end());
    plan.info[k].rowsrc.assign(src_rows.begin(), src_rows.end());
    plan.info[k].transposed = in.transposed;
    plan.items[k] = d;

    ItemInfo &inf = plan.info[k];
    inf.terms = T;
    // minors: row update, prefix / suffix / combine products (3 (m_c - 1)
    // complex multiplies, extra powers 2 (n + 1 - m_c)), m_c + 1 accumulations
    // gradients: row update 2 m_c, exclusion products 18 (n - 1), the
    // accumulation 2 m_c, one row flush 4 m_c, the weight 4 per term
    inf.flops = opt.minors
                    ? static_cast<double>(T) * (2.0 * mc_alg + 6.0 * (3.0 * (mc_alg - 1) +
                                                                   2.0 * (n + 1 - mc_alg)) +
                                                4.0 * (mc_alg + 1))
                : opt.grad ? static_cast<double>(T) * (8.0 * mc_alg + 18.0 * (n - 1) + 4.0)
                           : static_cast<double>(T) * (2.0 * mc_alg + 6.0 * (n - 1) + 4.0);
    if (opt.minors || opt.grad) {
      inf.unit_cost = ((1.0 + D) * 2.0 * mcp + Dt * 40.0 + 400.0) +
                      static_cast<double>(S_in) * step_cost;
    } else {
      // LPT key of a forward unit: the VALU instructions its wave issues.
      // Per state the row update (2 mcp), the product chain (4 (mcp - 1)),
      // the column powers (grouped: 4 (cmax - 1) multiplies, plus one scalar
      // bit test per column at half weight; counts: 4 (n - m_c)) and the
      // accumulation (8); per unit the start state (2 mcp per row of B, the
      // top-digit decode) and the wave reduction.  The choice of the split
      // above keeps the coarser model; this one orders the units, where an
      // overestimate of the grouped items' products (about 4 (n - m_c) for
      // 4 (cmax - 1)) would run long units last and lengthen the launch's
      // tail.
      const double powers = cmax == 1 ? 0.0
                            : grouped ? 4.0 * (cmax - 1) + 0.5 * mcp
                                      : 4.0 * (n - mc);
      const double state = 2.0 * mcp + 4.0 * (mcp - 1) + powers + 8.0;
      inf.unit_cost = (1.0 + D) * 2.0 * mcp + Dt * (top_binary ? 4.0 : 60.0) + 300.0 +
                      static_cast<double>(S_in) * state;
    }
    plan.total_flops += inf.flops;
    plan.total_terms += T;
    item_units[k] = d.units;
    // column-minor items: unit-multiplicity kernels in buckets 0-4 (widths
    // 1-8, 9-16, 17-18, 19-20, 21-32), the general ones in 5-8 (1-8, 9-16,
    // 17-24, 25-32)
    if (opt.minors)
      item_bucket[k] = cmax > 1 ? bucket + 5
                                : (mc <= 8 ? 0 : mc <= 16 ? 1 : mc <= 18 ? 2 : mc <= 20 ? 3 : 4);
    else if (opt.grad)
      item_bucket[k] = mc <= 8 ? 0 : mc <= 16 ? 1 : mc <= 19 ? 2 : mc == 20 ? 3 : 4;
    else
      item_bucket[k] = bucket;
  }
  plan.n_partials = partial_cursor;
  plan.all_direct = !plan.items.empty();
  for (const ItemDesc &d : plan.items) {
    plan.max_units = std::max(plan.max_units, d.units);
    if (!(d.flags & kDirectOut)) plan.all_direct = false;
  }
  plan.n_out = out_cursor;
  plan.minors = opt.minors;
  plan.grad = opt.grad;

  // --- work units: grouped by kernel bucket, LPT (costliest units first) ---
  std::vector<int64_t> idx(n_items);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
    if (item_bucket[x] != item_bucket[y]) return item_bucket[x] < item_bucket[y];
    return plan.info[x].unit_cost > plan.info[y].unit_cost;
  });
  int64_t total_units = 0;
  for (int64_t k = 0; k < n_items; ++k) total_units += item_units[k];
  if (total_units > INT32_MAX) {
    err = "too many work units";
    return kErrRange;
  }
  plan.units.reserve(static_cast<size_t>(total_units));
  for (int64_t k : idx) {
    if (item_units[k] == 0) continue;
    const int b = item_bucket[k];
    if (plan.ranges.empty() || plan.ranges.back().bucket != b)
      plan.ranges.push_back(Plan::Range{b, static_cast<int32_t>(plan.units.size()), 0, 0.0});
    plan.ranges.back().cost += plan.info[k].unit_cost * static_cast<double>(item_units[k]);
    for (int64_t l = 0; l < item_units[k]; ++l) {
      plan.units.push_back(WorkUnit{static_cast<int32_t>(k), static_cast<int32_t>(l)});
      plan.bucket_of_unit.push_back(b);
    }
    plan.ranges.back().count += static_cast<int32_t>(item_units[k]);
  }
  // XCD-aware order within each range: one-block-per-unit launches hand
  // block b to XCD b mod kXcds, and each XCD has its own L2.  Cut every
  // item's units into chunks, deal the chunks in LPT order to the XCD list
  // with the fewest units, and interleave the lists, so that the units of
  // one chunk run on one XCD and its tables are fetched into one L2 instead
  // of eight (approximately LPT order overall; only the tail, where lists
  // run out unevenly, leaves the mapping).
  {
    std::vector<WorkUnit> lists[kXcds];
    std::vector<WorkUnit> tmp;
    for (Plan::Range &r : plan.ranges) {
      const int64_t chunk = std::max<int64_t>(1, r.count / (kXcds * 16));
      for (auto &l : lists) l.clear();
      size_t fill[kXcds] = {};
      for (int32_t u = r.first; u < r.first + r.count;) {
        // a run of one item's units, at most `chunk` long
        int32_t v = u + 1;
        while (v < r.first + r.count && v - u < chunk && plan.units[v].item == plan.units[u].item) ++v;
        int x = 0;
        for (int y = 1; y < kXcds; ++y)
          if (fill[y] < fill[x]) x = y;
        lists[x].insert(lists[x].end(), plan.units.begin() + u, plan.units.begin() + v);
        fill[x] += static_cast<size_t>(v - u);
        u = v;
      }
      tmp.clear();
      for (size_t round = 0; tmp.size() < static_cast<size_t>(r.count); ++round)
        for (int x = 0; x < kXcds; ++x)
          if (round < lists[x].size()) tmp.push_back(lists[x][round]);
      std::copy(tmp.begin(), tmp.end(), plan.units.begin() + r.first);
    }
  }
  plan.units_off = arena_alloc(plan.arena,
                               int64_t(std::max<size_t>(plan.units.size(), 1)) * sizeof(WorkUnit));
  if (!plan.units.empty())
    std::memcpy(plan.arena.data() + plan.units_off, plan.units.data(),
                plan.units.size() * sizeof(WorkUnit));
  std::memcpy(plan.arena.data() + plan.items_off, plan.items.data(),
              plan.items.size() * sizeof(ItemDesc));
  return kOk;
}

void plan_refresh_matrix(Plan &plan, const std::vector<PermInput> &inputs) {
  const size_t rsz = plan.fp32 ? sizeof(float) : sizeof(double);
  for (size_t k = 0; k < plan.items.size() && k < inputs.size(); ++k) {
    const ItemInfo &inf = plan.info[k];
    if (inf.trivial_one) continue;
    const ItemDesc &d = plan.items[k];
    const PermInput &in = inputs[k];
    const int mc = static_cast<int>(inf.colsrc.size());
    char *base = plan.arena.data() + d.b_off;
    for (size_t r = 0; r < inf.rowsrc.size(); ++r)
      for (int j = 0; j < mc; ++j) {
        const int64_t rs = inf.rowsrc[r], cs = inf.colsrc[static_cast<size_t>(j)];
        const int64_t e = (inf.transposed ? cs * in.ld + rs : rs * in.ld + cs) * 2;
        double re, im;
        if (in.a_is_float) {
          re = static_cast<const float *>(in.a)[e];
          im = static_cast<const float *>(in.a)[e + 1];
        } else {
          re = static_cast<const double *>(in.a)[e];
          im = static_cast<const double *>(in.a)[e + 1];
        }
        const int64_t o = (static_cast<int64_t>(r) * d.mcp + j) * 2;
        if (rsz == sizeof(float)) {
          reinterpret_cast<float *>(base)[o] = static_cast<float>(re);
          reinterpret_cast<float *>(base)[o + 1] = static_cast<float>(im);
        } else {
          reinterpret_cast<double *>(base)[o] = re;
          reinterpret_cast<double *>(base)[o + 1] = im;
        }
      }
  }
}

}  // namespace pqp
