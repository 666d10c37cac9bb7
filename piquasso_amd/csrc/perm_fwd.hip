// Forward permanent kernels for gfx950 (CDNA4).
//
// Glynn / BBFG with multiplicities over a reflected mixed-radix Gray code,
// the method of src/permanent.cpp:49-264 and src/n_aryGrayCodeCounter.hpp,
// re-organised for 64-wide wavefronts:
//
//  * The Gray digits of a permanent are split into inner digits and top
//    digits.  A wave takes 64 consecutive top states (one per lane); every
//    lane then sweeps all inner states in the same order.  The row that
//    changes at each inner step, its sign and the weight are therefore
//    wave-uniform: they come from scalar loads, and the row update is one
//    v_add_f64 per real component with a scalar operand.
//  * Each lane keeps its m_c complex row sums in VGPRs (compile-time width
//    MCP, fully unrolled), recomputes the column product with FMA complex
//    multiplies (4 VALU ops each) and accumulates coef * product in FP64.
//    Every kDdFold terms the FP64 run sum is folded into a double-double
//    accumulator (COMP = true); the plain baseline keeps one FP64 sum.
//  * Lane results are scaled by the lane's exact integer top weight and
//    sign, reduced across the wave with a butterfly of dd additions (the
//    same value lands in every lane, so the result does not depend on
//    scheduling), and written to a per-unit partial slot.  A finalize
//    kernel sums an item's partials in a fixed order and scales by 2^-(n-1).
//  * The persistent kernel pulls work units from a global queue (work
//    stealing across items of uneven size, LPT order from the planner); the
//    static kernel maps one block to one unit.
//
// No MFMA: the inner loop is a product of sums, not a contraction.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "dd.hpp"
#include "perm_kernels.hpp"
#include "pqp_types.hpp"
#include "sweep_common.hpp"

namespace pqp {

namespace {

// prod_j cs_j^{c_j}, columns sorted by multiplicity (descending).
//
// Grouped form (kPowGrouped): with P_j the running prefix product over
// columns 0..j and d_j = c_j - c_{j+1} (c_mc = 1),
//
//   prod_j cs_j^{c_j} = prod_j P_j^{d_j} * P_{mc-1},
//
// i.e. one pass over the columns (m_c - 1 multiplies) plus cmax - 1
// multiplies into a second accumulator at the columns where the
// multiplicity drops, instead of sum_j c_j - 1 = n - 1 multiplies.  The
// drop positions come as three wave-uniform bit masks (bit j of m[k]:
// d_j > k; d_j <= 3), loaded once per unit; the test at each column is a
// scalar bit test and a branch that is rarely taken (the empty volatile asm
// keeps the compiler from if-converting the rare work into masked
// multiplies at every column).  Items with a larger drop, or wider than
// kPowGroupedMaxCols, use the count form: all columns once, then column j
// again for each r = 2..c_j (columns with c_j >= r are the prefix
// [0, cnt_r)).
struct PowState {
  uint32_t m1, m2, m3;  // grouped form: snapshot masks
  int first;            // lowest set bit of m1
};

template <int MCP, typename R>
__device__ __forceinline__ void column_product(const R (&cr)[MCP], const R (&ci)[MCP],
                                               R &pr, R &pi, int flags, int cmax,
                                               const PowState &ps, cptr<int32_t> cpow) {
  if (flags & kColsAllOnes) {
    // all multiplicities 1: two interleaved chains for ILP
    R ar = cr[0], ai = ci[0];
    if constexpr (MCP > 1) {
      R br = cr[1], bi = ci[1];
#pragma unroll
      for (int j = 2; j + 1 < MCP; j += 2) {
        cmul(ar, ai, cr[j], ci[j]);
        cmul(br, bi, cr[j + 1], ci[j + 1]);
      }
      if constexpr (MCP % 2 == 1 && MCP > 2) cmul(ar, ai, cr[MCP - 1], ci[MCP - 1]);
      cmul(ar, ai, br, bi);
    }
    pr = ar;
    pi = ai;
    return;
  }
  R ar = cr[0], ai = ci[0];
  if (MCP <= kPowGroupedMaxCols && (flags & kPowGrouped)) {
    // The masks pass through an opaque copy each term: otherwise the
    // compiler hoists one precomputed condition per column out of the loop
    // into SGPR pairs, and the scalar register file overflows.
    uint32_t m1 = __builtin_amdgcn_readfirstlane(ps.m1);
    uint32_t m2 = __builtin_amdgcn_readfirstlane(ps.m2);
    uint32_t m3 = __builtin_amdgcn_readfirstlane(ps.m3);
    int first = __builtin_amdgcn_readfirstlane(ps.first);
    __asm__ volatile("" : "+s"(m1), "+s"(m2), "+s"(m3), "+s"(first));
    R qr = R(1), qi = R(0);
#pragma unroll
    for (int j = 0; j < MCP; ++j) {
      if (j > 0) cmul(ar, ai, cr[j], ci[j]);
      if (__builtin_expect((m1 >> j) & 1u, 0)) {
        __asm__ volatile("");
        if (j == first) {
          __asm__ volatile("");
          qr = ar;
          qi = ai;
        } else {
          __asm__ volatile("");
          cmul(qr, qi, ar, ai);
        }
        if (__builtin_expect((m2 >> j) & 1u, 0)) {
          __asm__ volatile("");
          cmul(qr, qi, ar, ai);
          if ((m3 >> j) & 1u) {
            __asm__ volatile("");
            cmul(qr, qi, ar, ai);
          }
        }
      }
    }
    cmul(ar, ai, qr, qi);
  } else {
#pragma unroll
    for (int j = 1; j < MCP; ++j) cmul(ar, ai, cr[j], ci[j]);
    for (int r = 2; r <= cmax; ++r) {
      const int cnt = cpow[r - 2];
#pragma unroll
      for (int j = 0; j < MCP; ++j) {
        if (j < cnt) {
          __asm__ volatile("");
          cmul(ar, ai, cr[j], ci[j]);
        }
      }
    }
  }
  pr = ar;
  pi = ai;
}

// A unit's sum over its 64 lanes (every lane holds it after the butterfly)
// and the magnitudes of its terms.
struct WaveTotal {
  cdd s;
  double abs;
};

// One wave unit: top states [64*local, 64*local+64) of item `item`.  ABS
// (compensated static launches): also sum the magnitudes of the terms.
// `total`, when given, receives the unit's sum as well.
template <int MCP, typename R, bool COMP, bool ABS, bool UNIT_BASE = false>
__device__ __forceinline__ void sweep_unit(const LaunchArgs &a, int item, int local,
                                           WaveTotal *total = nullptr) {
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const int64_t U = it->U;
  const int S = uniform(static_cast<int>(it->S_in));  // <= 2^16 inner states
  const int L = it->L;
  const int Dt = it->Dt;
  const int flags = uniform(it->flags);
  const int cmax = it->cmax;

  int64_t u = static_cast<int64_t>(local) * kWave + lane;
  const bool active = u < U;
  if (!active) u = U - 1;

  const cptr<R> Bin = b_rows<R>(a, it) + 2 * MCP;  // inner rows of B (after the prepended row)
  const cptr<InnerEntry> tab = as_const<InnerEntry>(a.arena, it->step_off);
  const cptr<int32_t> cpow = as_const<int32_t>(a.arena, it->cpow_off + 24);
  PowState ps{0u, 0u, 0u, 0};
  if (flags & kPowGrouped) {
    const cptr<uint64_t> pm = as_const<uint64_t>(a.arena, it->cpow_off);
    ps.m1 = static_cast<uint32_t>(pm[0]);
    ps.m2 = static_cast<uint32_t>(pm[1]);
    ps.m3 = static_cast<uint32_t>(pm[2]);
    ps.first = __builtin_ctz(ps.m1 | 0x80000000u);
  }

  R cr[MCP], ci[MCP];
  double w_top;
  start_state<MCP, R, UNIT_BASE>(a, it, u, L, Dt, flags, cr, ci, w_top);
  // (-1)^{|g_top|} = (-1)^u: consecutive top states differ by one unit step.
  double lane_w = (u & 1) ? -w_top : w_top;
  if (!active) lane_w = 0.0;

  // Inner sweep in blocks of kDdFold states: FP64 run sums inside a block,
  // folded into the double-double accumulators between blocks.  The inner
  // row of B added (times +-2) after state v (its first PF columns; the
  // scalar registers hold at most 16 complex values) is fetched before state
  // v's product, and the table entry of state v + 1 one state ahead, so
  // scalar-load latency hides behind the product's VALU work.  The update
  // after the last state uses the padding entry (0 x row 0) and is never
  // read.
  constexpr int PF = MCP < 16 ? MCP : 16;
  dd accr{0.0, 0.0}, acci{0.0, 0.0};
  // magnitudes of the terms (ABS only): sum_v |coef| (|Re p| + |Im p|),
  // two FMAs with |.| source modifiers into one accumulator (two VGPRs)
  double abs_acc = 0.0;
  InnerEntry e = inner_entry(tab, 0);
  for (int v0 = 0; v0 < S; v0 += kDdFold) {
    const int vend = v0 + kDdFold < S ? v0 + kDdFold : S;
    double sr = 0.0, si = 0.0;
    for (int v = v0; v < vend; ++v) {
      const cptr<R> row = Bin + static_cast<int64_t>(e.row) * (2 * MCP);
      R ur[PF], ui[PF];
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        ur[j] = row[2 * j];
        ui[j] = row[2 * j + 1];
      }
      const InnerEntry en = inner_entry(tab, v + 1);
      R pr, pi;
      column_product<MCP, R>(cr, ci, pr, pi, flags, cmax, ps, cpow);
      sr = __builtin_fma(e.coef, static_cast<double>(pr), sr);
      si = __builtin_fma(e.coef, static_cast<double>(pi), si);
      if constexpr (ABS) {
        abs_acc = __builtin_fma(__builtin_fabs(e.coef), __builtin_fabs(static_cast<double>(pr)), abs_acc);
        abs_acc = __builtin_fma(__builtin_fabs(e.coef), __builtin_fabs(static_cast<double>(pi)), abs_acc);
      }
      // row sums += (+-2) x inner row: exact doubling, one rounding (FMA)
      const R s2 = in_vgpr(static_cast<R>(e.sign2));
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        cr[j] = fmaR(s2, ur[j], cr[j]);
        ci[j] = fmaR(s2, ui[j], ci[j]);
      }
#pragma unroll
      for (int j = PF; j < MCP; ++j) {
        cr[j] = fmaR(s2, static_cast<R>(row[2 * j]), cr[j]);
        ci[j] = fmaR(s2, static_cast<R>(row[2 * j + 1]), ci[j]);
      }
      e = en;
    }
    if constexpr (COMP) {
      accr = dd_add_d(accr, sr);
      acci = dd_add_d(acci, si);
    } else {
      accr.hi += sr;
      acci.hi += si;
    }
  }
  cdd lane_sum;
  if constexpr (COMP) {
    lane_sum = cdd{dd_mul_d(accr, lane_w), dd_mul_d(acci, lane_w)};
  } else {
    lane_sum = cdd{{accr.hi * lane_w, 0.0}, {acci.hi * lane_w, 0.0}};
  }
  double abs_sum = abs_acc * __builtin_fabs(lane_w);
  if constexpr (COMP) {
    lane_sum = wave_reduce<0>(lane_sum, lane);
    if constexpr (ABS) abs_sum = wave_sum(abs_sum, lane);
  } else {
    // plain FP64 butterfly
    lane_sum.re.hi = wave_sum(lane_sum.re.hi, lane);
    lane_sum.im.hi = wave_sum(lane_sum.im.hi, lane);
  }
  if (total) {
    total->s = lane_sum;
    total->abs = abs_sum;
  }
  if (lane == 0) {
    if (flags & kDirectOut) {
      // the item's only unit: scale by 2^-(n-1) and write the result
      const int e = -(it->n - 1);
      const double re_hi = ldexp(lane_sum.re.hi, e), re_lo = ldexp(lane_sum.re.lo, e);
      const double im_hi = ldexp(lane_sum.im.hi, e), im_lo = ldexp(lane_sum.im.lo, e);
      const int o = it->out_index;
      a.out[2 * o] = re_hi + re_lo;
      a.out[2 * o + 1] = im_hi + im_lo;
      if (a.out_dd) {
        a.out_dd[4 * o] = re_hi;
        a.out_dd[4 * o + 1] = re_lo;
        a.out_dd[4 * o + 2] = im_hi;
        a.out_dd[4 * o + 3] = im_lo;
      }
      if (ABS && a.out_abs) a.out_abs[o] = ldexp(abs_sum, e);
    } else {
      double *p = a.partials + 4 * (it->partial_base + local);
      p[0] = lane_sum.re.hi;
      p[1] = lane_sum.re.lo;
      p[2] = lane_sum.im.hi;
      p[3] = lane_sum.im.lo;
      if (ABS && a.part_abs) a.part_abs[it->partial_base + local] = abs_sum;
    }
  }
}

template <int LO, int HI, typename R, bool COMP, bool ABS, bool UNIT_BASE = false>
__device__ __forceinline__ void dispatch(const LaunchArgs &a, int mcp, int item, int local,
                                         WaveTotal *total = nullptr) {
  if constexpr (LO <= HI) {
    if (mcp == LO) {
      sweep_unit<LO, R, COMP, ABS, UNIT_BASE>(a, item, local, total);
    } else {
      dispatch<LO + 1, HI, R, COMP, ABS, UNIT_BASE>(a, mcp, item, local, total);
    }
  }
}

}  // namespace

// Persistent, work-stealing launch: every wave loops over the queue.  It
// keeps no term magnitudes (the widest variant has no registers to spare).
template <int LO, int HI, typename R, bool COMP>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(waves_for(HI, sizeof(R) == 4), 8))) void perm_persistent_kernel(LaunchArgs a, int first_unit,
                                                                int n_units, int head_slot) {
  // The first unit of each wave is its block index; later units come from
  // the queue, whose next index is requested before the current unit runs
  // so the atomic's round trip overlaps the sweep.
  int k = static_cast<int>(blockIdx.x);
  for (;;) {
    if (k >= n_units) break;
    int next = 0;
    if (threadIdx.x == 0) next = atomicAdd(a.queue_head + head_slot, 1);
    const cptr<WorkUnit> wu = as_const<WorkUnit>(a.units) + (first_unit + k);
    const int item = uniform(wu->item);
    const int local = uniform(wu->local);
    const int mcp = uniform(as_const<ItemDesc>(a.items)[item].mcp);
    dispatch<LO, HI, R, COMP, false>(a, mcp, item, local);
    k = __builtin_amdgcn_readfirstlane(next) + static_cast<int>(gridDim.x);
  }
}


// The block's sum of one item's partials: thread t adds partials t, t + NT,
// ... (double-double), each wave reduces with a butterfly, and the
// per-wave sums are added in wave order, so the result does not depend on
// scheduling; then the 2^-(n-1) scale, the complex128 result, its dd parts
// and the term magnitudes (thread 0 writes them).
template <bool COMP, int NT>
__device__ __forceinline__ void finalize_item(const LaunchArgs &a, int item) {
  constexpr int NW = NT / kWave;
  __shared__ double part[NW][5];
  const int tid = static_cast<int>(threadIdx.x);
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int units = it->units;
  const int64_t base = it->partial_base;
  cdd s = cdd_zero();
  double abs_sum = 0.0;  // magnitudes of the terms (COMP with part_abs)
  const bool with_abs = COMP && a.part_abs && a.out_abs;
  if (it->n == 0) {
    // empty selection: permanent 1 by definition
    if (tid == 0) s.re.hi = abs_sum = 1.0;
  } else {
    for (int k = tid; k < units; k += NT) {
      const double *p = a.partials + 4 * (base + k);
      if constexpr (COMP) {
        s = cdd_add(s, cdd{{p[0], p[1]}, {p[2], p[3]}});
        if (with_abs) abs_sum += a.part_abs[base + k];
      } else {
        s.re.hi += p[0];
        s.im.hi += p[2];
      }
    }
  }
  if constexpr (COMP) {
    s = wave_reduce<0>(s, lane);
    abs_sum = wave_sum(abs_sum, lane);
  } else {
    s.re.hi = wave_sum(s.re.hi, lane);
    s.im.hi = wave_sum(s.im.hi, lane);
  }
  if constexpr (NW > 1) {
    if (lane == 0) {
      part[wave][0] = s.re.hi;
      part[wave][1] = s.re.lo;
      part[wave][2] = s.im.hi;
      part[wave][3] = s.im.lo;
      part[wave][4] = abs_sum;
    }
    __syncthreads();
    if (tid == 0) {
      s = cdd{{part[0][0], part[0][1]}, {part[0][2], part[0][3]}};
      abs_sum = part[0][4];
      for (int w = 1; w < NW; ++w) {
        const cdd o{{part[w][0], part[w][1]}, {part[w][2], part[w][3]}};
        abs_sum += part[w][4];
        if constexpr (COMP) {
          s = cdd_add(s, o);
        } else {
          s.re.hi += o.re.hi;
          s.im.hi += o.im.hi;
        }
      }
    }
  }
  if (tid == 0) {
    const int e = it->n > 0 ? -(it->n - 1) : 0;
    const double re_hi = ldexp(s.re.hi, e), re_lo = ldexp(s.re.lo, e);
    const double im_hi = ldexp(s.im.hi, e), im_lo = ldexp(s.im.lo, e);
    const int o = it->out_index;
    a.out[2 * o] = re_hi + re_lo;
    a.out[2 * o + 1] = im_hi + im_lo;
    if (a.out_dd) {
      a.out_dd[4 * o] = re_hi;
      a.out_dd[4 * o + 1] = re_lo;
      a.out_dd[4 * o + 2] = im_hi;
      a.out_dd[4 * o + 3] = im_lo;
    }
    if (with_abs) a.out_abs[o] = ldexp(abs_sum, e);
  }
}

// One block of NT threads per item (finalize_item).  NT = 64 for batches of
// small items, 1024 when an item has thousands of units.  Block 0 also
// rewinds the work-queue heads for the next launch of the same plan.
// Static launch: block b sweeps unit first_unit + b.  A one-item launch of a
// single permanent (a.last_counter set) also finalizes the item in the block
// that finishes last: every block makes its partial visible device-wide,
// counts itself, and the block that counts last reads all partials (the
// fences order the partial writes, on any XCD, before the count and the
// count before the reads), writes the result to mapped host memory, resets
// the counter for the next launch on the stream and publishes the sequence
// number, so the caller polls instead of waiting for a finalize launch.
template <int LO, int HI, typename R, bool COMP>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(waves_for(HI, sizeof(R) == 4), 8))) void perm_static_kernel(LaunchArgs a, int first_unit) {
  const cptr<WorkUnit> wu = as_const<WorkUnit>(a.units) + (first_unit + static_cast<int>(blockIdx.x));
  const int item = uniform(wu->item);
  const int local = uniform(wu->local);
  const int mcp = uniform(as_const<ItemDesc>(a.items)[item].mcp);
  dispatch<LO, HI, R, COMP, COMP>(a, mcp, item, local);
  // widths up to 16 only (kLastFinalizeMaxWidth): in the wider kernels the
  // finalize code took private memory
  if constexpr (COMP && HI <= kLastFinalizeMaxWidth) {
    if (a.last_counter) {
      __threadfence();
      int old = 0;
      if (threadIdx.x == 0) old = atomicAdd(a.last_counter, 1);
      old = __builtin_amdgcn_readfirstlane(old);
      if (old == static_cast<int>(gridDim.x) - 1) {
        __threadfence();
        finalize_item<true, kWave>(a, item);
        if (threadIdx.x == 0) {
          *a.last_counter = 0;
          __threadfence_system();
          __hip_atomic_store(a.done, a.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
      }
    }
  }
}

template <bool COMP, int NT>
__global__ __launch_bounds__(NT) void perm_finalize_kernel(LaunchArgs a, int item0, int n_heads) {
  const int item = item0 + static_cast<int>(blockIdx.x);
  if (blockIdx.x == 0 && static_cast<int>(threadIdx.x) < n_heads) a.queue_head[threadIdx.x] = 0;
  if (item >= a.n_items) return;
  finalize_item<COMP, NT>(a, item);
}

// One small permanent (item 0 of a one-item plan) in one launch, for the
// single-call path: the block's waves sweep the item's units (wave w: units
// w, w + W, ...), the block adds their sums (a one-unit item has written
// its result from the sweep), and thread 0 then publishes `seq` to `done`
// (host-visible memory) after a system-scope fence, so the host can poll
// for the result instead of synchronising the stream.
//
// A lone small sweep is latency-bound: every first touch of a table line
// is a scalar-cache miss, and the sweep's loads depend on each other
// (descriptor -> table -> row), so misses would queue up one after another.
// The waves therefore first touch every 64-byte line of the plan's arena
// (wave w: lines w, w + W, ..., four independent loads in flight per wave)
// and meet at a barrier, after which the sweep's scalar loads hit.
//
// The rows of B may come in the kernel arguments (SingleArgs::b): the
// argument struct is read through the kernel-argument segment pointer, as
// scalar loads like the arena.
template <int LO, int HI, typename R, bool UNIT_BASE>
__global__ __launch_bounds__(kSingleBlockThreads) void perm_single_block_kernel(SingleArgs args) {
  constexpr int W = kSingleBlockThreads / kWave;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const uintptr_t b_addr = reinterpret_cast<uintptr_t>(__builtin_amdgcn_kernarg_segment_ptr()) +
                           offsetof(SingleArgs, b);
  LaunchArgs a = args.a;
  const int arena_lines = args.arena_lines;
  int b_lines = 0;
  if (args.b_in_args) {
    a.b_rows = reinterpret_cast<const char *>(b_addr);
    b_lines = kSingleArgBBytes / 64;
  }
  if (UNIT_BASE) a.unit_base = reinterpret_cast<const char *>(b_addr + args.base_off);
  {
    const cptr<uint32_t> lines = as_const<uint32_t>(a.arena);
    const cptr<uint32_t> blines = reinterpret_cast<cptr<uint32_t>>(b_addr);
    // straight-line loads with indices clamped to the last line, consumed
    // together by one asm statement, so all of them are in flight before
    // a single wait (the first 4 W lines of the arena and the kernel
    // arguments' rows of B; larger arenas continue in a loop)
    const int la = arena_lines - 1, lb = b_lines > 0 ? b_lines - 1 : 0;
    const cptr<uint32_t> bl = b_lines > 0 ? blines : lines;
    const int lbx = b_lines > 0 ? lb : la;
    const uint32_t v0 = lines[16 * min(wave, la)], v1 = lines[16 * min(wave + W, la)];
    const uint32_t v2 = lines[16 * min(wave + 2 * W, la)], v3 = lines[16 * min(wave + 3 * W, la)];
    const uint32_t v4 = bl[16 * min(wave, lbx)], v5 = bl[16 * min(wave + W, lbx)];
    const uint32_t v6 = bl[16 * min(wave + 2 * W, lbx)], v7 = bl[16 * min(wave + 3 * W, lbx)];
    __asm__ volatile("" ::"s"(v0), "s"(v1), "s"(v2), "s"(v3), "s"(v4), "s"(v5), "s"(v6), "s"(v7));
    uint32_t t = 0;
    for (int l = wave + 4 * W; l < arena_lines; l += W) t ^= lines[16 * l];
    __asm__ volatile("" ::"s"(t));
  }
  __syncthreads();
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items);
  const int units = uniform(it->units);
  const int mcp = uniform(it->mcp);
  if (units <= W) {
    // one unit per wave: the waves' totals meet in LDS and the first 16
    // lanes of wave 0 add them with a fixed butterfly (a one-unit item has
    // written its result itself)
    __shared__ double part[W][5];
    WaveTotal total{cdd_zero(), 0.0};
    if (wave < units) dispatch<LO, HI, R, true, true, UNIT_BASE>(a, mcp, 0, wave, &total);
    if ((threadIdx.x & (kWave - 1)) == 0 && wave < units) {
      part[wave][0] = total.s.re.hi;
      part[wave][1] = total.s.re.lo;
      part[wave][2] = total.s.im.hi;
      part[wave][3] = total.s.im.lo;
      part[wave][4] = total.abs;
    }
    __syncthreads();
    if (wave == 0 && !(uniform(it->flags) & kDirectOut)) {
      // lane l < units holds unit l's total, the others zero; butterfly
      // steps over lanes l ^ 1, l ^ 2, 7 - l and 15 - l leave the sum of
      // lanes 0..15 (W = 16 waves) in every one of them, added in a fixed
      // order; lane 0 writes the result
      static_assert(W == 16, "the butterfly below covers 16 unit totals");
      const int lane = static_cast<int>(threadIdx.x);
      const bool has = lane < units;
      cdd s = has ? cdd{{part[lane][0], part[lane][1]}, {part[lane][2], part[lane][3]}} : cdd_zero();
      double abs_sum = has ? part[lane][4] : 0.0;
      s = wave_reduce<0, 4>(s, lane);
      abs_sum = wave_sum<0, 4>(abs_sum, lane);
      if (lane == 0) {
        const int e = -(it->n - 1);
        const double re_hi = ldexp(s.re.hi, e), re_lo = ldexp(s.re.lo, e);
        const double im_hi = ldexp(s.im.hi, e), im_lo = ldexp(s.im.lo, e);
        const int o = it->out_index;
        a.out[2 * o] = re_hi + re_lo;
        a.out[2 * o + 1] = im_hi + im_lo;
        if (a.out_dd) {
          a.out_dd[4 * o] = re_hi;
          a.out_dd[4 * o + 1] = re_lo;
          a.out_dd[4 * o + 2] = im_hi;
          a.out_dd[4 * o + 3] = im_lo;
        }
        if (a.out_abs) a.out_abs[o] = ldexp(abs_sum, e);
      }
    }
  } else {
    for (int u = wave; u < units; u += W) dispatch<LO, HI, R, true, true>(a, mcp, 0, u);
    __syncthreads();
    if (!(uniform(it->flags) & kDirectOut)) finalize_item<true, kSingleBlockThreads>(a, 0);
  }
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(args.done, args.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

}  // namespace pqp

// ---------------------------------------------------------------------------
// Instantiation for one column-width bucket and one arithmetic variant.  This
// file is compiled once per (bucket, variant) pair with -DPQP_BUCKET_ID=k and
// -DPQP_VARIANT=v; the bucket table matches plan.cpp:
//   buckets 0..3: widths 1..8, 9..16, 17..24, 25..32 (narrow ranges keep the
//   register allocation, hence the waves per SIMD, close to each width's
//   need); 4..7: 40, 48, 56, 64 (padded widths);
//   variant 0: FP64 + double-double, 1: plain FP64, 2: FP32 products + dd.
// ---------------------------------------------------------------------------
#ifndef PQP_BUCKET_ID
#error "compile with -DPQP_BUCKET_ID=<0..7>"
#endif
#ifndef PQP_VARIANT
#error "compile with -DPQP_VARIANT=<0..2>"
#endif

#if defined(PQP_LO) && defined(PQP_HI)
// explicit width range (kernel analysis builds)
#elif PQP_BUCKET_ID == 0
#define PQP_LO 1
#define PQP_HI 8
#elif PQP_BUCKET_ID == 1
#define PQP_LO 9
#define PQP_HI 16
#elif PQP_BUCKET_ID == 2
#define PQP_LO 17
#define PQP_HI 24
#elif PQP_BUCKET_ID == 3
#define PQP_LO 25
#define PQP_HI 32
#elif PQP_BUCKET_ID == 4
#define PQP_LO 40
#define PQP_HI 40
#elif PQP_BUCKET_ID == 5
#define PQP_LO 48
#define PQP_HI 48
#elif PQP_BUCKET_ID == 6
#define PQP_LO 56
#define PQP_HI 56
#elif PQP_BUCKET_ID == 7
#define PQP_LO 64
#define PQP_HI 64
#endif

#if PQP_VARIANT == 0
#define PQP_R double
#define PQP_COMP true
#define PQP_SFX d
#elif PQP_VARIANT == 1
#define PQP_R double
#define PQP_COMP false
#define PQP_SFX p
#else
#define PQP_R float
#define PQP_COMP true
#define PQP_SFX f
#endif

#define PQP_PASTE_(base, k, v) base##k##_##v
#define PQP_PASTE(base, k, v) PQP_PASTE_(base, k, v)
#define PQP_NAME(base) PQP_PASTE(base, PQP_BUCKET_ID, PQP_SFX)

namespace pqp {

hipError_t PQP_NAME(launch_fwd_b)(const FwdLaunch &l) {
  if (l.n_units <= 0) return hipSuccess;
  if (l.persistent) {
    hipLaunchKernelGGL((perm_persistent_kernel<PQP_LO, PQP_HI, PQP_R, PQP_COMP>), dim3(l.grid),
                       dim3(kWave), 0, l.stream, l.args, l.first_unit, l.n_units, l.head_slot);
  } else {
    hipLaunchKernelGGL((perm_static_kernel<PQP_LO, PQP_HI, PQP_R, PQP_COMP>), dim3(l.n_units),
                       dim3(kWave), 0, l.stream, l.args, l.first_unit);
  }
  return hipGetLastError();
}

hipError_t PQP_NAME(occupancy_fwd_b)(bool persistent, int *blocks) {
  if (persistent)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks, perm_persistent_kernel<PQP_LO, PQP_HI, PQP_R, PQP_COMP>, kWave, 0);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks, perm_static_kernel<PQP_LO, PQP_HI, PQP_R, PQP_COMP>, kWave, 0);
}

#if PQP_BUCKET_ID <= 1 && PQP_VARIANT != 1
hipError_t PQP_NAME(launch_single_block_b)(const SingleArgs &args, hipStream_t stream) {
  // the instantiation with per-unit start rows only where the host made them
  if (args.base_in_args)
    hipLaunchKernelGGL((perm_single_block_kernel<PQP_LO, PQP_HI, PQP_R, true>), dim3(1),
                       dim3(kSingleBlockThreads), 0, stream, args);
  else
    hipLaunchKernelGGL((perm_single_block_kernel<PQP_LO, PQP_HI, PQP_R, false>), dim3(1),
                       dim3(kSingleBlockThreads), 0, stream, args);
  return hipGetLastError();
}
#endif

#if PQP_BUCKET_ID == 0 && PQP_VARIANT == 0
hipError_t launch_finalize(bool comp, const LaunchArgs &a, int item0, int n_items,
                           int n_heads, int max_units, hipStream_t stream) {
  const int grid = n_items > 0 ? n_items : 1;
  const bool wide = max_units > 256;
  if (comp && wide)
    hipLaunchKernelGGL((perm_finalize_kernel<true, 1024>), dim3(grid), dim3(1024), 0, stream, a,
                       item0, n_heads);
  else if (comp)
    hipLaunchKernelGGL((perm_finalize_kernel<true, kWave>), dim3(grid), dim3(kWave), 0, stream, a,
                       item0, n_heads);
  else if (wide)
    hipLaunchKernelGGL((perm_finalize_kernel<false, 1024>), dim3(grid), dim3(1024), 0, stream, a,
                       item0, n_heads);
  else
    hipLaunchKernelGGL((perm_finalize_kernel<false, kWave>), dim3(grid), dim3(kWave), 0, stream,
                       a, item0, n_heads);
  return hipGetLastError();
}
#endif

}  // namespace pqp
