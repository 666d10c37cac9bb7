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

#include <cstdint>

#include "dd.hpp"
#include "perm_kernels.hpp"
#include "pqp_types.hpp"

namespace pqp {

namespace {

__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// (ar + i ai) *= (xr + i xi): 2 multiplies + 2 FMAs.
template <typename R>
__device__ __forceinline__ void cmul(R &ar, R &ai, R xr, R xi) {
  const R t1 = ai * xi;
  const R t2 = ar * xi;
  const R nr = fmaR(ar, xr, -t1);
  const R ni = fmaR(ai, xr, t2);
  ar = nr;
  ai = ni;
}

// Read-only tables are addressed through the constant address space so
// that wave-uniform accesses become scalar loads (s_load) regardless of the
// stores the kernel makes to its output buffers.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ cptr<T> as_const(const void *p, int64_t byte_off = 0) {
  return reinterpret_cast<cptr<T>>(reinterpret_cast<uintptr_t>(p) + static_cast<uintptr_t>(byte_off));
}

template <typename T>
__device__ __forceinline__ T uniform(T v) {
  return v;
}

// Copy a (wave-uniform) value into a VGPR.  A VALU instruction reads at
// most one scalar source; keeping coefficients in VGPRs lets the row data
// stay scalar operands instead of being copied to VGPRs column by column.
__device__ __forceinline__ double in_vgpr(double x) {
  double r;
  __asm__("v_mov_b64 %0, %1" : "=v"(r) : "s"(x));
  return r;
}
__device__ __forceinline__ float in_vgpr(float x) {
  float r;
  __asm__("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
  return r;
}

// 1/i for the exact binomial recurrence (i <= 64).
__constant__ double kInvSmall[65] = {
    0.0,        1.0 / 1,  1.0 / 2,  1.0 / 3,  1.0 / 4,  1.0 / 5,  1.0 / 6,  1.0 / 7,
    1.0 / 8,    1.0 / 9,  1.0 / 10, 1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15,
    1.0 / 16,   1.0 / 17, 1.0 / 18, 1.0 / 19, 1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23,
    1.0 / 24,   1.0 / 25, 1.0 / 26, 1.0 / 27, 1.0 / 28, 1.0 / 29, 1.0 / 30, 1.0 / 31,
    1.0 / 32,   1.0 / 33, 1.0 / 34, 1.0 / 35, 1.0 / 36, 1.0 / 37, 1.0 / 38, 1.0 / 39,
    1.0 / 40,   1.0 / 41, 1.0 / 42, 1.0 / 43, 1.0 / 44, 1.0 / 45, 1.0 / 46, 1.0 / 47,
    1.0 / 48,   1.0 / 49, 1.0 / 50, 1.0 / 51, 1.0 / 52, 1.0 / 53, 1.0 / 54, 1.0 / 55,
    1.0 / 56,   1.0 / 57, 1.0 / 58, 1.0 / 59, 1.0 / 60, 1.0 / 61, 1.0 / 62, 1.0 / 63,
    1.0 / 64};
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// prod_j cs_j^{c_j}: all columns once (two interleaved chains for ILP),
// then the extra powers; columns are sorted by multiplicity, so the
// columns with c_j >= r form the prefix [0, cpow[r-2]).
template <int MCP, typename R>
__device__ __forceinline__ void column_product(const R (&cr)[MCP], const R (&ci)[MCP],
                                               R &pr, R &pi, bool ones, int cmax,
                                               cptr<int32_t> cpow) {
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
  if (!ones) {
    for (int r = 2; r <= cmax; ++r) {
      const int cnt = cpow[r - 2];
#pragma unroll
      for (int j = 0; j < MCP; ++j) {
        if (j < cnt) {
          // keep this a (uniform, scalar) branch: if-converting it would
          // turn every column into a masked multiply
          __asm__ volatile("" ::: "memory");
          cmul(ar, ai, cr[j], ci[j]);
        }
      }
    }
  }
  pr = ar;
  pi = ai;
}

// Butterfly reduction of a complex dd value over the 64 lanes.
__device__ __forceinline__ cdd wave_reduce(cdd v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    cdd o;
    o.re.hi = __shfl_xor(v.re.hi, off);
    o.re.lo = __shfl_xor(v.re.lo, off);
    o.im.hi = __shfl_xor(v.im.hi, off);
    o.im.lo = __shfl_xor(v.im.lo, off);
    v = cdd_add(v, o);
  }
  return v;
}

// One wave unit: top states [64*local, 64*local+64) of item `item`.
template <int MCP, typename R, bool COMP>
__device__ void sweep_unit(const LaunchArgs &a, int item, int local) {
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const int64_t U = it->U;
  const int64_t S = it->S_in;
  const int L = it->L;
  const int Dt = it->Dt;
  const int flags = it->flags;
  const int cmax = it->cmax;
  const bool ones = (flags & kColsAllOnes) != 0;

  int64_t u = static_cast<int64_t>(local) * kWave + lane;
  const bool active = u < U;
  if (!active) u = U - 1;

  const cptr<R> B = as_const<R>(a.arena, it->b_off);
  const cptr<R> B2 = as_const<R>(a.arena, it->b2_off);
  const cptr<double> alpha = as_const<double>(a.arena, it->alpha_off);
  const cptr<double> coef = as_const<double>(a.arena, it->coef_off);
  const cptr<uint32_t> steps = as_const<uint32_t>(a.arena, it->step_off);
  const cptr<int32_t> cpow = as_const<int32_t>(a.arena, it->cpow_off);

  // Row sums at the lane's first state,
  //   b_0 + sum_inner rho_k b_k + sum_top (rho_t - 2 g_t(u)) b_t,
  // accumulated in one loop over the rows of B from the highest top digit
  // down to the prepended row (one set of live row sums, one FMA block).
  // Row data are scalar operands; the coefficient is always a VGPR so each
  // update is a single v_fma with one scalar source.
  //
  // Top digits, mixed radix, top-down as the reference counter initialises
  // (src/n_aryGrayCodeCounter.hpp:178-193): counter digit
  // c_t = floor(u / P_t) - R_t floor(u / P_{t+1}); Gray digit g_t = c_t,
  // reflected to R_t - 1 - c_t when the higher Gray digits have odd sum.
  // floor(u / P) is evaluated as floor((u + 1/2) * (1/P)) in FP64, exact for
  // u < 2^50 (the fraction stays >= 1/(2P) away from an integer).  For
  // all-binary top digits this is the reflected binary code u ^ (u >> 1).
  R cr[MCP], ci[MCP];
#pragma unroll
  for (int j = 0; j < MCP; ++j) {
    cr[j] = R(0);
    ci[j] = R(0);
  }
  const cptr<double> top = as_const<double>(a.arena, it->top_off);
  const uint64_t gbin = static_cast<uint64_t>(u) ^ (static_cast<uint64_t>(u) >> 1);
  const double ud = static_cast<double>(u) + 0.5;
  double q_hi = 0.0;
  int parity = 0;
  double w_top = 1.0;
  for (int r = L + Dt; r >= 0; --r) {
    R al;
    if (r > L) {
      const int t = r - L - 1;
      if (flags & kTopBinary) {
        al = ((gbin >> t) & 1) ? R(-1) : R(1);
      } else {
        const double inv_place = top[4 * t];
        const double radix = top[4 * t + 1];
        const int rho = static_cast<int>(top[4 * t + 2]);
        const double q = __builtin_floor(ud * inv_place);
        const int c = static_cast<int>(__builtin_fma(-radix, q_hi, q));
        q_hi = q;
        const int g = parity ? static_cast<int>(radix) - 1 - c : c;
        parity ^= (g & 1);
        // C(rho, g) = prod_{i=1..g} (rho - g + i) / i; every partial product
        // is an integer, recovered exactly by rounding.
        double cb = 1.0;
        for (int i = 1; i <= rho; ++i) {
          const double f = (i <= g) ? static_cast<double>(rho - g + i) : 1.0;
          const double inv_i = (i <= g) ? kInvSmall[i] : 1.0;
          cb = __builtin_rint(cb * f * inv_i);
        }
        w_top *= cb;
        al = static_cast<R>(rho - 2 * g);
      }
    } else {
      al = in_vgpr(r > 0 ? static_cast<R>(alpha[r - 1]) : R(1));
    }
    const cptr<R> row = B + static_cast<int64_t>(r) * 2 * MCP;
#pragma unroll
    for (int j = 0; j < MCP; ++j) {
      cr[j] = fmaR(al, row[2 * j], cr[j]);
      ci[j] = fmaR(al, row[2 * j + 1], ci[j]);
    }
  }
  // (-1)^{|g_top|} = (-1)^u: consecutive top states differ by one unit step.
  double lane_w = (u & 1) ? -w_top : w_top;
  if (!active) lane_w = 0.0;

  // Inner sweep.
  double sr = 0.0, si = 0.0;
  dd accr{0.0, 0.0}, acci{0.0, 0.0};
  int fold = 0;
  for (int64_t v = 0;; ++v) {
    R pr, pi;
    column_product<MCP, R>(cr, ci, pr, pi, ones, cmax, cpow);
    const double cf = coef[v];
    sr = __builtin_fma(cf, static_cast<double>(pr), sr);
    si = __builtin_fma(cf, static_cast<double>(pi), si);
    if constexpr (COMP) {
      if (++fold == kDdFold) {
        accr = dd_add_d(accr, sr);
        acci = dd_add_d(acci, si);
        sr = 0.0;
        si = 0.0;
        fold = 0;
      }
    }
    if (v + 1 >= S) break;
    const cptr<R> row = B2 + static_cast<int64_t>(steps[v]) * 2 * MCP;
#pragma unroll
    for (int j = 0; j < MCP; ++j) {
      cr[j] += row[2 * j];
      ci[j] += row[2 * j + 1];
    }
  }
  cdd lane_sum;
  if constexpr (COMP) {
    accr = dd_add_d(accr, sr);
    acci = dd_add_d(acci, si);
    lane_sum = cdd{dd_mul_d(accr, lane_w), dd_mul_d(acci, lane_w)};
  } else {
    lane_sum = cdd{{sr * lane_w, 0.0}, {si * lane_w, 0.0}};
  }
  if constexpr (COMP) {
    lane_sum = wave_reduce(lane_sum);
  } else {
    // plain FP64 butterfly
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      lane_sum.re.hi += __shfl_xor(lane_sum.re.hi, off);
      lane_sum.im.hi += __shfl_xor(lane_sum.im.hi, off);
    }
  }
  if (lane == 0) {
    double *p = a.partials + 4 * (it->partial_base + local);
    p[0] = lane_sum.re.hi;
    p[1] = lane_sum.re.lo;
    p[2] = lane_sum.im.hi;
    p[3] = lane_sum.im.lo;
  }
}

template <int LO, int HI, typename R, bool COMP>
__device__ __forceinline__ void dispatch(const LaunchArgs &a, int mcp, int item, int local) {
  if constexpr (LO <= HI) {
    if (mcp == LO) {
      sweep_unit<LO, R, COMP>(a, item, local);
    } else {
      dispatch<LO + 1, HI, R, COMP>(a, mcp, item, local);
    }
  }
}

}  // namespace

// Waves per SIMD the register allocator must leave room for: the row sums
// take 4 VGPRs per column in FP64 (2 in FP32); the inner loop is pure VALU
// with scalar operands, so 2-3 waves per SIMD hide the scalar-load latency.
constexpr int waves_for(int hi, bool fp32) {
  const int w = fp32 ? hi / 2 : hi;
  return w <= 8 ? 6 : w <= 16 ? 4 : w <= 32 ? 3 : w <= 48 ? 2 : 1;
}

// Persistent, work-stealing launch: every wave loops over the queue.
template <int LO, int HI, typename R, bool COMP>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(waves_for(HI, sizeof(R) == 4), 8))) void perm_persistent_kernel(LaunchArgs a, int first_unit,
                                                                int n_units, int head_slot) {
  for (;;) {
    int k = 0;
    if (threadIdx.x == 0) k = atomicAdd(a.queue_head + head_slot, 1);
    k = __builtin_amdgcn_readfirstlane(k);
    if (k >= n_units) break;
    const cptr<WorkUnit> wu = as_const<WorkUnit>(a.units) + (first_unit + k);
    const int item = uniform(wu->item);
    const int local = uniform(wu->local);
    const int mcp = uniform(as_const<ItemDesc>(a.items)[item].mcp);
    dispatch<LO, HI, R, COMP>(a, mcp, item, local);
  }
}

// Static launch: block b sweeps unit first_unit + b.
template <int LO, int HI, typename R, bool COMP>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(waves_for(HI, sizeof(R) == 4), 8))) void perm_static_kernel(LaunchArgs a, int first_unit) {
  const cptr<WorkUnit> wu = as_const<WorkUnit>(a.units) + (first_unit + static_cast<int>(blockIdx.x));
  const int item = uniform(wu->item);
  const int local = uniform(wu->local);
  const int mcp = uniform(as_const<ItemDesc>(a.items)[item].mcp);
  dispatch<LO, HI, R, COMP>(a, mcp, item, local);
}

// One wave per item: sum the item's partials in a fixed order, scale by
// 2^-(n-1), write the complex128 result and its dd parts.  Block 0 also
// rewinds the work-queue heads for the next launch of the same plan.
template <bool COMP>
__global__ __launch_bounds__(kWave) void perm_finalize_kernel(LaunchArgs a, int item0,
                                                              int n_heads) {
  const int item = item0 + static_cast<int>(blockIdx.x);
  const int lane = static_cast<int>(threadIdx.x);
  if (blockIdx.x == 0 && lane < n_heads) a.queue_head[lane] = 0;
  if (item >= a.n_items) return;
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int units = it->units;
  const int64_t base = it->partial_base;
  cdd s = cdd_zero();
  if (it->n == 0) {
    // empty selection: permanent 1 by definition
    if (lane == 0) s.re.hi = 1.0;
  } else {
    for (int k = lane; k < units; k += kWave) {
      const double *p = a.partials + 4 * (base + k);
      if constexpr (COMP) {
        s = cdd_add(s, cdd{{p[0], p[1]}, {p[2], p[3]}});
      } else {
        s.re.hi += p[0];
        s.im.hi += p[2];
      }
    }
  }
  if constexpr (COMP) {
    s = wave_reduce(s);
  } else {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      s.re.hi += __shfl_xor(s.re.hi, off);
      s.im.hi += __shfl_xor(s.im.hi, off);
    }
  }
  if (lane == 0) {
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
  }
}

}  // namespace pqp

// ---------------------------------------------------------------------------
// Instantiation for one column-width bucket (this file is compiled once per
// bucket with -DPQP_BUCKET_ID=k; the bucket table matches plan.cpp).
// ---------------------------------------------------------------------------
#ifndef PQP_BUCKET_ID
#error "compile with -DPQP_BUCKET_ID=<0..7>"
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

#define PQP_CAT2(a, b) a##b
#define PQP_CAT(a, b) PQP_CAT2(a, b)

namespace pqp {

namespace {

template <typename R, bool COMP>
hipError_t launch_variant(const FwdLaunch &l) {
  if (l.n_units <= 0) return hipSuccess;
  if (l.persistent) {
    hipLaunchKernelGGL((perm_persistent_kernel<PQP_LO, PQP_HI, R, COMP>), dim3(l.grid),
                       dim3(kWave), 0, l.stream, l.args, l.first_unit, l.n_units,
                       l.head_slot);
  } else {
    hipLaunchKernelGGL((perm_static_kernel<PQP_LO, PQP_HI, R, COMP>), dim3(l.n_units),
                       dim3(kWave), 0, l.stream, l.args, l.first_unit);
  }
  return hipGetLastError();
}

template <typename R, bool COMP>
hipError_t occupancy_variant(bool persistent, int *blocks) {
  if (persistent)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks, perm_persistent_kernel<PQP_LO, PQP_HI, R, COMP>, kWave, 0);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks, perm_static_kernel<PQP_LO, PQP_HI, R, COMP>, kWave, 0);
}

}  // namespace

hipError_t PQP_CAT(launch_fwd_b, PQP_BUCKET_ID)(const FwdLaunch &l) {
  if (l.fp32) {
    if (!l.comp) return hipErrorInvalidValue;  // FP32 products always use dd accumulation
    return launch_variant<float, true>(l);
  }
  return l.comp ? launch_variant<double, true>(l) : launch_variant<double, false>(l);
}

hipError_t PQP_CAT(occupancy_fwd_b, PQP_BUCKET_ID)(bool persistent, bool fp32, bool comp,
                                                   int *blocks) {
  if (fp32) return occupancy_variant<float, true>(persistent, blocks);
  return comp ? occupancy_variant<double, true>(persistent, blocks)
              : occupancy_variant<double, false>(persistent, blocks);
}

#if PQP_BUCKET_ID == 0
hipError_t launch_fwd(int bucket, const FwdLaunch &l) {
  switch (bucket) {
    case 0: return launch_fwd_b0(l);
    case 1: return launch_fwd_b1(l);
    case 2: return launch_fwd_b2(l);
    case 3: return launch_fwd_b3(l);
    case 4: return launch_fwd_b4(l);
    case 5: return launch_fwd_b5(l);
    case 6: return launch_fwd_b6(l);
    case 7: return launch_fwd_b7(l);
    default: return hipErrorInvalidValue;
  }
}

hipError_t occupancy_fwd(int bucket, bool persistent, bool fp32, bool comp, int *blocks) {
  switch (bucket) {
    case 0: return occupancy_fwd_b0(persistent, fp32, comp, blocks);
    case 1: return occupancy_fwd_b1(persistent, fp32, comp, blocks);
    case 2: return occupancy_fwd_b2(persistent, fp32, comp, blocks);
    case 3: return occupancy_fwd_b3(persistent, fp32, comp, blocks);
    case 4: return occupancy_fwd_b4(persistent, fp32, comp, blocks);
    case 5: return occupancy_fwd_b5(persistent, fp32, comp, blocks);
    case 6: return occupancy_fwd_b6(persistent, fp32, comp, blocks);
    case 7: return occupancy_fwd_b7(persistent, fp32, comp, blocks);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_finalize(bool comp, const LaunchArgs &a, int item0, int n_items,
                           int n_heads, hipStream_t stream) {
  const int grid = n_items > 0 ? n_items : 1;
  if (comp)
    hipLaunchKernelGGL(perm_finalize_kernel<true>, dim3(grid), dim3(kWave), 0, stream, a,
                       item0, n_heads);
  else
    hipLaunchKernelGGL(perm_finalize_kernel<false>, dim3(grid), dim3(kWave), 0, stream, a,
                       item0, n_heads);
  return hipGetLastError();
}
#endif

}  // namespace pqp
