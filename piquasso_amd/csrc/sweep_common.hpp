// Device code shared by the sweep kernels (perm_fwd.hip, perm_minors.hip):
// scalar-load addressing of the read-only arena, FMA complex multiply,
// double-double wave reduction, and the start state of a lane (row sums,
// top-digit weight) for the Glynn sweep over the reflected mixed-radix Gray
// code of src/n_aryGrayCodeCounter.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd.hpp"
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

// Entry v of an inner table (scalar loads).
__device__ __forceinline__ InnerEntry inner_entry(cptr<InnerEntry> tab, int v) {
  InnerEntry e;
  e.coef = tab[v].coef;
  e.sign2 = tab[v].sign2;
  e.row = tab[v].row;
  return e;
}

// The partner's value at step STEP of a butterfly over the 64 lanes: lanes
// l ^ 1 and l ^ 2 (DPP quad permutations), then 7 - l and 15 - l within
// groups of 8 and 16 (DPP half-row and row mirrors: after the previous steps
// the two halves of each group hold the sums of disjoint lanes), then the
// rows 16 and 32 lanes apart (v_permlane16_swap / v_permlane32_swap, which
// exchange whole rows between a register pair; with the value in both
// registers a lane of the lower row finds its partner's value in the second
// result, a lane of the upper row in the first).  No LDS traffic: DPP and
// the permlane swaps are VALU operations.
template <int STEP>
__device__ __forceinline__ double wave_partner(double x, int lane) {
  const uint64_t v = __builtin_bit_cast(uint64_t, x);
  uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
  if constexpr (STEP < 4) {
    constexpr int kCtrl = STEP == 0 ? 0xB1 : STEP == 1 ? 0x4E : STEP == 2 ? 0x141 : 0x140;
    lo = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(lo), static_cast<int>(lo),
                                                           kCtrl, 0xF, 0xF, false));
    hi = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(hi), static_cast<int>(hi),
                                                           kCtrl, 0xF, 0xF, false));
  } else {
    const auto l = STEP == 4 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                             : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = STEP == 4 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                             : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    const bool upper = (lane & (STEP == 4 ? 16 : 32)) != 0;
    lo = upper ? l[0] : l[1];
    hi = upper ? h[0] : h[1];
  }
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

__device__ __forceinline__ int lane_id() {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
}

// Butterfly reduction of a complex dd value over the 64 lanes (every lane
// holds the sum; the additions are symmetric, so all lanes agree).  Every
// lane of the wave must be active.
// Steps STEP..END-1 only: END = 4 sums groups of 16 lanes.
template <int STEP = 0, int END = 6>
__device__ __forceinline__ cdd wave_reduce(cdd v, int lane) {
  if constexpr (STEP < END) {
    const cdd o{{wave_partner<STEP>(v.re.hi, lane), wave_partner<STEP>(v.re.lo, lane)},
                {wave_partner<STEP>(v.im.hi, lane), wave_partner<STEP>(v.im.lo, lane)}};
    return wave_reduce<STEP + 1, END>(cdd_add(v, o), lane);
  } else {
    return v;
  }
}
__device__ __forceinline__ cdd wave_reduce(cdd v) { return wave_reduce<0, 6>(v, lane_id()); }

// Plain FP64 sum over the 64 lanes (or groups of 16 for END = 4), same
// butterfly.
template <int STEP = 0, int END = 6>
__device__ __forceinline__ double wave_sum(double x, int lane) {
  if constexpr (STEP < END) {
    return wave_sum<STEP + 1, END>(x + wave_partner<STEP>(x, lane), lane);
  } else {
    return x;
  }
}

// Waves per SIMD the register allocator must leave room for: the row sums
// take 4 VGPRs per column in FP64 (2 in FP32); the inner loop is pure VALU
// with scalar operands, so 2-3 waves per SIMD hide the scalar-load latency.
constexpr int waves_for(int hi, bool fp32) {
  const int w = fp32 ? hi / 2 : hi;
  return w <= 8 ? 6 : w <= 16 ? 4 : w <= 32 ? 3 : w <= 48 ? 2 : 1;
}

// Row sums of lane state u (top state u, inner digits all 0) and the lane's
// top weight prod_top C(rho_t, g_t).
// Rows of B of item `it`: in the arena, or in the kernel arguments of a
// one-block single-permanent launch (a.b_rows).
template <typename R>
__device__ __forceinline__ cptr<R> b_rows(const LaunchArgs &a, const cptr<ItemDesc> it) {
  return a.b_rows ? as_const<R>(a.b_rows) : as_const<R>(a.arena, it->b_off);
}

template <int MCP, typename R, bool UNIT_BASE = false>
__device__ __forceinline__ void start_state(const LaunchArgs &a, const cptr<ItemDesc> it,
                                            int64_t u, int L, int Dt, int flags,
                                            R (&cr)[MCP], R (&ci)[MCP], double &w_top) {
  const cptr<R> B = b_rows<R>(a, it);
  const cptr<double> alpha = as_const<double>(a.arena, it->alpha_off);
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
  const uint64_t gbin = static_cast<uint64_t>(u) ^ (static_cast<uint64_t>(u) >> 1);
  if (UNIT_BASE && a.unit_base) {
    // one-block launch, binary top digits: the unit's start row (host sum of
    // the rows common to its 64 lanes) plus the lane-dependent digits
    const cptr<R> base = as_const<R>(a.unit_base) + (u / kWave) * 2 * MCP;
#pragma unroll
    for (int j = 0; j < MCP; ++j) {
      cr[j] = base[2 * j];
      ci[j] = base[2 * j + 1];
    }
    w_top = 1.0;
    for (int t = (Dt < kLaneDigits ? Dt : kLaneDigits) - 1; t >= 0; --t) {
      const R al = ((gbin >> t) & 1) ? R(-1) : R(1);
      const cptr<R> row = B + static_cast<int64_t>(L + 1 + t) * 2 * MCP;
#pragma unroll
      for (int j = 0; j < MCP; ++j) {
        cr[j] = fmaR(al, row[2 * j], cr[j]);
        ci[j] = fmaR(al, row[2 * j + 1], ci[j]);
      }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < MCP; ++j) {
    cr[j] = R(0);
    ci[j] = R(0);
  }
  const cptr<double> top = as_const<double>(a.arena, it->top_off);
  const double ud = static_cast<double>(u) + 0.5;
  double q_hi = 0.0;
  int parity = 0;
  w_top = 1.0;
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
}

}  // namespace

}  // namespace pqp
