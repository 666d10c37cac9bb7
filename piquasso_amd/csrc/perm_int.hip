// Exact permanents of integer matrices on gfx950: the Glynn sweep of
// perm_fwd.hip (same plan, same Gray order, src/permanent.cpp:49-264) with
// every product and sum taken modulo K <= 8 primes just below 2^49; the host
// recovers the integer from the residues by the Chinese remainder theorem.
//
// Row sums of an integer matrix are exact integers in FP64 (the host checks
// that they stay below 2^48), so the row updates are the ordinary ones.  The
// product prod_j cs_j (columns of multiplicity c are planned as c unit
// columns) is formed in groups of G columns exactly in FP64 (the host picks
// G with every group product below 2^49), and the group products are
// multiplied modulo each prime with a floating-point modular multiply:
//
//   h = x y, l = fma(x, y, -h)        (x y = h + l exactly)
//   q = rint(h * (1/p)), r = fma(-q, p, h) + l,
//
// exact for |x|, |y| <= 2p < 2^50: |h - q p| <= 1.5 p is an integer below
// 2^53 (the fma rounds nothing), |l| <= 2^46, so |r| <= 1.75 p stays a
// valid input.  Sums are reduced by r = fma(-rint(s / p), p, s) (exact for
// integers |s| < 2^53).  Each lane accumulates coef * product per prime,
// multiplies by its top weight, the wave reduces the residues with a
// butterfly, and the finalize kernel adds the units' residues, multiplies by
// 2^-(n-1) mod p and normalises to [0, p).

#include <hip/hip_runtime.h>

#include <cstdint>

#include "perm_kernels.hpp"
#include "pqp_types.hpp"
#include "sweep_common.hpp"

namespace pqp {

namespace {

__device__ __forceinline__ double reduce_mod(double x, double p, double pinv) {
  const double q = __builtin_rint(x * pinv);
  return __builtin_fma(-q, p, x);
}

__device__ __forceinline__ double mul_mod(double x, double y, double p, double pinv) {
  const double h = x * y;
  const double l = __builtin_fma(x, y, -h);
  const double q = __builtin_rint(h * pinv);
  return __builtin_fma(-q, p, h) + l;
}

template <int MCP, int G>
__device__ __forceinline__ void int_unit(const LaunchArgs &a, const IntArgs &ia, int item,
                                         int local) {
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const int64_t U = it->U;
  const int S = uniform(static_cast<int>(it->S_in));
  const int L = it->L;
  const int Dt = it->Dt;
  const int flags = uniform(it->flags);
  const int K = ia.n_primes;
  const cptr<double> mod = as_const<double>(a.arena, ia.mod_off);  // p[8], 1/p[8], inv 2^(n-1)[8]

  int64_t u = static_cast<int64_t>(local) * kWave + lane;
  const bool active = u < U;
  if (!active) u = U - 1;

  const cptr<double> Bin = as_const<double>(a.arena, it->b2_off);  // inner rows of B (complex)
  const cptr<InnerEntry> tab = as_const<InnerEntry>(a.arena, it->step_off);

  double cr[MCP], ci[MCP];
  double w_top;
  start_state<MCP, double>(a, it, u, L, Dt, flags, cr, ci, w_top);
  const double lane_w = active ? ((u & 1) ? -w_top : w_top) : 0.0;

  double acc[kIntMaxPrimes];
#pragma unroll
  for (int k = 0; k < kIntMaxPrimes; ++k) acc[k] = 0.0;
  constexpr int NG = (MCP + G - 1) / G;

  InnerEntry e = inner_entry(tab, 0);
  for (int v = 0; v < S; ++v) {
    const cptr<double> row = Bin + static_cast<int64_t>(e.row) * (2 * MCP);
    const InnerEntry en = inner_entry(tab, v + 1);
    // exact group products (|product| <= 2^49)
    double gp[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      gp[g] = cr[g * G];
#pragma unroll
      for (int q = 1; q < G; ++q)
        if (g * G + q < MCP) gp[g] *= cr[g * G + q];
    }
    const double coef = e.coef;
    const bool small = __builtin_fabs(coef) <= 8.0;
#pragma unroll
    for (int k = 0; k < kIntMaxPrimes; ++k) {
      if (k < K) {
        const double p = mod[k], pinv = mod[kIntMaxPrimes + k];
        double x = gp[0];
#pragma unroll
        for (int g = 1; g < NG; ++g) x = mul_mod(x, gp[g], p, pinv);
        if (small) {
          acc[k] = reduce_mod(__builtin_fma(coef, x, acc[k]), p, pinv);
        } else {
          __asm__ volatile("");
          acc[k] = reduce_mod(acc[k] + mul_mod(x, reduce_mod(coef, p, pinv), p, pinv), p, pinv);
        }
      }
    }
    const double s2 = in_vgpr(static_cast<double>(e.sign2));
#pragma unroll
    for (int j = 0; j < MCP; ++j) cr[j] = __builtin_fma(s2, row[2 * j], cr[j]);
    e = en;
  }
  // lane weight, then the wave's sum of residues (butterfly, reduced each step)
  double *out = a.partials + static_cast<int64_t>(kIntMaxPrimes) * (it->partial_base + local);
#pragma unroll
  for (int k = 0; k < kIntMaxPrimes; ++k) {
    if (k < K) {
      const double p = mod[k], pinv = mod[kIntMaxPrimes + k];
      double r = mul_mod(acc[k], reduce_mod(lane_w, p, pinv), p, pinv);
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) r = reduce_mod(r + __shfl_xor(r, off), p, pinv);
      if (lane == 0) out[k] = r;
    }
  }
}

template <int MCP>
__device__ __forceinline__ void int_dispatch(const LaunchArgs &a, const IntArgs &ia, int item,
                                             int local) {
  if (ia.group == 4)
    int_unit<MCP, 4>(a, ia, item, local);
  else if (ia.group == 2)
    int_unit<MCP, 2>(a, ia, item, local);
  else
    int_unit<MCP, 1>(a, ia, item, local);
}

}  // namespace

// One block (one wave) per work unit; MCP is the plan's padded width.
template <int MCP>
__global__ __launch_bounds__(kWave) void perm_int_kernel(LaunchArgs a, IntArgs ia, int first_unit) {
  const cptr<WorkUnit> wu = as_const<WorkUnit>(a.units) + (first_unit + static_cast<int>(blockIdx.x));
  const int item = uniform(wu->item);
  const int local = uniform(wu->local);
  int_dispatch<MCP>(a, ia, item, local);
}

// One block of kIntFinalizeThreads per item: for each prime, the threads
// add strided subsets of the units' residues, the block reduces them
// through LDS (modular sums are exact, so the order does not matter), and
// thread 0 multiplies by 2^-(n-1) mod p and writes the residue in [0, p).
constexpr int kIntFinalizeThreads = 256;

__global__ __launch_bounds__(kIntFinalizeThreads) void perm_int_finalize_kernel(LaunchArgs a,
                                                                               IntArgs ia) {
  __shared__ double red[kIntFinalizeThreads];
  const int item = static_cast<int>(blockIdx.x);
  const int tid = static_cast<int>(threadIdx.x);
  if (item >= a.n_items) return;
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const cptr<double> mod = as_const<double>(a.arena, ia.mod_off);
  const int units = it->units;
  const double *part = a.partials + static_cast<int64_t>(kIntMaxPrimes) * it->partial_base;
  for (int k = 0; k < ia.n_primes; ++k) {
    const double p = mod[k], pinv = mod[kIntMaxPrimes + k], inv2 = mod[2 * kIntMaxPrimes + k];
    double s = 0.0;
    for (int u = tid; u < units; u += kIntFinalizeThreads)
      s = reduce_mod(s + part[static_cast<int64_t>(kIntMaxPrimes) * u + k], p, pinv);
    red[tid] = s;
    __syncthreads();
    for (int w = kIntFinalizeThreads / 2; w > 0; w >>= 1) {
      if (tid < w) red[tid] = reduce_mod(red[tid] + red[tid + w], p, pinv);
      __syncthreads();
    }
    if (tid == 0) {
      double r = 1.0;  // empty selection: 1
      if (it->n > 0) {
        r = reduce_mod(mul_mod(red[0], inv2, p, pinv), p, pinv);
        if (r < 0.0) r += p;
      }
      a.out[static_cast<int64_t>(kIntMaxPrimes) * it->out_index + k] = r;
    }
    __syncthreads();
  }
}

hipError_t launch_int(int mcp, const LaunchArgs &a, const IntArgs &ia, int first_unit,
                      int n_units, hipStream_t stream) {
  if (n_units <= 0) return hipSuccess;
  switch (mcp) {
#define PQP_INT_CASE(W)                                                                          \
  case W:                                                                                        \
    hipLaunchKernelGGL(perm_int_kernel<W>, dim3(n_units), dim3(kWave), 0, stream, a, ia,         \
                       first_unit);                                                              \
    break;
    PQP_INT_CASE(8)
    PQP_INT_CASE(16)
    PQP_INT_CASE(24)
    PQP_INT_CASE(32)
    PQP_INT_CASE(40)
    PQP_INT_CASE(48)
    PQP_INT_CASE(56)
    PQP_INT_CASE(64)
#undef PQP_INT_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_int_finalize(const LaunchArgs &a, const IntArgs &ia, int n_items,
                               hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(perm_int_finalize_kernel, dim3(n_items), dim3(kIntFinalizeThreads), 0, stream,
                     a, ia);
  return hipGetLastError();
}

}  // namespace pqp
