// Host-side launch interface of the permanent kernels.  The kernel
// templates are instantiated once per column-width bucket and arithmetic
// variant (perm_fwd.hip is compiled once per pair); launch_fwd() in
// dispatch.cpp routes a launch to the right translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include "pqp_types.hpp"
#include "tensor_types.hpp"
#include "poly_types.hpp"

namespace pqp {

struct FwdLaunch {
  LaunchArgs args;
  int first_unit = 0;   // first unit of this bucket's range in args.units
  int n_units = 0;      // units in the range
  int head_slot = 0;    // queue head used by a persistent launch
  int grid = 0;         // persistent launches: number of blocks
  bool persistent = true;
  bool fp32 = false;    // FP32 row sums and products
  bool comp = true;     // double-double accumulation (false: plain FP64)
  hipStream_t stream = nullptr;
};

// Per (bucket, variant) entry points, defined in perm_fwd.hip builds;
// variant suffix d = FP64 + dd, p = plain FP64, f = FP32 products + dd.
#define PQP_DECLARE(k, v)                                         \
  hipError_t launch_fwd_b##k##_##v(const FwdLaunch &l);           \
  hipError_t occupancy_fwd_b##k##_##v(bool persistent, int *blocks_per_cu);
#define PQP_DECLARE_BUCKET(k) PQP_DECLARE(k, d) PQP_DECLARE(k, p) PQP_DECLARE(k, f)
PQP_DECLARE_BUCKET(0)
PQP_DECLARE_BUCKET(1)
PQP_DECLARE_BUCKET(2)
PQP_DECLARE_BUCKET(3)
PQP_DECLARE_BUCKET(4)
PQP_DECLARE_BUCKET(5)
PQP_DECLARE_BUCKET(6)
PQP_DECLARE_BUCKET(7)
#undef PQP_DECLARE_BUCKET
#undef PQP_DECLARE

// One-launch single permanents (perm_single_block_kernel), widths 1..16:
// d = FP64 + dd, f = FP32 products + dd.
hipError_t launch_single_block_b0_d(const SingleArgs &args, hipStream_t s);
hipError_t launch_single_block_b0_f(const SingleArgs &args, hipStream_t s);
hipError_t launch_single_block_b1_d(const SingleArgs &args, hipStream_t s);
hipError_t launch_single_block_b1_f(const SingleArgs &args, hipStream_t s);
hipError_t launch_single_block(int bucket, bool fp32, const SingleArgs &args, hipStream_t stream);

constexpr int kBuckets = 8;        // forward kernel width buckets
constexpr int kMinorBuckets = 9;   // column-minor kernel buckets
constexpr int kMaxRanges = 9;      // concurrent launches of one plan (one per bucket)

// Column-minor kernels (perm_minors.hip), buckets 0..3 (widths up to 32).
hipError_t launch_minors_b0(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b1(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b2(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b3(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b4(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b5(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b6(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b7(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors_b8(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_minors(int bucket, const LaunchArgs &a, int first_unit, int n_units,
                         hipStream_t stream);
hipError_t launch_minors_finalize(const LaunchArgs &a, int n_items, int max_units,
                                  hipStream_t stream);

// Gradient kernels (perm_grad.hip), unit column multiplicities: buckets for
// widths 1-8, 9-16, 17-19, 20, 21-32.
constexpr int kGradBuckets = 5;
hipError_t launch_grad_b0(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_grad_b1(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_grad_b2(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_grad_b3(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_grad_b4(const LaunchArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_grad(int bucket, const LaunchArgs &a, int first_unit, int n_units,
                       hipStream_t stream);
// max_values: the largest 1 + rows of B x m_c over the items (the
// permanent and its gradient entries)
hipError_t launch_grad_finalize(const LaunchArgs &a, int n_items, int max_values,
                                hipStream_t stream);

// Exact integer permanents (perm_int.hip): one launch per padded width
// (a multiple of 8), then the finalize; 8 residues per unit and per item.
hipError_t launch_int(int mcp, const LaunchArgs &a, const IntArgs &ia, int first_unit,
                      int n_units, hipStream_t stream);
hipError_t launch_int_finalize(const LaunchArgs &a, const IntArgs &ia, int n_items,
                               hipStream_t stream);

// Tensor permanents (tensor_perm.hip): one launch per matrix order n over
// units [first_unit, first_unit + n_units), then one finalize for all items.
hipError_t launch_tensor_p0(int n, const TensorArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_tensor_p1(int n, const TensorArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_tensor(int n, const TensorArgs &a, int first_unit, int n_units, hipStream_t s);
hipError_t launch_tensor_finalize(const TensorArgs &a, hipStream_t stream);

// Matrix-polynomial permanent (perm_poly.hip): the subset sweep over
// a.n_wg workgroups (the polynomial in LDS when `lds`; else in global memory
// with `group` (1 or kPolyGroup) subsets interleaved; dynamic LDS of
// lds_bytes), then the reduction of the workgroups' accumulators.
hipError_t launch_perm_poly(const PolyArgs &a, bool lds, int group, size_t lds_bytes,
                            hipStream_t stream);

hipError_t launch_fwd(int bucket, const FwdLaunch &l);
hipError_t occupancy_fwd(int bucket, bool persistent, bool fp32, bool comp, int *blocks_per_cu);
// Finalize items [item0, item0 + n_items); block 0 also rewinds n_heads
// queue heads (pass 0 when a persistent launch may still be running).
hipError_t launch_finalize(bool comp, const LaunchArgs &a, int item0, int n_items,
                           int n_heads, int max_units, hipStream_t stream);

}  // namespace pqp
