// Host-side launch interface of the permanent kernels.  The kernel
// templates are instantiated once per column-width bucket (perm_fwd.hip is
// compiled once per bucket with -DPQP_BUCKET_ID=k); launch_fwd() routes a
// launch to the right translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include "pqp_types.hpp"

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

#define PQP_DECLARE_BUCKET(k)                                            \
  hipError_t launch_fwd_b##k(const FwdLaunch &l);                        \
  hipError_t occupancy_fwd_b##k(bool persistent, bool fp32, bool comp,   \
                                int *blocks_per_cu);
PQP_DECLARE_BUCKET(0)
PQP_DECLARE_BUCKET(1)
PQP_DECLARE_BUCKET(2)
PQP_DECLARE_BUCKET(3)
PQP_DECLARE_BUCKET(4)
PQP_DECLARE_BUCKET(5)
PQP_DECLARE_BUCKET(6)
PQP_DECLARE_BUCKET(7)
#undef PQP_DECLARE_BUCKET

hipError_t launch_fwd(int bucket, const FwdLaunch &l);
hipError_t occupancy_fwd(int bucket, bool persistent, bool fp32, bool comp, int *blocks_per_cu);
// Finalize items [item0, item0 + n_items); block 0 also rewinds n_heads
// queue heads (pass 0 when a persistent launch may still be running).
hipError_t launch_finalize(bool comp, const LaunchArgs &a, int item0, int n_items,
                           int n_heads, hipStream_t stream);

}  // namespace pqp
