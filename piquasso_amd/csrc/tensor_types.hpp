// Device image of a batch of tensor permanents (uniformly partially
// distinguishable bosons), shared by the planner in capi.cpp and the
// kernels in tensor_perm.hip.
//
// T(A, x) = 4^-(n-1) sum over sign vectors delta, eps in {+1,-1}^n with
//           delta_0 = eps_0 = +1 of (prod delta)(prod eps)
//           prod_i ( x R_delta[i] conj(R_eps[i]) + (1 - x) Q_{delta eps}[i] ),
// the value piquasso sums over column subsets
// (piquasso/_simulators/passive/probabilities.py:193-237).
//
// Work split: sign vector delta is s (bit b set: delta_{b+1} = -1), one per
// lane, s = (s_hi << ls) | lane; a wave walks eps = t = (t_blk << kt) |
// gray(k), k = 0 .. 2^kt - 1, flipping one sign per step.  Term (t, s) is
// the conjugate of term (s, t), so only Re is summed, with weight 2 when
// t's lane block (t >> ls) is above s_hi, 1 when it is s_hi, and pairs
// below are not visited.
#pragma once

#include <cstdint>

namespace pqp {

constexpr int kTensorMaxN = 20;  // widest matrix the kernels are built for

struct alignas(16) TensorItem {
  int64_t cols_off;   // arena byte offset: n x n complex, column-major [c][i] (re, im)
  int64_t pw_off;     // arena byte offset: n x n real, (1 - x) |A[i, c]|^2 at [c][i]
  double x;           // particle overlap
  int32_t n;          // matrix order (particles)
  int32_t ls;         // lane bits: 2^ls lanes carry s (ls = min(n - 1, 6))
  int32_t kt;         // walk bits: 2^kt Gray steps of t per work unit
  int32_t unit0;      // first work unit of this item
  int32_t n_units;    // work units of this item
  int32_t pad;
};
static_assert(sizeof(TensorItem) == 48, "TensorItem layout");

struct TensorUnit {
  int32_t item;
  int32_t s_hi;       // s = (s_hi << ls) | lane
  int32_t t_blk;      // t = (t_blk << kt) | gray(k)
  int32_t mixed;      // 1: some steps have t >> ls <= s_hi (weight per step)
};

struct TensorArgs {
  const char *arena;
  const TensorItem *items;
  const TensorUnit *units;
  double *partials;   // 3 doubles per unit: dd sum (hi, lo), sum of |terms|
  double *out;        // 2 doubles per item: T, sum over pairs of weight x |Re term|
  int32_t n_units;
  int32_t n_items;
};

}  // namespace pqp
