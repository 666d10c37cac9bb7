// Routes kernel launches and occupancy queries to the translation unit that
// instantiates the requested (column-width bucket, arithmetic variant).
#include "perm_kernels.hpp"

namespace pqp {

#define PQP_CASE(k)                                        \
  case k:                                                  \
    if (l.fp32) return launch_fwd_b##k##_f(l);             \
    return l.comp ? launch_fwd_b##k##_d(l) : launch_fwd_b##k##_p(l);

hipError_t launch_single_block(int bucket, bool fp32, const SingleArgs &args, hipStream_t stream) {
  switch (bucket) {
    case 0:
      return fp32 ? launch_single_block_b0_f(args, stream) : launch_single_block_b0_d(args, stream);
    case 1:
      return fp32 ? launch_single_block_b1_f(args, stream) : launch_single_block_b1_d(args, stream);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_fwd(int bucket, const FwdLaunch &l) {
  if (l.fp32 && !l.comp) return hipErrorInvalidValue;  // FP32 products always use dd
  switch (bucket) {
    PQP_CASE(0)
    PQP_CASE(1)
    PQP_CASE(2)
    PQP_CASE(3)
    PQP_CASE(4)
    PQP_CASE(5)
    PQP_CASE(6)
    PQP_CASE(7)
    default:
      return hipErrorInvalidValue;
  }
}
#undef PQP_CASE

#define PQP_CASE(k)                                                        \
  case k:                                                                  \
    if (fp32) return occupancy_fwd_b##k##_f(persistent, blocks);           \
    return comp ? occupancy_fwd_b##k##_d(persistent, blocks)               \
                : occupancy_fwd_b##k##_p(persistent, blocks);

hipError_t occupancy_fwd(int bucket, bool persistent, bool fp32, bool comp, int *blocks) {
  switch (bucket) {
    PQP_CASE(0)
    PQP_CASE(1)
    PQP_CASE(2)
    PQP_CASE(3)
    PQP_CASE(4)
    PQP_CASE(5)
    PQP_CASE(6)
    PQP_CASE(7)
    default:
      return hipErrorInvalidValue;
  }
}
#undef PQP_CASE

hipError_t launch_grad(int bucket, const LaunchArgs &a, int first_unit, int n_units,
                       hipStream_t stream) {
  switch (bucket) {
    case 0: return launch_grad_b0(a, first_unit, n_units, stream);
    case 1: return launch_grad_b1(a, first_unit, n_units, stream);
    case 2: return launch_grad_b2(a, first_unit, n_units, stream);
    case 3: return launch_grad_b3(a, first_unit, n_units, stream);
    case 4: return launch_grad_b4(a, first_unit, n_units, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_minors(int bucket, const LaunchArgs &a, int first_unit, int n_units,
                         hipStream_t stream) {
  switch (bucket) {
    case 0: return launch_minors_b0(a, first_unit, n_units, stream);
    case 1: return launch_minors_b1(a, first_unit, n_units, stream);
    case 2: return launch_minors_b2(a, first_unit, n_units, stream);
    case 3: return launch_minors_b3(a, first_unit, n_units, stream);
    case 4: return launch_minors_b4(a, first_unit, n_units, stream);
    case 5: return launch_minors_b5(a, first_unit, n_units, stream);
    case 6: return launch_minors_b6(a, first_unit, n_units, stream);
    case 7: return launch_minors_b7(a, first_unit, n_units, stream);
    case 8: return launch_minors_b8(a, first_unit, n_units, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_tensor(int n, const TensorArgs &a, int first_unit, int n_units, hipStream_t s) {
  if (n >= 1 && n <= 12) return launch_tensor_p0(n, a, first_unit, n_units, s);
  if (n >= 13 && n <= kTensorMaxN) return launch_tensor_p1(n, a, first_unit, n_units, s);
  return hipErrorInvalidValue;
}

}  // namespace pqp
