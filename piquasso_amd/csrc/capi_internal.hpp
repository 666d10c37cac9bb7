// Helpers of the C ABI shared between capi.cpp and the translation units
// that define further entry points (noalloc.hip): the thread-local error
// message, the device check, and the planning options of single-permanent
// calls, so that every single-permanent entry plans the same way.
#pragma once

#include <cstdint>
#include <string>

#include "plan.hpp"

namespace pqp {

// Record `msg` as the calling thread's pqp_last_error() and return `code`.
int capi_fail(int code, const std::string &msg);

// PQP_OK when a HIP device is usable, else PQP_ERR_DEVICE with the message.
int capi_require_device();

// Planning options of a single-permanent call with these PQP_FLAG_* flags:
// about 1.4 sqrt(T) waves (capped per compute unit), one-wave permanents
// write their result from the sweep kernel.
PlanOptions single_plan_options(uint32_t flags);

}  // namespace pqp
