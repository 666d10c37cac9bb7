// pqp_permanent_c128_noalloc: the single-permanent entry point without any
// device or pinned-host allocation, for callers that may not allocate
// (serving loops that forbid hipMalloc, harnesses that account for every
// allocation).
//
// All device memory the call needs beyond the caller's buffers (the plan
// arena the sweep kernels read, their per-wave partial sums, the result
// slots) is carved out of `scratch`, a file-scope device array of
// kScratchBytes.  Each GPU holds its own copy of the array for the life of
// the process, at its own address: the host looks that address up once per
// device with hipGetSymbolAddress (the name `scratch` in host code is only a
// host-side stand-in) and passes the looked-up pointer to the kernels.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; nothing in it has to start at zero (the arena is uploaded by each
// call, the sweep writes every partial slot it owns before the finalize
// reads it, and the result slots are written by the kernels).  Calls from
// this library serialise on a mutex; callers that drive the same array from
// several threads through other means would race.
//
// The kernels are those of pqp_permanent_c128 for the same plan (the
// one-block launch for small permanents, else static launches of one block
// per work unit and the finalize kernel), on the null stream, so the
// results are bit-identical to that entry point's.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>

#include "capi_internal.hpp"
#include "perm_kernels.hpp"
#include "plan.hpp"
#include "pqp_permanent.h"
#include "pqp_types.hpp"

namespace {

// 8 MiB holds the plan of any permanent of up to 2^39 Gray terms (n = 40 at
// unit multiplicities, seconds of GPU time): at most 131072 work units (8
// bytes each, plus 32 bytes of partial sums) and a 2^16-entry inner table of
// 16-byte entries.  Larger permanents are rejected.
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

using namespace pqp;

std::mutex g_mu;
char *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

size_t align256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

// Device address of the current device's copy of `scratch` (looked up on
// the first call on that device).
int scratch_of_current_device(char **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return capi_fail(PQP_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return capi_fail(PQP_ERR_DEVICE, "device index " + std::to_string(dev) + " beyond the scratch table");
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return capi_fail(PQP_ERR_DEVICE,
                       std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<char *>(p);
  }
  *ptr = g_scratch[dev];
  return PQP_OK;
}

struct Range {
  explicit Range(const char *name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
};

#define PQP_NA_HIP(expr)                                                                \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return capi_fail(PQP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" int pqp_permanent_c128_noalloc(const double *matrix, int64_t m_r, int64_t m_c,
                                          int64_t ld, const int32_t *rows, const int32_t *cols,
                                          uint32_t flags, double *out, double *out_dd) {
  Range range_("pqp_permanent_c128_noalloc");
  if (!out) return capi_fail(PQP_ERR_INVALID, "out is null");
  const PermInput in{matrix, false, m_r, m_c, ld, rows, cols};
  Plan plan;
  std::string err;
  if (m_r < 0 || (m_r > 0 && !rows)) return capi_fail(PQP_ERR_INVALID, "invalid shape or null pointer");
  int st = plan_build(std::vector<PermInput>{in}, single_plan_options(flags, in), plan, err);
  if (st != PQP_OK) return capi_fail(st, err);

  // scratch layout: [arena][partials: 4 doubles per slot][term magnitudes:
  // 1 double per slot][results: 2 + 4 + 1 doubles, then the done slot of a
  // one-block launch]
  const size_t arena_bytes = plan.arena.size();
  const size_t partials_off = align256(arena_bytes);
  const size_t n_slots = static_cast<size_t>(std::max<int64_t>(plan.n_partials, 1));
  const size_t results_off = align256(partials_off + n_slots * 5 * sizeof(double));
  const size_t needed = results_off + 8 * sizeof(double);
  if (needed > kScratchBytes)
    return capi_fail(PQP_ERR_RANGE, "this permanent needs " + std::to_string(needed) +
                                        " bytes of device scratch; the allocation-free entry "
                                        "point has " + std::to_string(kScratchBytes));
  if ((st = capi_require_device()) != PQP_OK) return st;

  std::lock_guard<std::mutex> lock(g_mu);
  char *base = nullptr;
  if ((st = scratch_of_current_device(&base)) != PQP_OK) return st;

  // pageable host source and destination: synchronous copies on the null stream
  PQP_NA_HIP(hipMemcpy(base, plan.arena.data(), arena_bytes, hipMemcpyHostToDevice));
  LaunchArgs a;
  a.arena = base;
  a.items = reinterpret_cast<const ItemDesc *>(base + plan.items_off);
  a.units = reinterpret_cast<const WorkUnit *>(base + plan.units_off);
  a.partials = reinterpret_cast<double *>(base + partials_off);
  a.part_abs = a.partials + 4 * n_slots;
  a.out = reinterpret_cast<double *>(base + results_off);
  a.out_dd = a.out + 2;
  a.out_abs = a.out + 6;
  a.queue_head = nullptr;  // static launches use no queue
  a.n_units = static_cast<int32_t>(plan.units.size());
  a.n_items = static_cast<int32_t>(plan.items.size());
  const bool comp = !(flags & PQP_FLAG_PLAIN) || plan.fp32;
  int bucket = -1;
  if (single_block_eligible(plan, flags, &bucket)) {
    // the kernel pqp_permanent_c128 runs for this plan (same sums, same
    // order: bit-identical results); its done slot is scratch here, and the
    // synchronous copy below waits for the launch
    static thread_local SingleArgs args;
    single_block_args(plan, a, reinterpret_cast<uint64_t *>(a.out + 7), 0, args);
    PQP_NA_HIP(launch_single_block(bucket, plan.fp32, args, nullptr));
  } else {
    for (const Plan::Range &r : plan.ranges) {
      FwdLaunch l;
      l.args = a;
      l.first_unit = r.first;
      l.n_units = r.count;
      l.persistent = false;
      l.fp32 = plan.fp32;
      l.comp = comp;
      l.stream = nullptr;
      PQP_NA_HIP(launch_fwd(r.bucket, l));
    }
    if (!plan.all_direct)
      PQP_NA_HIP(launch_finalize(comp, a, 0, a.n_items, 0, plan.max_units, nullptr));
  }
  double res[7];
  PQP_NA_HIP(hipMemcpy(res, base + results_off, sizeof(res), hipMemcpyDeviceToHost));
  out[0] = res[0];
  out[1] = res[1];
  if (out_dd) std::memcpy(out_dd, res + 2, 4 * sizeof(double));
  const int n = plan.items.empty() ? 0 : plan.items[0].n;
  capi_set_last_bound(n == 0 ? 0.0
                      : comp ? sweep_error_bound(n, plan.fp32, res[6], res[0], res[1])
                             : std::numeric_limits<double>::quiet_NaN());
  return PQP_OK;
}
