// Coefficients of the permanent of a matrix polynomial on gfx950,
//
//   c_e = [x^e] perm(B_0 + sum_m x_m B_m),
//
// the quantity behind piquasso's lossy, partially distinguishable detection
// probabilities (B_0 = G o (I - T^dag T), B_m = G o t_m t_m^dag;
// piquasso/_simulators/passive/probabilities.py:244-384).  Ryser's formula
// over the non-empty column subsets S turns every row into a linear factor,
//
//   c_e = sum_S (-1)^(n - |S|) [x^e] prod_i ( c_i(S) + sum_m l_{m,i}(S) x_m ),
//   c_i(S) = sum_{j in S} B_0[i, j],   l_{m,i}(S) = sum_{j in S} B_m[i, j],
//
// and the product is built row by row as a truncated polynomial.  The
// reference keeps the polynomial on the box 0 <= a_m <= max e_m; here it is
// kept on the downward closure of the requested exponents (every monomial
// that divides one of them), which is all the recursion ever reads and, for
// a table of all outputs with at most n photons, the simplex instead of the
// box ((n + M choose M) instead of (n + 1)^M coefficients).  The monomials
// are stored by total degree; multiplying by a factor updates degree D from
// degree D - 1 (its "links" a - e_m) and itself, so the levels are updated
// from the top down with a workgroup barrier between levels, in place.
//
// One workgroup per group of K subsets at a time (groups strided over the
// grid).  The polynomial lives in LDS when it fits (kLds, K = 1; 256
// threads, several workgroups per CU), else in a per-workgroup global-memory
// slice (1024 threads, two workgroups per CU, so the slices and
// accumulators stay few; K = 2 subsets interleaved coefficient by
// coefficient, so a gather fetches 32 contiguous bytes for two lanes; each
// lane issues four links' loads and gathers before its first FMA waits).
// Each workgroup adds its subsets' signed coefficients into its own FP64
// accumulator; perm_poly_reduce_kernel then sums the workgroups'
// accumulators in double-double in workgroup order, so results do not
// depend on timing.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd.hpp"
#include "perm_kernels.hpp"
#include "poly_types.hpp"

namespace pqp {

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(__builtin_fma(a.x, b.x, __builtin_fma(-a.y, b.y, c.x)),
                      __builtin_fma(a.x, b.y, __builtin_fma(a.y, b.x, c.y)));
}

template <bool kLds, int kThreads, int K>
__global__ __launch_bounds__(kThreads) void perm_poly_kernel(PolyArgs a) {
  extern __shared__ double2 smem[];
  const int n = a.n, M = a.M, L = a.L;
  const int tid = static_cast<int>(threadIdx.x);
  const int wg = static_cast<int>(blockIdx.x);
  const int per = n * (M + 1);  // row sums of one subset
  // LDS: row sums rs[(k * n + i) * (M + 1) + m] of the group's K subsets
  // (m = M: the constant term), then the polynomial when it is kept in LDS.
  // Coefficient a of subset k is poly[a * K + k]: the K subsets of a group
  // share each link and read K adjacent coefficients per gather.
  double2 *rs = smem;
  double2 *poly = kLds ? smem + K * per
                       : a.scratch + static_cast<int64_t>(wg) * a.n_mono * K;
  double2 *acc = a.acc + static_cast<int64_t>(wg) * a.n_out;
  const uint32_t n_subsets = (1u << n) - 1u;
  const uint32_t n_groups = (n_subsets + K - 1) / K;
  bool first = true;
  for (uint32_t g = static_cast<uint32_t>(wg); g < n_groups; g += static_cast<uint32_t>(a.n_wg)) {
    // the previous group's accumulation has read the polynomials
    __syncthreads();
    for (int t = tid; t < K * per; t += kThreads) {
      const int k = t / per, r = t - k * per;
      const int i = r / (M + 1), m = r - i * (M + 1);
      // a subset index past 2^n - 1 (the last group's tail) is computed like
      // the others and added with sign 0
      const uint32_t s = 1u + g * K + static_cast<uint32_t>(k);
      const double2 *B = (m == M) ? a.b0 : a.bm + static_cast<int64_t>(m) * n * n;
      double2 v = make_double2(0.0, 0.0);
      for (int j = 0; j < n; ++j)
        if (s >> j & 1u) {
          const double2 b = B[i * n + j];
          v.x += b.x;
          v.y += b.y;
        }
      rs[t] = v;
    }
    if (tid < K) poly[tid] = make_double2(1.0, 0.0);
    __syncthreads();
    for (int i = 0; i < n; ++i) {
      const int top = (i + 1 < L) ? i + 1 : L;
      for (int D = top; D >= 0; --D) {
        // degree i + 1 is new in this row: it has no old value to scale
        const bool fresh = (D == i + 1);
        const int lo = a.lvl_off[D], hi = a.lvl_off[D + 1];
        for (int t = tid; t < (hi - lo) * K; t += kThreads) {
          const int m_idx = lo + t / K, k = t % K;
          const double2 *r = rs + (k * n + i) * (M + 1);
          double2 v = fresh ? make_double2(0.0, 0.0) : cmul(r[M], poly[m_idx * K + k]);
          const int q1 = a.link_ptr[m_idx + 1];
          int q = a.link_ptr[m_idx];
          // four links at a time: their loads, then the four gathers, are
          // issued before the first FMA waits
          for (; q + 4 <= q1; q += 4) {
            const int e0 = a.links[q], e1 = a.links[q + 1], e2 = a.links[q + 2], e3 = a.links[q + 3];
            const double2 p0 = poly[(e0 >> 8) * K + k], p1 = poly[(e1 >> 8) * K + k];
            const double2 p2 = poly[(e2 >> 8) * K + k], p3 = poly[(e3 >> 8) * K + k];
            v = cfma(r[e0 & 0xff], p0, v);
            v = cfma(r[e1 & 0xff], p1, v);
            v = cfma(r[e2 & 0xff], p2, v);
            v = cfma(r[e3 & 0xff], p3, v);
          }
          for (; q < q1; ++q) {
            const int e = a.links[q];
            v = cfma(r[e & 0xff], poly[(e >> 8) * K + k], v);
          }
          poly[m_idx * K + k] = v;
        }
        __syncthreads();
      }
    }
    double sign[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t s = 1u + g * K + static_cast<uint32_t>(k);
      sign[k] = (s > n_subsets) ? 0.0 : (((n - __builtin_popcount(s)) & 1) ? -1.0 : 1.0);
    }
    for (int p = tid; p < a.n_out; p += kThreads) {
      const int64_t base = static_cast<int64_t>(a.out_mono[p]) * K;
      double2 v = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double2 c = poly[base + k];
        v.x += sign[k] * c.x;
        v.y += sign[k] * c.y;
      }
      if (first) {
        acc[p] = v;
      } else {
        double2 w = acc[p];
        w.x += v.x;
        w.y += v.y;
        acc[p] = w;
      }
    }
    first = false;
  }
}

__global__ __launch_bounds__(256) void perm_poly_reduce_kernel(PolyArgs a) {
  const int p = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (p >= a.n_out) return;
  dd re{0.0, 0.0}, im{0.0, 0.0};
  for (int w = 0; w < a.n_wg; ++w) {
    const double2 v = a.acc[static_cast<int64_t>(w) * a.n_out + p];
    re = dd_add_d(re, v.x);
    im = dd_add_d(im, v.y);
  }
  a.out[2 * p] = re.hi + re.lo;
  a.out[2 * p + 1] = im.hi + im.lo;
}

}  // namespace

hipError_t launch_perm_poly(const PolyArgs &a, bool lds, int group, size_t lds_bytes,
                            hipStream_t stream) {
  if (lds)
    hipLaunchKernelGGL((perm_poly_kernel<true, kPolyThreads, 1>), dim3(a.n_wg), dim3(kPolyThreads),
                       lds_bytes, stream, a);
  else if (group == kPolyGroup)
    hipLaunchKernelGGL((perm_poly_kernel<false, kPolyThreadsGlobal, kPolyGroup>), dim3(a.n_wg),
                       dim3(kPolyThreadsGlobal), lds_bytes, stream, a);
  else
    hipLaunchKernelGGL((perm_poly_kernel<false, kPolyThreadsGlobal, 1>), dim3(a.n_wg),
                       dim3(kPolyThreadsGlobal), lds_bytes, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(perm_poly_reduce_kernel, dim3((a.n_out + 255) / 256), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace pqp
