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
// One workgroup per subset at a time (subsets strided over the grid).  The
// polynomial lives in LDS when it fits (kLds; 256 threads, several
// workgroups per CU), else in a per-workgroup global-memory slice (1024
// threads, about one workgroup per CU, so the slices and accumulators stay
// few).  Each workgroup adds its subsets' signed coefficients
// into its own FP64 accumulator; perm_poly_reduce_kernel then sums the
// workgroups' accumulators in double-double in workgroup order, so results
// do not depend on timing.
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

template <bool kLds, int kThreads>
__global__ __launch_bounds__(kThreads) void perm_poly_kernel(PolyArgs a) {
  extern __shared__ double2 smem[];
  const int n = a.n, M = a.M, L = a.L;
  const int tid = static_cast<int>(threadIdx.x);
  const int wg = static_cast<int>(blockIdx.x);
  // LDS: row sums rs[i * (M + 1) + m] (m = M: the constant term), then the
  // polynomial when it is kept in LDS
  double2 *rs = smem;
  double2 *poly = kLds ? smem + n * (M + 1)
                       : a.scratch + static_cast<int64_t>(wg) * a.n_mono;
  double2 *acc = a.acc + static_cast<int64_t>(wg) * a.n_out;
  const uint32_t n_subsets = (1u << n) - 1u;
  bool first = true;
  for (uint32_t s = 1u + static_cast<uint32_t>(wg); s <= n_subsets; s += static_cast<uint32_t>(a.n_wg)) {
    // the previous subset's accumulation has read the polynomial
    __syncthreads();
    for (int t = tid; t < n * (M + 1); t += kThreads) {
      const int i = t / (M + 1), m = t - i * (M + 1);
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
    if (tid == 0) poly[0] = make_double2(1.0, 0.0);
    __syncthreads();
    for (int i = 0; i < n; ++i) {
      const double2 *r = rs + i * (M + 1);
      const double2 c = r[M];
      const int top = (i + 1 < L) ? i + 1 : L;
      for (int D = top; D >= 0; --D) {
        // degree i + 1 is new in this row: it has no old value to scale
        const bool fresh = (D == i + 1);
        const int lo = a.lvl_off[D], hi = a.lvl_off[D + 1];
        for (int m_idx = lo + tid; m_idx < hi; m_idx += kThreads) {
          double2 v = fresh ? make_double2(0.0, 0.0) : cmul(c, poly[m_idx]);
          const int k1 = a.link_ptr[m_idx + 1];
          for (int k = a.link_ptr[m_idx]; k < k1; ++k) {
            const int e = a.links[k];
            v = cfma(r[e & 0xff], poly[e >> 8], v);
          }
          poly[m_idx] = v;
        }
        __syncthreads();
      }
    }
    const double sign = ((n - __builtin_popcount(s)) & 1) ? -1.0 : 1.0;
    for (int p = tid; p < a.n_out; p += kThreads) {
      const double2 v = poly[a.out_mono[p]];
      if (first) {
        acc[p] = make_double2(sign * v.x, sign * v.y);
      } else {
        double2 w = acc[p];
        w.x += sign * v.x;
        w.y += sign * v.y;
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

hipError_t launch_perm_poly(const PolyArgs &a, bool lds, size_t lds_bytes, hipStream_t stream) {
  if (lds)
    hipLaunchKernelGGL((perm_poly_kernel<true, kPolyThreads>), dim3(a.n_wg), dim3(kPolyThreads),
                       lds_bytes, stream, a);
  else
    hipLaunchKernelGGL((perm_poly_kernel<false, kPolyThreadsGlobal>), dim3(a.n_wg),
                       dim3(kPolyThreadsGlobal), lds_bytes, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(perm_poly_reduce_kernel, dim3((a.n_out + 255) / 256), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace pqp
