// Tensor permanent of uniformly partially distinguishable bosons on gfx950,
// the quantity behind piquasso's detection probability of such inputs
// (piquasso/_simulators/passive/probabilities.py:156-237; Tichy,
// arXiv:1410.7687 Sec. III).  The reference sums over pairs of column
// subsets (s, t); here the same value is summed over pairs of sign vectors
// (delta, eps) in {+1, -1}^n with delta_0 = eps_0 = +1, the form Glynn's
// formula gives the permanent:
//
//   T = 4^-(n-1) sum_{delta, eps} (prod delta)(prod eps)
//       prod_i ( x R_delta[i] conj(R_eps[i]) + (1 - x) Q_{delta eps}[i] ),
//   R_delta[i] = sum_c delta_c A[i, c],
//   Q_{delta eps}[i] = sum_c delta_c eps_c |A[i, c]|^2.
//
// Averaging prod delta prod_i delta_{sigma(i)} over the signs keeps exactly
// the maps sigma that are permutations, as for the subsets, so both sums
// expand to sum_{sigma, tau} prod_i A[i, sigma(i)] conj(A[i, tau(i)])
// x^#{i: sigma(i) != tau(i)}.  The sign form has 4^(n-1) pairs instead of
// 4^n and far less cancellation (sum |term| / |T| ~ 10 instead of ~1e5 at
// n = 8).
//
// One wave per work unit (tensor_types.hpp: delta's bits as s, eps's as t).
// Each lane owns one delta and keeps x R_delta (2n doubles) in registers;
// the wave walks 2^kt eps in Gray order, so a step flips one sign eps_c:
// R_eps -+= 2 A[:, c] (wave-uniform, A's column read with scalar loads) and
// Q -+= 2 delta_c (1 - x)|A[:, c]|^2.  The term costs 8 VALU per row; only
// its real part is accumulated (the (delta, eps) and (eps, delta) terms are
// conjugate).  FP64 runs of 64 steps are folded into a double-double lane
// sum; the wave and then the item's units are reduced in double-double in a
// fixed order.  Each unit also sums |term| (times its weight) for a
// forward-error bound of the result.
//
// Compiled twice: -DPQP_TENSOR_PART=0 instantiates n = 1..12 and the
// finalize kernel, =1 instantiates n = 13..20.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd.hpp"
#include "sweep_common.hpp"
#include "tensor_types.hpp"

namespace pqp {

namespace {

__device__ __forceinline__ dd wave_reduce_dd(dd v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    dd o;
    o.hi = __shfl_xor(v.hi, off);
    o.lo = __shfl_xor(v.lo, off);
    v = dd_add(v, o);
  }
  return v;
}

template <int N>
__global__ __launch_bounds__(kWave) void tensor_perm_kernel(TensorArgs a, int first_unit) {
  const int u_idx = first_unit + static_cast<int>(blockIdx.x);
  const cptr<TensorUnit> up = as_const<TensorUnit>(a.units) + u_idx;
  const int item = up->item;
  const uint32_t s_hi = static_cast<uint32_t>(up->s_hi);
  const uint32_t t_blk = static_cast<uint32_t>(up->t_blk);
  const int mixed = up->mixed;
  const cptr<TensorItem> it = as_const<TensorItem>(a.items) + item;
  const int ls = it->ls;
  const int kt = it->kt;
  const double x = it->x;
  const cptr<double> col = as_const<double>(a.arena, it->cols_off);  // [c][i] (re, im)
  const cptr<double> pw = as_const<double>(a.arena, it->pw_off);     // [c][i]

  const int lane = static_cast<int>(threadIdx.x);
  const bool active = lane < (1 << ls);
  // sign vectors: column 0 fixed at +1, column b + 1 is -1 where bit b is set
  const uint32_t s = (s_hi << ls) | static_cast<uint32_t>(active ? lane : 0);
  const uint32_t t0 = t_blk << kt;

  // x R_delta, R_eps at the walk's first eps, (1 - x) Q_{delta eps}
  double xr[N], xi[N], rr[N], ri[N], q[N];
#pragma unroll
  for (int i = 0; i < N; ++i) xr[i] = xi[i] = rr[i] = ri[i] = q[i] = 0.0;
#pragma unroll 1
  for (int c = 0; c < N; ++c) {
    const bool dneg = c > 0 && ((s >> (c - 1)) & 1u);
    const bool eneg = c > 0 && ((t0 >> (c - 1)) & 1u);  // wave-uniform
    const double dx = dneg ? -x : x;
    const double es = in_vgpr(eneg ? -1.0 : 1.0);
    const double de = (dneg != eneg) ? -1.0 : 1.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      xr[i] = fmaR(dx, col[2 * (c * N + i)], xr[i]);
      xi[i] = fmaR(dx, col[2 * (c * N + i) + 1], xi[i]);
      rr[i] = fmaR(es, col[2 * (c * N + i)], rr[i]);
      ri[i] = fmaR(es, col[2 * (c * N + i) + 1], ri[i]);
      q[i] = fmaR(de, pw[c * N + i], q[i]);
    }
  }
  const double lane_sign = (__builtin_popcount(s) & 1) ? -1.0 : 1.0;
  int par_t = __builtin_popcount(t0) & 1;

  dd acc_dd{0.0, 0.0};
  double acc = 0.0;
  double aabs = 0.0;  // sum of |weight x term| (forward-error scale)
  const int steps = 1 << kt;
  // The step after state k flips eps at column c = ctz(k + 1) + 1:
  // R_eps -+= 2 A[:, c], Q -+= 2 delta_c (1 - x)|A[:, c]|^2.  A step's
  // column is 3n doubles of scalar operands, more than the SGPR file holds
  // at once for large n: the first PF rows of A's column are loaded before
  // the product of state k (their latency hides behind it), the rest after.
  constexpr int PF = N < 8 ? N : 8;
  for (int k = 0; k < steps; ++k) {
    const bool more = k + 1 < steps;  // wave-uniform
    const int b = __builtin_ctz(static_cast<unsigned>(k + 1));
    const int c = more ? b + 1 : 0;
    double ur[PF], ui[PF];
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      ur[i] = col[2 * (c * N + i)];
      ui[i] = col[2 * (c * N + i) + 1];
    }
    double w = 2.0;
    if (mixed) {
      const uint32_t tb = (t0 | static_cast<uint32_t>(k ^ (k >> 1))) >> ls;
      w = tb > s_hi ? 2.0 : (tb == s_hi ? 1.0 : 0.0);
    }
    if (w != 0.0) {
      // factor i: x R_delta[i] conj(R_eps[i]) + Q[i]
      double pr = fmaR(xi[0], ri[0], fmaR(xr[0], rr[0], q[0]));
      double pi = fmaR(xi[0], rr[0], -(xr[0] * ri[0]));
#pragma unroll
      for (int i = 1; i < N - 1; ++i) {
        const double fr = fmaR(xi[i], ri[i], fmaR(xr[i], rr[i], q[i]));
        const double fi = fmaR(xi[i], rr[i], -(xr[i] * ri[i]));
        cmul(pr, pi, fr, fi);
      }
      if (N > 1) {
        const double fr = fmaR(xi[N - 1], ri[N - 1], fmaR(xr[N - 1], rr[N - 1], q[N - 1]));
        const double fi = fmaR(xi[N - 1], rr[N - 1], -(xr[N - 1] * ri[N - 1]));
        pr = fmaR(pr, fr, -(pi * fi));  // real part of the product
      }
      const double coef = (par_t ? -w : w) * lane_sign;
      acc = fmaR(coef, pr, acc);
      aabs = fmaR(w, __builtin_fabs(pr), aabs);
    }
    if ((k & (kDdFold - 1)) == kDdFold - 1) {
      acc_dd = dd_add_d(acc_dd, acc);
      acc = 0.0;
    }
    if (more) {
      const uint32_t g = static_cast<uint32_t>((k + 1) ^ ((k + 1) >> 1));
      const double sd2 = in_vgpr(((g >> b) & 1u) ? -2.0 : 2.0);
      const double sq2 = ((s >> b) & 1u) ? -sd2 : sd2;
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        rr[i] = fmaR(sd2, ur[i], rr[i]);
        ri[i] = fmaR(sd2, ui[i], ri[i]);
      }
#pragma unroll
      for (int i = PF; i < N; ++i) {
        rr[i] = fmaR(sd2, col[2 * (c * N + i)], rr[i]);
        ri[i] = fmaR(sd2, col[2 * (c * N + i) + 1], ri[i]);
      }
#pragma unroll
      for (int i = 0; i < N; ++i) q[i] = fmaR(sq2, pw[c * N + i], q[i]);
      par_t ^= 1;
    }
  }
  acc_dd = dd_add_d(acc_dd, acc);
  if (!active) {
    acc_dd = dd{0.0, 0.0};
    aabs = 0.0;
  }
  acc_dd = wave_reduce_dd(acc_dd);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) aabs += __shfl_xor(aabs, off);
  if (lane == 0) {
    a.partials[3 * u_idx] = acc_dd.hi;
    a.partials[3 * u_idx + 1] = acc_dd.lo;
    a.partials[3 * u_idx + 2] = aabs;
  }
}

#if PQP_TENSOR_PART == 0
// One wave per item: the dd sum of the item's unit partials in unit order
// (lane-strided, then a fixed butterfly), T = hi + lo, and the sum of the
// units' |term| sums.
__global__ __launch_bounds__(kWave) void tensor_finalize_kernel(TensorArgs a) {
  const int item = static_cast<int>(blockIdx.x);
  if (item >= a.n_items) return;
  const cptr<TensorItem> it = as_const<TensorItem>(a.items) + item;
  const int lane = static_cast<int>(threadIdx.x);
  dd s{0.0, 0.0};
  double sa = 0.0;
  if (it->n == 0) {
    if (lane == 0) s.hi = sa = 1.0;  // the empty product (probabilities.py:197-198)
  } else {
    for (int u = lane; u < it->n_units; u += kWave) {
      const double *p = a.partials + 3 * (static_cast<int64_t>(it->unit0) + u);
      s = dd_add(s, dd{p[0], p[1]});
      sa += p[2];
    }
  }
  s = wave_reduce_dd(s);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) sa += __shfl_xor(sa, off);
  if (lane == 0) {
    // Glynn scale 4^-(n-1) (exact)
    const int e = it->n > 0 ? -2 * (it->n - 1) : 0;
    a.out[2 * item] = ldexp(s.hi + s.lo, e);
    a.out[2 * item + 1] = ldexp(sa, e);
  }
}
#endif

}  // namespace

#if PQP_TENSOR_PART == 0
#define PQP_TENSOR_CASES X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
hipError_t launch_tensor_p0(int n, const TensorArgs &a, int first_unit, int n_units,
                            hipStream_t stream) {
#else
#define PQP_TENSOR_CASES X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)
hipError_t launch_tensor_p1(int n, const TensorArgs &a, int first_unit, int n_units,
                            hipStream_t stream) {
#endif
  if (n_units <= 0) return hipSuccess;
  switch (n) {
#define X(k)                                                                              \
  case k:                                                                                 \
    hipLaunchKernelGGL(tensor_perm_kernel<k>, dim3(n_units), dim3(kWave), 0, stream, a,  \
                       first_unit);                                                      \
    return hipGetLastError();
    PQP_TENSOR_CASES
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

#if PQP_TENSOR_PART == 0
hipError_t launch_tensor_finalize(const TensorArgs &a, hipStream_t stream) {
  if (a.n_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(tensor_finalize_kernel, dim3(a.n_items), dim3(kWave), 0, stream, a);
  return hipGetLastError();
}
#endif

}  // namespace pqp
