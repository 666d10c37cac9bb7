// Column-minor ("Laplace") kernels for gfx950: for one row multiplicity
// vector r and column multiplicities c with sum(c) = sum(r) + 1, every
//
//   M_l = perm(A; r, c - e_l) = 2^-(n-1) sum_g (-1)^{|g|} w(g)
//                                 prod_k cs_k(g)^{c_k - [k = l]},   n = sum(r),
//
// in one Gray sweep (Lemma 1 of arXiv:2005.04214; the reference's
// permanent_laplace_cpp, src/permanent_laplace.cpp:43-240, recomputes all
// m_c products per term, O(m_c^2); here prefix and suffix products make it
// O(m_c)).  The same launch also accumulates the uncut product
// prod_k cs_k^{c_k}, which is what the reference returns for a column with
// c_l = 0.
//
// The gradient of a permanent is built from these: d perm(A; r, c) / d A_ij
// = r_i c_j perm(A; r - e_i, c - e_j), i.e. one column-minor item per row
// with r_i > 0 (src/permanent.cpp:271-300 computes the same n^2 minors with
// one full permanent each).
//
// Lane organisation, start state and Gray tables are those of the forward
// kernels (perm_fwd.hip).  Per lane the row sums, the prefix products and
// the m_c + 1 FP64 accumulators live in VGPRs; lane sums are scaled by the
// lane's top weight and sign, reduced across the wave in double-double and
// stored per work unit; a finalize kernel sums each item's units in a fixed
// order (deterministic) and scales by 2^-(n-1).

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd.hpp"
#include "perm_kernels.hpp"
#include "pqp_types.hpp"
#include "sweep_common.hpp"

namespace pqp {

namespace {

template <int MCP, bool ONES>
__device__ __forceinline__ void minors_unit(const LaunchArgs &a, int item, int local) {
  using R = double;
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const int64_t U = it->U;
  const int S = uniform(static_cast<int>(it->S_in));
  const int L = it->L;
  const int Dt = it->Dt;
  const int flags = it->flags;
  constexpr bool ones = ONES;

  int64_t u = static_cast<int64_t>(local) * kWave + lane;
  const bool active = u < U;
  if (!active) u = U - 1;

  const cptr<R> Bin = as_const<R>(a.arena, it->b2_off);  // inner rows of B
  const cptr<InnerEntry> tab = as_const<InnerEntry>(a.arena, it->step_off);
  const cptr<int32_t> cmult = as_const<int32_t>(a.arena, it->cmul_off);

  R cr[MCP], ci[MCP];
  double w_top;
  start_state<MCP, R>(a, it, u, L, Dt, flags, cr, ci, w_top);
  double lane_w = (u & 1) ? -w_top : w_top;
  if (!active) lane_w = 0.0;

  double acr[MCP + 1], aci[MCP + 1];
#pragma unroll
  for (int k = 0; k <= MCP; ++k) {
    acr[k] = 0.0;
    aci[k] = 0.0;
  }

  constexpr int PF = MCP < 8 ? MCP : 8;  // row columns fetched ahead of the products
  InnerEntry e = inner_entry(tab, 0);
  for (int v = 0; v < S; ++v) {
    const cptr<R> row = Bin + static_cast<int64_t>(e.row) * (2 * MCP);
    R ur[PF], ui[PF];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      ur[j] = row[2 * j];
      ui[j] = row[2 * j + 1];
    }
    const InnerEntry en = inner_entry(tab, v + 1);
    const double cf = e.coef;

    if constexpr (ONES) {
      // unit multiplicities: E_k = cs_k, F_k = 1.  Prefix products
      // P_k = prod_{k' < k} cs_k' are kept for even k only (P_0 = 1); an odd
      // P_{2m+1} = P_{2m} cs_{2m} is rebuilt in the backward pass, which
      // carries the running suffix Q: minor l = P_l Q_l, and the final Q is
      // the uncut product.  (Half the prefix registers for m_c / 2 extra
      // multiplies keeps the kernel within one register budget step.)
      constexpr int NP = (MCP + 1) / 2;  // P_0, P_2, ..., P_{2(NP-1)}
      R pr[NP], pi[NP];
      pr[0] = R(1);
      pi[0] = R(0);
      if constexpr (NP > 1) {
        pr[1] = cr[0];
        pi[1] = ci[0];
        cmul(pr[1], pi[1], cr[1], ci[1]);
#pragma unroll
        for (int m = 2; m < NP; ++m) {
          pr[m] = pr[m - 1];
          pi[m] = pi[m - 1];
          cmul(pr[m], pi[m], cr[2 * m - 2], ci[2 * m - 2]);
          cmul(pr[m], pi[m], cr[2 * m - 1], ci[2 * m - 1]);
        }
      }
      R qr = R(1), qi = R(0);
#pragma unroll
      for (int k = MCP - 1; k >= 0; --k) {
        R mr, mi;
        if (k % 2 == 0) {
          mr = pr[k / 2];
          mi = pi[k / 2];
        } else {
          mr = pr[k / 2];
          mi = pi[k / 2];
          if (k > 1) {
            cmul(mr, mi, cr[k - 1], ci[k - 1]);
          } else {
            mr = cr[0];
            mi = ci[0];
          }
        }
        if (k < MCP - 1) {
          if (k == 0) {
            mr = qr;
            mi = qi;
          } else {
            cmul(mr, mi, qr, qi);
          }
        }
        acr[k] = __builtin_fma(cf, mr, acr[k]);
        aci[k] = __builtin_fma(cf, mi, aci[k]);
        if (k == MCP - 1) {
          qr = cr[k];
          qi = ci[k];
        } else {
          cmul(qr, qi, cr[k], ci[k]);
        }
      }
      acr[MCP] = __builtin_fma(cf, qr, acr[MCP]);
      aci[MCP] = __builtin_fma(cf, qi, aci[MCP]);
    } else {
      // F_k = cs_k^{c_k - 1} (1 for unit multiplicity), E_k = F_k cs_k, both
      // recomputed where needed (wave-uniform powers; no per-column arrays).
      // Prefix P_k = prod_{k' < k} E_k', suffix Q_k = prod_{k' > k} E_k';
      // minor product l: P_l F_l Q_l.
      auto F = [&](int k, R &fr_, R &fi_) {
        fr_ = cr[k];
        fi_ = ci[k];
        const int ck = cmult[k];
        for (int q = 2; q < ck; ++q) cmul(fr_, fi_, cr[k], ci[k]);
      };
      auto E = [&](int k, R &er, R &ei) {
        er = cr[k];
        ei = ci[k];
        const int ck = cmult[k];
        for (int q = 1; q < ck; ++q) cmul(er, ei, cr[k], ci[k]);
      };
      R pr[MCP], pi[MCP];
      pr[0] = R(1);
      pi[0] = R(0);
      if constexpr (MCP > 1) {
        E(0, pr[1], pi[1]);
#pragma unroll
        for (int k = 2; k < MCP; ++k) {
          R er, ei;
          E(k - 1, er, ei);
          pr[k] = pr[k - 1];
          pi[k] = pi[k - 1];
          cmul(pr[k], pi[k], er, ei);
        }
      }
      R qr = R(1), qi = R(0);
#pragma unroll
      for (int k = MCP - 1; k >= 0; --k) {
        R mr, mi;
        if (k == MCP - 1) {
          mr = pr[k];
          mi = pi[k];
        } else if (k == 0) {
          mr = qr;
          mi = qi;
        } else {
          mr = pr[k];
          mi = pi[k];
          cmul(mr, mi, qr, qi);
        }
        if (cmult[k] > 1) {
          R fr_, fi_;
          F(k, fr_, fi_);
          cmul(mr, mi, fr_, fi_);
        }
        acr[k] = __builtin_fma(cf, mr, acr[k]);
        aci[k] = __builtin_fma(cf, mi, aci[k]);
        R er, ei;
        E(k, er, ei);
        if (k == MCP - 1) {
          qr = er;
          qi = ei;
        } else {
          cmul(qr, qi, er, ei);
        }
      }
      // uncut product prod_k E_k = Q_{-1}
      acr[MCP] = __builtin_fma(cf, qr, acr[MCP]);
      aci[MCP] = __builtin_fma(cf, qi, aci[MCP]);
    }

    // row sums += (+-2) x inner row (after the last state the padding
    // entry adds 0 x row 0; never read)
    const R s2 = in_vgpr(static_cast<R>(e.sign2));
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      cr[j] = fmaR(s2, ur[j], cr[j]);
      ci[j] = fmaR(s2, ui[j], ci[j]);
    }
#pragma unroll
    for (int j = PF; j < MCP; ++j) {
      cr[j] = fmaR(s2, static_cast<R>(row[2 * j]), cr[j]);
      ci[j] = fmaR(s2, static_cast<R>(row[2 * j + 1]), ci[j]);
    }
    e = en;
  }

  double *p = a.partials + 4 * (it->partial_base + static_cast<int64_t>(local) * (MCP + 1));
#pragma unroll
  for (int k = 0; k <= MCP; ++k) {
    cdd s = cdd{{acr[k] * lane_w, 0.0}, {aci[k] * lane_w, 0.0}};
    // the products above are exact for integer lane weights < 2^53 only up
    // to rounding; carry the rounding error of the scaling into the low part
    s.re.lo = __builtin_fma(acr[k], lane_w, -s.re.hi);
    s.im.lo = __builtin_fma(aci[k], lane_w, -s.im.hi);
    s = wave_reduce(s);
    if (lane == 0) {
      p[4 * k] = s.re.hi;
      p[4 * k + 1] = s.re.lo;
      p[4 * k + 2] = s.im.hi;
      p[4 * k + 3] = s.im.lo;
    }
  }
}

template <int LO, int HI, bool ONES>
__device__ __forceinline__ void minors_dispatch(const LaunchArgs &a, int mcp, int item,
                                                int local) {
  if constexpr (LO <= HI) {
    if (mcp == LO) {
      minors_unit<LO, ONES>(a, item, local);
    } else {
      minors_dispatch<LO + 1, HI, ONES>(a, mcp, item, local);
    }
  }
}

// Waves per SIMD the register budget is sized for: unit-multiplicity
// kernels keep m_c row sums, m_c - 1 prefix products and m_c + 1
// accumulators (about 12 m_c VGPRs); the general kernels also recompute
// column powers.
constexpr int minor_waves(int hi, bool ones) {
  return hi <= 8 ? 4 : (ones && hi <= 20) ? 2 : 1;
}

}  // namespace

// One block per work unit.  ONES: every column multiplicity is 1 (the
// planner puts such items in the buckets 0-3, the others in 4-7), so the
// unit-multiplicity kernel's registers are not sized by the general path.
template <int LO, int HI, bool ONES>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(minor_waves(HI, ONES), 8))) void perm_minors_kernel(
    LaunchArgs a, int first_unit) {
  const cptr<WorkUnit> wu = as_const<WorkUnit>(a.units) + (first_unit + static_cast<int>(blockIdx.x));
  const int item = uniform(wu->item);
  const int local = uniform(wu->local);
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int mcp = uniform(it->mcp);
  minors_dispatch<LO, HI, ONES>(a, mcp, item, local);
}

#if PQP_BUCKET_ID == 0
// One block of NT threads per item.  With NT = 64 (items of at most 64
// units, as in batches) lane l sums column slot l's unit partials in unit
// order.  With NT = 256 (a few items of many units, as in single calls),
// for each slot thread t sums the partials of units t, t + NT, ... in
// double-double and the block combines the thread sums in a fixed order
// (butterfly within each wave, waves in index order).  Both orders are
// fixed, so results do not depend on timing.
template <int NT>
__global__ __launch_bounds__(NT) void perm_minors_finalize_kernel(LaunchArgs a) {
  constexpr int kWaves = NT / kWave;
  __shared__ cdd wave_sum[kWaves];
  const int item = static_cast<int>(blockIdx.x);
  const int tid = static_cast<int>(threadIdx.x);
  if (item >= a.n_items) return;
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int mc = it->mc;
  const int mcp = it->mcp;
  const int units = it->units;
  const int64_t base = it->partial_base;
  const int e = it->n > 0 ? -(it->n - 1) : 0;
  if constexpr (NT == kWave) {
    // few units per item (batches): lane l sums slot l's partials in unit order
    for (int l = tid; l <= mc; l += kWave) {
      const int slot = (l == mc) ? mcp : l;
      cdd t = cdd_zero();
      if (it->n == 0) {
        t.re.hi = 1.0;
      } else {
        for (int k = 0; k < units; ++k) {
          const double *p = a.partials + 4 * (base + static_cast<int64_t>(k) * (mcp + 1) + slot);
          t = cdd_add(t, cdd{{p[0], p[1]}, {p[2], p[3]}});
        }
      }
      const double re_hi = ldexp(t.re.hi, e), re_lo = ldexp(t.re.lo, e);
      const double im_hi = ldexp(t.im.hi, e), im_lo = ldexp(t.im.lo, e);
      const int64_t o = static_cast<int64_t>(it->out_index) + l;
      a.out[2 * o] = re_hi + re_lo;
      a.out[2 * o + 1] = im_hi + im_lo;
      if (a.out_dd) {
        a.out_dd[4 * o] = re_hi;
        a.out_dd[4 * o + 1] = re_lo;
        a.out_dd[4 * o + 2] = im_hi;
        a.out_dd[4 * o + 3] = im_lo;
      }
    }
    return;
  }
  for (int l = 0; l <= mc; ++l) {
    // kernel slot mcp holds the uncut product; slots mc..mcp-1 are padding
    const int slot = (l == mc) ? mcp : l;
    cdd s = cdd_zero();
    if (it->n == 0) {
      if (tid == 0) s.re.hi = 1.0;  // perm of an empty selection
    } else {
      for (int k = tid; k < units; k += NT) {
        const double *p = a.partials + 4 * (base + static_cast<int64_t>(k) * (mcp + 1) + slot);
        s = cdd_add(s, cdd{{p[0], p[1]}, {p[2], p[3]}});
      }
    }
    s = wave_reduce(s);
    if ((tid & (kWave - 1)) == 0) wave_sum[tid / kWave] = s;
    __syncthreads();
    if (tid == 0) {
      cdd t = wave_sum[0];
      for (int w = 1; w < kWaves; ++w) t = cdd_add(t, wave_sum[w]);
      const double re_hi = ldexp(t.re.hi, e), re_lo = ldexp(t.re.lo, e);
      const double im_hi = ldexp(t.im.hi, e), im_lo = ldexp(t.im.lo, e);
      const int64_t o = static_cast<int64_t>(it->out_index) + l;
      a.out[2 * o] = re_hi + re_lo;
      a.out[2 * o + 1] = im_hi + im_lo;
      if (a.out_dd) {
        a.out_dd[4 * o] = re_hi;
        a.out_dd[4 * o + 1] = re_lo;
        a.out_dd[4 * o + 2] = im_hi;
        a.out_dd[4 * o + 3] = im_lo;
      }
    }
    __syncthreads();  // wave_sum is rewritten for the next slot
  }
}

#endif

// ---------------------------------------------------------------------------
// Instantiation per bucket (-DPQP_BUCKET_ID=0..8): unit column
// multiplicities in 0-4 (widths 1-8, 9-16, 17-18, 19-20, 21-32; up to 20
// columns two waves per SIMD fit without spills), general ones in 5-8
// (1-8, 9-16, 17-24, 25-32).
// ---------------------------------------------------------------------------
#ifndef PQP_BUCKET_ID
#error "compile with -DPQP_BUCKET_ID=<0..8>"
#endif
#if PQP_BUCKET_ID == 0 || PQP_BUCKET_ID == 5
#define PQP_LO 1
#define PQP_HI 8
#elif PQP_BUCKET_ID == 1 || PQP_BUCKET_ID == 6
#define PQP_LO 9
#define PQP_HI 16
#elif PQP_BUCKET_ID == 2
#define PQP_LO 17
#define PQP_HI 18
#elif PQP_BUCKET_ID == 3
#define PQP_LO 19
#define PQP_HI 20
#elif PQP_BUCKET_ID == 4
#define PQP_LO 21
#define PQP_HI 32
#elif PQP_BUCKET_ID == 7
#define PQP_LO 17
#define PQP_HI 24
#else
#define PQP_LO 25
#define PQP_HI 32
#endif
#define PQP_ONES (PQP_BUCKET_ID < 5)

#define PQP_PASTE_(base, k) base##k
#define PQP_PASTE(base, k) PQP_PASTE_(base, k)

hipError_t PQP_PASTE(launch_minors_b, PQP_BUCKET_ID)(const LaunchArgs &a, int first_unit,
                                                      int n_units, hipStream_t stream) {
  if (n_units <= 0) return hipSuccess;
  hipLaunchKernelGGL((perm_minors_kernel<PQP_LO, PQP_HI, PQP_ONES>), dim3(n_units), dim3(kWave), 0,
                     stream, a, first_unit);
  return hipGetLastError();
}

#if PQP_BUCKET_ID == 0
hipError_t launch_minors_finalize(const LaunchArgs &a, int n_items, int max_units,
                                  hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  // a few large items (single calls) get 256 threads each, many small ones a wave
  if (max_units > kWave)
    hipLaunchKernelGGL(perm_minors_finalize_kernel<256>, dim3(n_items), dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL(perm_minors_finalize_kernel<kWave>, dim3(n_items), dim3(kWave), 0, stream, a);
  return hipGetLastError();
}
#endif

}  // namespace pqp
