// Gradient kernels for gfx950: d perm(A; r, c) / d A for every entry from
// ONE Glynn sweep over the reflected mixed-radix Gray code (the sweep of
// perm_fwd.hip), together with the permanent itself.
//
// With the reference's prepend (src/permanent.cpp:54-95) the permanent is
//
//   perm = 2^-(n-1) sum_t s(t) prod_j cs_j(t)^{c_j},   cs_j = sum_i alpha_i(t) b_ij,
//
// alpha_0 = 1 for the prepended row and alpha_i = rho_i - 2 g_i(t) for the
// others, s(t) = (-1)^{|g|} prod_i C(rho_i, g_i).  Its derivative by b_ij is
//
//   G_ij = 2^-(n-1) sum_t alpha_i(t) Y_j(t),   Y_j(t) = s(t) prod_{k != j} cs_k,
//
// for unit column multiplicities (the planner expands a column of
// multiplicity c into c columns): an n x m_c contraction of the per-term
// coefficient vector alpha(t) with the per-term exclusion products Y(t)
// (prefix and suffix products give all of Y in O(m_c)).  The gradient by A
// follows on the host: A's row i0 (the prepended copy) collects rows 0 and
// i0 of B, and the copies of a column add up.
// The reference's grad_perm (src/permanent.cpp:271-300) computes the same
// G_ij = r_i c_j perm(A; r - e_i, c - e_j) with one full permanent per entry.
//
// How the contraction runs on a wave (lanes = 64 top states, every lane
// sweeping the same inner states, as in the forward kernels):
//
//  * top rows: alpha_i is constant over a lane's sweep, so G_i = sum_u
//    alpha_i(u) W_u with W_u = sum_t Y(u, t), the lane's running sum of
//    exclusion products (m_c complex accumulators per lane).  At the end of
//    the unit each top row's alpha-weighted W is summed over the lanes.
//  * inner rows: alpha_i(t) is the same in every lane, so G_i = sum_t
//    alpha_i(t) Yw(t) with Yw(t) = sum_u Y(u, t), the wave-wide sum of the
//    term's exclusion products.  Every step reduce-scatters Y over the 64
//    lanes (two permlane32/16 swap stages, then an exchange through LDS
//    within each group of 16 lanes): each lane ends up owning one of the
//    2 m_c real components of Yw(t).  Only one inner digit moves per step, so the lane
//    keeps the running sum V(t) of its component and, per inner row, an
//    accumulator K_i in LDS: when row i's coefficient moves by d after term
//    t, K_i -= d V(t), and at the end G_i = K_i + alpha_i(final) V(end)
//    (the segments of constant alpha_i, telescoped).  Per step that is one
//    LDS read-modify-write per lane, whatever the number of inner rows.
//  * the prepended row: alpha_0 = 1, so its gradient is V(end) itself.
//
// Per unit, every (row of B, real component) value is written as one FP64
// partial; a finalize kernel sums an item's units in double-double in unit
// order (deterministic), adds the permanent's dd partials, and scales by
// 2^-(n-1).  No MFMA: the per-term work is a product of sums; the
// contraction over terms is folded into running sums.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd.hpp"
#include "perm_kernels.hpp"
#include "pqp_types.hpp"
#include "sweep_common.hpp"

#ifndef PQP_GRAD_PF
#define PQP_GRAD_PF 12
#endif

namespace pqp {

namespace {

__device__ __forceinline__ void split64(double x, uint32_t &lo, uint32_t &hi) {
  const uint64_t v = __builtin_bit_cast(uint64_t, x);
  lo = static_cast<uint32_t>(v);
  hi = static_cast<uint32_t>(v >> 32);
}

__device__ __forceinline__ double join64(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// One reduce-scatter step over lane pairs (l, l ^ (1 << B)), B = 5 or 4:
// lanes with bit B clear return a + a(partner), lanes with it set return
// b + b(partner).  v_permlane32_swap / v_permlane16_swap exchange the two
// values between the partners in one instruction per dword.
template <int B>
__device__ __forceinline__ double xadd(double a, double b) {
  static_assert(B == 5 || B == 4, "permlane swaps cover lane bits 5 and 4");
  uint32_t al, ah, bl, bh;
  split64(a, al, ah);
  split64(b, bl, bh);
  // swap(x, y): first result = x in the lower partner, y of the lower
  // partner in the upper one; second = x of the upper partner in the lower
  // one, y in the upper one (lower = bit B clear)
  auto lo = B == 5 ? __builtin_amdgcn_permlane32_swap(al, bl, false, false)
                   : __builtin_amdgcn_permlane16_swap(al, bl, false, false);
  auto hi = B == 5 ? __builtin_amdgcn_permlane32_swap(ah, bh, false, false)
                   : __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
  return join64(lo[0], hi[0]) + join64(lo[1], hi[1]);
}

// Reduce-scatter of an M-column complex vector over the wave, laid out as
// [re_0..re_{M-1}, im_0..im_{M-1}]:
//   stage 1 (lane bit 5) pairs re_k with im_k: the lower half of the wave
//     keeps real parts, the upper half imaginary parts (M values per lane);
//   stage 2 (lane bit 4) pairs columns 2c and 2c + 1 (the missing partner of
//     an odd M is zero): the lower 16 lanes of each half keep the even
//     columns, the upper 16 the odd ones; H = ceil(M / 2) values per lane,
//     each a sum over the lanes q, q ^ 16, q ^ 32, q ^ 48;
//   the 16 lanes of each group g = q >> 4 then hold partial sums of the same
//     H values; they exchange them through LDS (row c of a per-wave table,
//     one column per lane) and lane 16 g + c (c < H) adds the 16 partials of
//     value c in a fixed order.
// Lane q ends up owning real component (q & 32 ? M : 0) + 2 (q & 15) +
// (q & 16 ? 1 : 0) when that column is below M.
constexpr int kTStride = kWave + 2;  // LDS row stride in doubles: rows fall on different banks

template <int M>
__device__ __forceinline__ int rs_component(int lane, bool &valid) {
  const int col = 2 * (lane & 15) + ((lane & 16) ? 1 : 0);
  valid = col < M;
  return ((lane & 32) ? M : 0) + col;
}

// The exchange in two halves: rs_write stores the lane's H partials in its
// column of the table, rs_read adds up the 16 partials of the lane's value.
// Between the two, the table must not be written again (the sweep keeps
// two tables and reads one step late, so the store's latency is hidden).
template <int M>
__device__ __forceinline__ void rs_write(const double (&s2)[(M + 1) / 2], double *T, int lane) {
  constexpr int H = (M + 1) / 2;
  static_assert(H <= 16, "a 16-lane group holds at most 16 values");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int c = 0; c < H; ++c) T[c * kTStride + lane] = s2[c];
  __builtin_amdgcn_wave_barrier();
}

template <int M>
__device__ __forceinline__ double rs_read(const double *T, int lane) {
  constexpr int H = (M + 1) / 2;
  const int c = lane & 15;
  const double *row = T + (c < H ? c : 0) * kTStride + (lane & ~15);
  // two interleaved chains (even and odd lanes of the group), fixed order
  double s0 = row[0], s1 = row[1];
#pragma unroll
  for (int i = 2; i < 16; i += 2) {
    s0 += row[i];
    s1 += row[i + 1];
  }
  __builtin_amdgcn_wave_barrier();
  return c < H ? s0 + s1 : 0.0;
}

template <int M>
__device__ __forceinline__ double rs_finish(const double (&s2)[(M + 1) / 2], double *T, int lane) {
  rs_write<M>(s2, T, lane);
  return rs_read<M>(T, lane);
}

// Stages 1 and 2 fed in descending column order (the backward pass over
// the columns): column 2c + 1 waits in a register until column 2c arrives.
template <int M>
struct RsStream {
  static constexpr int H = (M + 1) / 2;
  double held;
  double s2[H];
  template <int K>
  __device__ __forceinline__ void put(double v) {
    if constexpr (K % 2 == 1) {
      held = v;
    } else {
      s2[K / 2] = xadd<4>(v, (K + 1 < M) ? held : 0.0);
    }
  }
};

// rs.put<k>(v) for a k known only after unrolling (k < M).
template <int M, int K = M - 1>
__device__ __forceinline__ void rs_put(RsStream<M> &rs, int k, double v) {
  if constexpr (K >= 0) {
    if (k == K) {
      rs.template put<K>(v);
    } else {
      rs_put<M, K - 1>(rs, k, v);
    }
  }
}

// Reduce-scatter of an M-column complex vector (wave sum, one real
// component per lane).
template <int M>
__device__ __forceinline__ double rs_vector(const double (&vr)[M], const double (&vi)[M], double *T,
                                            int lane) {
  RsStream<M> rs;
#pragma unroll
  for (int k = M - 1; k >= 0; --k) rs_put<M>(rs, k, xadd<5>(vr[k], vi[k]));
  return rs_finish<M>(rs.s2, T, lane);
}

// One wave unit of a gradient item: top states [64 local, 64 local + 64).
template <int MCP>
__device__ __forceinline__ void grad_unit(const LaunchArgs &a, int item, int local,
                                          double (*K)[kWave], double *T) {
  using R = double;
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const int64_t U = it->U;
  const int S = uniform(static_cast<int>(it->S_in));
  const int L = uniform(it->L);
  const int Dt = uniform(it->Dt);
  const int flags = uniform(it->flags);

  int64_t u = static_cast<int64_t>(local) * kWave + lane;
  const bool active = u < U;
  if (!active) u = U - 1;

  const cptr<R> Bin = as_const<R>(a.arena, it->b2_off);  // inner rows of B
  const cptr<InnerEntry> tab = as_const<InnerEntry>(a.arena, it->step_off);
  const cptr<double> alpha = as_const<double>(a.arena, it->alpha_off);  // start, then final

  R cr[MCP], ci[MCP];
  double w_top;
  start_state<MCP, R>(a, it, u, L, Dt, flags, cr, ci, w_top);
  double lane_w = (u & 1) ? -w_top : w_top;
  if (!active) lane_w = 0.0;

  for (int i = 0; i < (L > 0 ? L : 1); ++i) K[i][lane] = 0.0;
  double Wr[MCP], Wi[MCP];  // running sums of the lane's exclusion products
#pragma unroll
  for (int k = 0; k < MCP; ++k) {
    Wr[k] = 0.0;
    Wi[k] = 0.0;
  }
  double accr = 0.0, acci = 0.0;  // the lane's share of the permanent
  double vw = 0.0;                // running wave sum of Y, this lane's component
  // The exchange of step t is read back at the top of step t + 1 (two
  // tables, T and T + TB), together with the accumulator of step t's row
  // move: all of these LDS reads are in flight at once, while the registers
  // of the products are still free.
  constexpr int TB = ((MCP + 1) / 2) * kTStride;
  {
    double zero[(MCP + 1) / 2];
#pragma unroll
    for (int c = 0; c < (MCP + 1) / 2; ++c) zero[c] = 0.0;
    rs_write<MCP>(zero, T + TB, lane);  // "step -1" contributes nothing
  }
  int prow = 0;          // row moved after the previous step
  double ps2 = 0.0;      // its coefficient change (0: no previous step)

  // Row columns fetched at the top of the step (scalar loads), the rest
  // after the prefix products: the scheduling barriers keep the loads there,
  // so their latency hides behind the products instead of stalling the row
  // update at the end of the step.
  constexpr int PF = MCP < PQP_GRAD_PF ? MCP : PQP_GRAD_PF;
  InnerEntry e = inner_entry(tab, 0);
  for (int v = 0; v < S; ++v) {
    {
      // the previous step's exchange and its row move (ps2 after its term)
      const double kold = K[prow][lane];
      vw += rs_read<MCP>(T + ((v & 1) ? 0 : TB), lane);
      K[prow][lane] = __builtin_fma(-ps2, vw, kold);
    }
    const cptr<R> row = Bin + static_cast<int64_t>(e.row) * (2 * MCP);
    R ur[PF], ui[PF];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      ur[j] = row[2 * j];
      ui[j] = row[2 * j + 1];
    }
    const InnerEntry en = inner_entry(tab, v + 1);
    __builtin_amdgcn_sched_barrier(0);
    const double cf = e.coef * lane_w;  // exact: integer weights below 2^53

    // Exclusion products X_k = cf * prod_{k' != k} cs_k': prefix
    // P_k = cf * prod_{k' < k} cs_k' (kept for even k, odd ones rebuilt),
    // suffix Q_k = prod_{k' > k} cs_k' carried backwards.  Stage 1 of the
    // reduce-scatter pairs re_k with im_k as soon as X_k exists.
    RsStream<MCP> rs;  // stage 2, fed by stage 1 in descending column order
    constexpr int NP = (MCP + 1) / 2;  // P_0, P_2, ..., P_{2(NP-1)}
    R pr[NP], pi[NP];
    auto E = [&](int k, R &er, R &ei) {
      er = cr[k];
      ei = ci[k];
    };
    pr[0] = cf;
    pi[0] = R(0);
#pragma unroll
    for (int m = 1; m < NP; ++m) {
      R e0r, e0i, e1r, e1i;
      E(2 * m - 2, e0r, e0i);
      E(2 * m - 1, e1r, e1i);
      pr[m] = pr[m - 1];
      pi[m] = pi[m - 1];
      if (m == 1) {
        pr[m] = cf * e0r;
        pi[m] = cf * e0i;
      } else {
        cmul(pr[m], pi[m], e0r, e0i);
      }
      cmul(pr[m], pi[m], e1r, e1i);
    }
    R ur2[MCP - PF > 0 ? MCP - PF : 1], ui2[MCP - PF > 0 ? MCP - PF : 1];
#pragma unroll
    for (int j = PF; j < MCP; ++j) {
      ur2[j - PF] = row[2 * j];
      ui2[j - PF] = row[2 * j + 1];
    }
    if constexpr (MCP > PF) __builtin_amdgcn_sched_barrier(0);
    R qr = R(1), qi = R(0);
#pragma unroll
    for (int k = MCP - 1; k >= 0; --k) {
      R xr, xi;
      if (k % 2 == 0) {
        xr = pr[k / 2];
        xi = pi[k / 2];
      } else {
        R er, ei;
        E(k - 1, er, ei);
        xr = pr[k / 2];
        xi = pi[k / 2];
        if (k == 1) {
          xr = cf * er;
          xi = cf * ei;
        } else {
          cmul(xr, xi, er, ei);
        }
      }
      if (k < MCP - 1) cmul(xr, xi, qr, qi);
      Wr[k] += xr;
      Wi[k] += xi;
      rs_put<MCP>(rs, k, xadd<5>(xr, xi));
      R er, ei;
      E(k, er, ei);
      if (k == MCP - 1) {
        qr = er;
        qi = ei;
      } else {
        cmul(qr, qi, er, ei);
      }
    }
    // the term itself: cf * prod_k cs_k
    accr = __builtin_fma(cf, qr, accr);
    acci = __builtin_fma(cf, qi, acci);
    // this step's exchange, read back at the top of the next step
    rs_write<MCP>(rs.s2, T + ((v & 1) ? TB : 0), lane);
    const double s2v = in_vgpr(static_cast<double>(e.sign2));
    prow = static_cast<int>(e.row);
    ps2 = static_cast<double>(e.sign2);

    // row sums += (+-2) x inner row (after the last state the padding entry
    // adds 0 x row 0; never read)
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      cr[j] = fmaR(s2v, ur[j], cr[j]);
      ci[j] = fmaR(s2v, ui[j], ci[j]);
    }
#pragma unroll
    for (int j = PF; j < MCP; ++j) {
      cr[j] = fmaR(s2v, ur2[j - PF], cr[j]);
      ci[j] = fmaR(s2v, ui2[j - PF], ci[j]);
    }
    e = en;
  }

  // after the sweep the lane index is computed afresh (v_mbcnt: every lane
  // is active here) rather than kept live through the sweep
  int lane_e = static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
  __asm__ volatile("" : "+v"(lane_e));
  // the last step's exchange (its row move is the padding entry's: none)
  vw += rs_read<MCP>(T + ((S & 1) ? 0 : TB), lane_e);

  // ---- unit results: [perm dd (4)] [rows of B x 2 MCP components] ----
  const int rows_B = 1 + L + Dt;
  const int64_t stride = grad_unit_stride(rows_B, MCP);
  double *p = a.partials + 4 * it->partial_base + stride * local;
  cdd t = cdd{{accr, 0.0}, {acci, 0.0}};
  t = wave_reduce(t);
  if (lane_e == 0) {
    p[0] = t.re.hi;
    p[1] = t.re.lo;
    p[2] = t.im.hi;
    p[3] = t.im.lo;
  }
  double *g = p + 4;
  bool valid;
  const int comp = rs_component<MCP>(lane_e, valid);
  // prepended row (alpha = 1)
  if (valid) g[comp] = vw;
  // inner rows: K_i + alpha_i(final) V(end)
  for (int i = 0; i < L; ++i)
    if (valid) g[static_cast<int64_t>(1 + i) * 2 * MCP + comp] = __builtin_fma(alpha[L + i], vw, K[i][lane_e]);
  // top rows: sum over lanes of alpha_i(u) W_u, digits decoded as in
  // start_state (the lane's top state recomputed here rather than kept
  // live through the sweep)
  const cptr<double> top = as_const<double>(a.arena, it->top_off);
  int64_t ut = static_cast<int64_t>(local) * kWave + lane_e;
  if (ut >= U) ut = U - 1;
  const uint64_t gbin = static_cast<uint64_t>(ut) ^ (static_cast<uint64_t>(ut) >> 1);
  const double ud = static_cast<double>(ut) + 0.5;
  double q_hi = 0.0;
  int parity = 0;
  for (int tt = Dt - 1; tt >= 0; --tt) {
    double al;
    if (flags & kTopBinary) {
      al = ((gbin >> tt) & 1) ? -1.0 : 1.0;
    } else {
      const double inv_place = top[4 * tt];
      const double radix = top[4 * tt + 1];
      const int rho = static_cast<int>(top[4 * tt + 2]);
      const double q = __builtin_floor(ud * inv_place);
      const int c = static_cast<int>(__builtin_fma(-radix, q_hi, q));
      q_hi = q;
      const int gg = parity ? static_cast<int>(radix) - 1 - c : c;
      parity ^= (gg & 1);
      al = static_cast<double>(rho - 2 * gg);
    }
    double xr[MCP], xi[MCP];
#pragma unroll
    for (int k = 0; k < MCP; ++k) {
      xr[k] = al * Wr[k];
      xi[k] = al * Wi[k];
    }
    const double y = rs_vector<MCP>(xr, xi, T, lane_e);
    if (valid) g[static_cast<int64_t>(1 + L + tt) * 2 * MCP + comp] = y;
  }
}

template <int LO, int HI>
__device__ __forceinline__ void grad_dispatch(const LaunchArgs &a, int mcp, int item, int local,
                                              double (*K)[kWave], double *T) {
  if constexpr (LO <= HI) {
    if (mcp == LO) {
      grad_unit<LO>(a, item, local, K, T);
    } else {
      grad_dispatch<LO + 1, HI>(a, mcp, item, local, K, T);
    }
  }
}

// Waves per SIMD the register budget is sized for: up to 20 columns the row
// sums, the lane's running sums of exclusion products and the prefix
// products fit two waves' 256 VGPRs.
constexpr int grad_waves(int hi) { return hi <= 8 ? 4 : hi <= 20 ? 2 : 1; }

}  // namespace

// One block (one wave) per work unit.
template <int LO, int HI>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(grad_waves(HI), 8))) void perm_grad_kernel(LaunchArgs a, int first_unit) {
  __shared__ double K[kGradMaxInner][kWave];
  __shared__ __align__(16) double T[2 * ((HI + 1) / 2) * kTStride];  // reduce-scatter exchange tables
  const cptr<WorkUnit> wu = as_const<WorkUnit>(a.units) + (first_unit + static_cast<int>(blockIdx.x));
  const int item = uniform(wu->item);
  const int local = uniform(wu->local);
  const int mcp = uniform(as_const<ItemDesc>(a.items)[item].mcp);
  grad_dispatch<LO, HI>(a, mcp, item, local, K, T);
}

#if PQP_BUCKET_ID == 0
// Finalize of gradient items.  Output slot out_index holds the permanent,
// then row b of B, column j at out_index + 1 + b * mc + j: value v = 0 is
// the permanent, v = 1 + b * mc + j a gradient entry.  Block (item, c) of
// 256 threads takes values [64 c, 64 c + 64) of one item (blocks past the
// item's values return at once): wave w sums, for its lane's value, the
// partials of units w, w + 4, ... in double-double, and the four wave sums
// are added in wave order, so results do not depend on scheduling.
constexpr int kGradFinWaves = 4;

__global__ __launch_bounds__(kGradFinWaves * kWave) void perm_grad_finalize_kernel(LaunchArgs a) {
  __shared__ cdd part[kGradFinWaves][kWave];
  const int item = static_cast<int>(blockIdx.x);
  if (item >= a.n_items) return;
  const cptr<ItemDesc> it = as_const<ItemDesc>(a.items) + item;
  const int mc = it->mc;
  const int mcp = it->mcp;
  const int units = it->units;
  const int rows_B = it->n > 0 ? 1 + it->L + it->Dt : 0;
  const int n_vals = 1 + rows_B * mc;
  const int v0 = static_cast<int>(blockIdx.y) * kWave;
  if (v0 >= n_vals) return;
  const int64_t stride = grad_unit_stride(rows_B, mcp);
  const double *base = a.partials + 4 * it->partial_base;
  const int tid = static_cast<int>(threadIdx.x);
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int v = v0 + lane;
  cdd s = cdd_zero();
  if (v < n_vals) {
    if (v == 0) {
      if (it->n == 0) {
        if (wave == 0) s.re.hi = 1.0;  // perm of an empty selection
      } else {
        for (int k = wave; k < units; k += kGradFinWaves) {
          const double *p = base + stride * k;
          s = cdd_add(s, cdd{{p[0], p[1]}, {p[2], p[3]}});
        }
      }
    } else {
      const int b = (v - 1) / mc, j = (v - 1) % mc;
      const double *g = base + 4 + static_cast<int64_t>(b) * 2 * mcp + j;
      int k = wave;
      for (; k + 3 * kGradFinWaves < units; k += 4 * kGradFinWaves) {
        double xr[4], xi[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          xr[q] = g[stride * (k + q * kGradFinWaves)];
          xi[q] = g[stride * (k + q * kGradFinWaves) + mcp];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          s.re = dd_add_d(s.re, xr[q]);
          s.im = dd_add_d(s.im, xi[q]);
        }
      }
      for (; k < units; k += kGradFinWaves) {
        s.re = dd_add_d(s.re, g[stride * k]);
        s.im = dd_add_d(s.im, g[stride * k + mcp]);
      }
    }
  }
  part[wave][lane] = s;
  __syncthreads();
  if (wave != 0 || v >= n_vals) return;
  s = part[0][lane];
#pragma unroll
  for (int w = 1; w < kGradFinWaves; ++w) s = cdd_add(s, part[w][lane]);
  const int e = it->n > 0 ? -(it->n - 1) : 0;
  const double re_hi = ldexp(s.re.hi, e), re_lo = ldexp(s.re.lo, e);
  const double im_hi = ldexp(s.im.hi, e), im_lo = ldexp(s.im.lo, e);
  const int64_t o = it->out_index + v;
  a.out[2 * o] = re_hi + re_lo;
  a.out[2 * o + 1] = im_hi + im_lo;
  if (a.out_dd) {
    a.out_dd[4 * o] = re_hi;
    a.out_dd[4 * o + 1] = re_lo;
    a.out_dd[4 * o + 2] = im_hi;
    a.out_dd[4 * o + 3] = im_lo;
  }
}
#endif

// ---------------------------------------------------------------------------
// Instantiation per bucket (-DPQP_BUCKET_ID=0..4): widths 1-8, 9-16,
// 17-19, 20 (two waves per SIMD; width 20 alone keeps its allocation within
// 256 VGPRs) and 21-32 (one wave per SIMD).
// ---------------------------------------------------------------------------
#ifndef PQP_BUCKET_ID
#error "compile with -DPQP_BUCKET_ID=<0..4>"
#endif
#if defined(PQP_LO) && defined(PQP_HI)
// explicit width range (kernel analysis builds)
#elif PQP_BUCKET_ID == 0
#define PQP_LO 1
#define PQP_HI 8
#elif PQP_BUCKET_ID == 1
#define PQP_LO 9
#define PQP_HI 16
#elif PQP_BUCKET_ID == 2
#define PQP_LO 17
#define PQP_HI 19
#elif PQP_BUCKET_ID == 3
#define PQP_LO 20
#define PQP_HI 20
#else
#define PQP_LO 21
#define PQP_HI 32
#endif

#define PQP_PASTE_(base, k) base##k
#define PQP_PASTE(base, k) PQP_PASTE_(base, k)

hipError_t PQP_PASTE(launch_grad_b, PQP_BUCKET_ID)(const LaunchArgs &a, int first_unit,
                                                    int n_units, hipStream_t stream) {
  if (n_units <= 0) return hipSuccess;
  hipLaunchKernelGGL((perm_grad_kernel<PQP_LO, PQP_HI>), dim3(n_units), dim3(kWave), 0,
                     stream, a, first_unit);
  return hipGetLastError();
}

#if PQP_BUCKET_ID == 0
hipError_t launch_grad_finalize(const LaunchArgs &a, int n_items, int max_values,
                                hipStream_t stream) {
  if (n_items <= 0) return hipSuccess;
  const int chunks = (max_values + kWave - 1) / kWave;
  hipLaunchKernelGGL(perm_grad_finalize_kernel, dim3(n_items, chunks > 0 ? chunks : 1),
                     dim3(kGradFinWaves * kWave), 0, stream, a);
  return hipGetLastError();
}
#endif

}  // namespace pqp
