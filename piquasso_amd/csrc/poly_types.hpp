// Device-side layout of one matrix-polynomial permanent problem
// (perm_poly.hip): coefficients [x^e] perm(B_0 + sum_m x_m B_m) by Ryser's
// formula over the column subsets, with each subset's polynomial kept on a
// downward-closed set of monomials stored level by level (total degree
// 0, 1, ..., L).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pqp {

constexpr int kPolyMaxN = 24;         // 2^24 - 1 column subsets at most
constexpr int kPolyMaxVars = 255;     // variable index in the low 8 bits of a link
constexpr int kPolyMaxMono = 1 << 23; // monomial index in the upper 23 bits of a link
constexpr int kPolyThreads = 256;          // polynomial in LDS
constexpr int kPolyThreadsGlobal = 1024;   // polynomial in global memory
constexpr int kPolyGroup = 2;              // subsets interleaved per workgroup (global memory)

struct PolyArgs {
  const double2 *b0;       // n x n, row-major
  const double2 *bm;       // M matrices n x n, one after another
  const int32_t *link_ptr; // n_mono + 1: links of monomial a are [link_ptr[a], link_ptr[a + 1])
  const int32_t *links;    // (index of a - e_m) << 8 | m, for every m with a_m > 0
  const int32_t *lvl_off;  // L + 2: monomials of total degree D are [lvl_off[D], lvl_off[D + 1])
  const int32_t *out_mono; // n_out: monomial of each requested coefficient
  double2 *scratch;        // global-memory polynomials, n_mono x group per workgroup (when not in LDS)
  double2 *acc;            // n_wg x n_out: each workgroup's signed subset sum
  double *out;             // 2 per requested coefficient
  int32_t n, M, L, n_mono, n_out, n_wg;
};

}  // namespace pqp
