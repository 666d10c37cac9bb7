// Host planning of a matrix-polynomial permanent (perm_poly.hip): the
// requested exponent vectors, the variables they use, and the downward
// closure of the requested monomials ordered by total degree, with the
// links a -> a - e_m the row recursion reads.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pqp {

struct PolyPlan {
  int n = 0;                       // matrix order (rows = particles)
  int M = 0;                       // variables with a non-zero exponent somewhere
  int L = 0;                       // largest total degree requested (<= n)
  int64_t n_mono = 0;
  std::vector<int32_t> vars;       // caller's index of each kept variable
  std::vector<int32_t> link_ptr;   // n_mono + 1
  std::vector<int32_t> links;      // (monomial << 8) | variable
  std::vector<int32_t> lvl_off;    // L + 2
  std::vector<int32_t> out_mono;   // per kept output: its monomial
  std::vector<int64_t> out_index;  // per kept output: the caller's index
  int64_t flops_per_subset = 0;    // algorithmic FP64 flops of one subset's product
};

// Exponent rows with total degree above n are dropped (their coefficient is
// zero).  Returns a Status code and an error text.
int poly_plan_build(int64_t n, int64_t n_vars, int64_t n_out, const int32_t *exponents,
                    PolyPlan &plan, std::string &err);

}  // namespace pqp
