// Host planners under AddressSanitizer / UBSan: random permanent plans
// (shapes, multiplicities, FP32 tables, column minors, column expansion,
// single-permanent wave counts, batch unit shares), their error cases, and
// matrix-polynomial closures (packed-key and byte-row paths).  Exits 0 when
// every call returned the expected status; the sanitizers abort otherwise.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "plan.hpp"
#include "poly_plan.hpp"

using namespace pqp;

int main() {
  std::mt19937_64 g(11);
  std::normal_distribution<double> nd;
  int failures = 0, built = 0, poly_built = 0;
  for (int trial = 0; trial < 400; ++trial) {
    const int count = 1 + static_cast<int>(g() % 6);
    std::vector<std::vector<double>> mats(count);
    std::vector<std::vector<int32_t>> rows(count), cols(count);
    std::vector<PermInput> in;
    const bool minors = trial % 4 == 1;
    for (int k = 0; k < count; ++k) {
      const int mr = 1 + static_cast<int>(g() % 10), mc = 1 + static_cast<int>(g() % 10);
      mats[k].resize(2 * mr * mc);
      for (double &x : mats[k]) x = nd(g);
      rows[k].assign(mr, 0);
      cols[k].assign(mc, 0);
      const int n = static_cast<int>(g() % 12);
      for (int p = 0; p < n; ++p) rows[k][g() % mr]++;
      for (int p = 0; p < n + (minors ? 1 : 0); ++p) cols[k][g() % mc]++;
      in.push_back(PermInput{mats[k].data(), false, mr, mc, mc, rows[k].data(), cols[k].data()});
    }
    PlanOptions opt;
    opt.minors = minors;
    opt.fp32_products = !minors && trial % 3 == 0;
    opt.expand_cols = static_cast<int>(g() % 3);
    if (trial % 5 == 0) opt.auto_units_cap = 1024;
    if (trial % 5 == 1) opt.total_units = 2048;
    if (trial % 7 == 0) opt.direct_out = true;
    Plan p;
    std::string err;
    const int st = plan_build(in, opt, p, err);
    built += st == kOk;
    if (st != kOk && st != kErrRange) {
      std::printf("trial %d: unexpected status %d (%s)\n", trial, st, err.c_str());
      ++failures;
    }
  }
  // error cases: sum mismatch, negative multiplicity, null matrix
  {
    std::vector<double> a(8, 1.0);
    int32_t r[2] = {1, 1}, c[2] = {2, 1}, neg[2] = {-1, 3};
    Plan p;
    std::string err;
    if (plan_build({PermInput{a.data(), false, 2, 2, 2, r, c}}, PlanOptions(), p, err) != kErrSumMismatch)
      ++failures;
    if (plan_build({PermInput{a.data(), false, 2, 2, 2, neg, c}}, PlanOptions(), p, err) == kOk) ++failures;
    if (plan_build({PermInput{nullptr, false, 2, 2, 2, r, r}}, PlanOptions(), p, err) != kErrInvalid)
      ++failures;
  }
  // matrix-polynomial closures
  for (int trial = 0; trial < 200; ++trial) {
    const int n = 1 + static_cast<int>(g() % 8);
    const int vars = trial % 10 == 0 ? 40 : static_cast<int>(g() % 7);
    const int rowsn = 1 + static_cast<int>(g() % 30);
    std::vector<int32_t> ex(static_cast<size_t>(rowsn) * vars);
    for (int32_t &x : ex) x = static_cast<int32_t>(g() % 3) - (trial % 17 == 0 && g() % 50 == 0 ? 3 : 0);
    PolyPlan pp;
    std::string err;
    const int st = poly_plan_build(n, vars, rowsn, ex.empty() ? nullptr : ex.data(), pp, err);
    poly_built += st == kOk;
    if (st != kOk && st != kErrInvalid) {
      std::printf("poly trial %d: unexpected status %d (%s)\n", trial, st, err.c_str());
      ++failures;
    }
  }
  std::printf("plans built %d/400, polynomial plans built %d/200\n", built, poly_built);
  std::printf("host planners: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
