// Downward closure of the requested monomials for perm_poly.hip.
//
// Monomials are packed into one 64-bit key when M x (bits of the largest
// degree) fits, else rows of M bytes (exponents <= n <= 24) in one flat
// array; either is found through an open-addressing hash table.  The closure
// is generated level by level from the highest total degree down: every
// monomial of degree D adds its divisors a - e_m to degree D - 1.  The final
// order is by degree (level 0 = the constant monomial first), which is the
// order the kernel updates them in.
#include "poly_plan.hpp"

#include <algorithm>
#include <cstring>

#include "plan.hpp"
#include "poly_types.hpp"

namespace pqp {

namespace {

struct MonoSet {
  int M;
  std::vector<uint8_t> rows;     // n_mono x M
  std::vector<int32_t> table;    // -1 = empty, else a row index
  size_t mask = 0;
  int64_t count = 0;

  explicit MonoSet(int m) : M(m) { rehash(1024); }

  uint64_t hash(const uint8_t *r) const {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int m = 0; m < M; ++m) {
      h ^= r[m];
      h *= 0x100000001B3ull;
      h ^= h >> 29;
    }
    return h;
  }

  void rehash(size_t cap) {
    table.assign(cap, -1);
    mask = cap - 1;
    for (int64_t k = 0; k < count; ++k) {
      size_t s = hash(&rows[static_cast<size_t>(k) * M]) & mask;
      while (table[s] >= 0) s = (s + 1) & mask;
      table[s] = static_cast<int32_t>(k);
    }
  }

  // index of row r, inserting it when absent; *added tells which
  int64_t find_or_add(const uint8_t *r, bool *added) {
    size_t s = hash(r) & mask;
    while (table[s] >= 0) {
      if (std::memcmp(&rows[static_cast<size_t>(table[s]) * M], r, static_cast<size_t>(M)) == 0) {
        *added = false;
        return table[s];
      }
      s = (s + 1) & mask;
    }
    *added = true;
    const int64_t k = count++;
    rows.insert(rows.end(), r, r + M);
    table[s] = static_cast<int32_t>(k);
    if (static_cast<size_t>(count) * 2 > table.size()) rehash(table.size() * 2);
    return k;
  }

  int64_t find(const uint8_t *r) const {
    size_t s = hash(r) & mask;
    while (table[s] >= 0) {
      if (std::memcmp(&rows[static_cast<size_t>(table[s]) * M], r, static_cast<size_t>(M)) == 0)
        return table[s];
      s = (s + 1) & mask;
    }
    return -1;
  }
};

// Monomials packed into one 64-bit key, `bits` bits per variable (used
// when M * bits <= 64): hashing and the step a -> a - e_v are single
// integer operations.
struct PackedSet {
  // open addressing with the key stored beside its insertion index, so a
  // probe touches one cache line instead of the slot and then keys[]
  struct Slot {
    uint64_t key;
    int64_t idx;  // -1 = empty
  };
  int bits;
  std::vector<uint64_t> keys;  // by insertion index
  std::vector<Slot> table;
  size_t mask = 0;
  int shift = 0;

  PackedSet(int b, size_t expect) : bits(b) {
    size_t cap = 1024;
    while (cap < 2 * expect) cap *= 2;
    keys.reserve(expect);
    rehash(cap);
  }

  size_t slot(uint64_t k) const { return static_cast<size_t>((k * 0x9E3779B97F4A7C15ull) >> shift); }

  void rehash(size_t cap) {
    table.assign(cap, Slot{0, -1});
    mask = cap - 1;
    shift = 64 - __builtin_ctzll(cap);
    for (size_t k = 0; k < keys.size(); ++k) {
      size_t s = slot(keys[k]);
      while (table[s].idx >= 0) s = (s + 1) & mask;
      table[s] = Slot{keys[k], static_cast<int64_t>(k)};
    }
  }

  int64_t find_or_add(uint64_t key, bool *added) {
    size_t s = slot(key);
    while (table[s].idx >= 0) {
      if (table[s].key == key) {
        *added = false;
        return table[s].idx;
      }
      s = (s + 1) & mask;
    }
    *added = true;
    const int64_t k = static_cast<int64_t>(keys.size());
    keys.push_back(key);
    table[s] = Slot{key, k};
    if (keys.size() * 2 > table.size()) rehash(table.size() * 2);
    return k;
  }
};

// The closure with packed keys; fills level lists (insertion indices) and,
// per monomial, its links (insertion index of a - e_v, and v) in the order
// the monomials were expanded.
int closure_packed(const PolyPlan &p, int64_t n_vars, const int32_t *exponents, int bits,
                   std::vector<std::vector<int64_t>> &level, std::vector<int64_t> &out_raw,
                   std::vector<int64_t> &link_first, std::vector<int32_t> &link_cnt,
                   std::vector<int64_t> &raw_links, int64_t &n_mono, std::string &err) {
  const int M = p.M;
  const uint64_t field = (uint64_t{1} << bits) - 1;
  // the closure holds at least the requested monomials
  PackedSet set(bits, p.out_index.size() + 1);
  link_first.reserve(p.out_index.size() + 1);
  link_cnt.reserve(p.out_index.size() + 1);
  raw_links.reserve(p.out_index.size() * static_cast<size_t>(std::min(M, p.L)) + 1);
  for (int64_t q : p.out_index) {
    uint64_t key = 0;
    int deg = 0;
    for (int v = 0; v < M; ++v) {
      const uint64_t e = static_cast<uint64_t>(exponents[q * n_vars + p.vars[v]]);
      key |= e << (bits * v);
      deg += static_cast<int>(e);
    }
    bool added = false;
    const int64_t k = set.find_or_add(key, &added);
    if (added) level[deg].push_back(k);
    out_raw.push_back(k);
  }
  for (int D = p.L; D >= 1; --D) {
    for (size_t t = 0; t < level[D].size(); ++t) {
      const int64_t k = level[D][t];
      const uint64_t key = set.keys[k];
      if (static_cast<int64_t>(link_first.size()) <= k) {
        link_first.resize(static_cast<size_t>(k) + 1, -1);
        link_cnt.resize(static_cast<size_t>(k) + 1, 0);
      }
      link_first[k] = static_cast<int64_t>(raw_links.size());
      for (int v = 0; v < M; ++v) {
        if (((key >> (bits * v)) & field) == 0) continue;
        bool added = false;
        const int64_t j = set.find_or_add(key - (uint64_t{1} << (bits * v)), &added);
        if (added) level[D - 1].push_back(j);
        raw_links.push_back(j << 8 | v);
      }
      link_cnt[k] = static_cast<int32_t>(static_cast<int64_t>(raw_links.size()) - link_first[k]);
      if (static_cast<int64_t>(set.keys.size()) > kPolyMaxMono) {
        err = "more than " + std::to_string(kPolyMaxMono) + " monomials";
        return kErrRange;
      }
    }
  }
  if (level[0].empty()) {
    bool added = false;
    level[0].push_back(set.find_or_add(0, &added));
  }
  n_mono = static_cast<int64_t>(set.keys.size());
  link_first.resize(static_cast<size_t>(n_mono), -1);
  link_cnt.resize(static_cast<size_t>(n_mono), 0);
  return kOk;
}

// One subset: row i updates degrees 0..min(i + 1, L); a fresh top degree
// costs a complex multiply and (links - 1) complex FMAs per monomial, every
// other one a complex multiply plus one FMA per link.
void count_flops(PolyPlan &p, const std::vector<int64_t> &lvl_links) {
  p.flops_per_subset = 0;
  for (int i = 0; i < p.n; ++i) {
    const int topd = std::min(i + 1, p.L);
    for (int D = 0; D <= topd; ++D) {
      const int64_t cnt = p.lvl_off[D + 1] - p.lvl_off[D];
      if (D == i + 1)
        p.flops_per_subset += 8 * lvl_links[D] - 2 * cnt;
      else
        p.flops_per_subset += 6 * cnt + 8 * lvl_links[D];
    }
  }
}

}  // namespace

int poly_plan_build(int64_t n, int64_t n_vars, int64_t n_out, const int32_t *exponents,
                    PolyPlan &p, std::string &err) {
  if (n < 1 || n_vars < 0 || n_out < 0 || (n_out > 0 && n_vars > 0 && !exponents)) {
    err = "invalid matrix polynomial arguments";
    return kErrInvalid;
  }
  if (n > kPolyMaxN) {
    err = "matrix order " + std::to_string(n) + " exceeds " + std::to_string(kPolyMaxN);
    return kErrRange;
  }
  p = PolyPlan();
  p.n = static_cast<int>(n);
  // kept outputs (total degree <= n) and the variables they use
  std::vector<int32_t> top(static_cast<size_t>(n_vars), 0);
  for (int64_t q = 0; q < n_out; ++q) {
    int64_t deg = 0;
    for (int64_t m = 0; m < n_vars; ++m) {
      const int32_t e = exponents[q * n_vars + m];
      if (e < 0) {
        err = "negative exponent in row " + std::to_string(q);
        return kErrInvalid;
      }
      deg += e;
    }
    if (deg > n) continue;
    p.out_index.push_back(q);
    p.L = std::max<int>(p.L, static_cast<int>(deg));
    for (int64_t m = 0; m < n_vars; ++m) top[m] = std::max(top[m], exponents[q * n_vars + m]);
  }
  for (int64_t m = 0; m < n_vars; ++m)
    if (top[m] > 0) p.vars.push_back(static_cast<int32_t>(m));
  p.M = static_cast<int>(p.vars.size());
  if (p.M > kPolyMaxVars) {
    err = std::to_string(p.M) + " variables in use exceed " + std::to_string(kPolyMaxVars);
    return kErrRange;
  }
  const int M = p.M, L = p.L;
  if (p.out_index.empty()) return kOk;
  const int bits = 32 - __builtin_clz(static_cast<unsigned>(std::max(p.L, 1)));
  if (M * bits <= 64) {
    std::vector<std::vector<int64_t>> level(static_cast<size_t>(L) + 1);
    std::vector<int64_t> out_raw, link_first, raw_links;
    std::vector<int32_t> link_cnt;
    out_raw.reserve(p.out_index.size());
    int st = closure_packed(p, n_vars, exponents, bits, level, out_raw, link_first, link_cnt,
                            raw_links, p.n_mono, err);
    if (st != kOk) return st;
    std::vector<int32_t> order_of(static_cast<size_t>(p.n_mono), -1);
    p.lvl_off.assign(static_cast<size_t>(L) + 2, 0);
    int32_t next = 0;
    for (int D = 0; D <= L; ++D) {
      p.lvl_off[D] = next;
      for (int64_t k : level[D]) order_of[k] = next++;
    }
    p.lvl_off[L + 1] = next;
    p.link_ptr.assign(static_cast<size_t>(p.n_mono) + 1, 0);
    p.links.resize(raw_links.size());
    std::vector<int64_t> lvl_links(static_cast<size_t>(L) + 1, 0);
    int32_t pos = 0;
    for (int D = 0; D <= L; ++D) {
      for (int64_t k : level[D]) {
        p.link_ptr[order_of[k]] = pos;
        for (int32_t c = 0; c < link_cnt[k]; ++c) {
          const int64_t e = raw_links[static_cast<size_t>(link_first[k] + c)];
          p.links[pos++] = order_of[e >> 8] << 8 | static_cast<int32_t>(e & 0xff);
        }
        lvl_links[D] += link_cnt[k];
      }
    }
    p.link_ptr[p.n_mono] = pos;
    p.out_mono.resize(out_raw.size());
    for (size_t q = 0; q < out_raw.size(); ++q) p.out_mono[q] = order_of[out_raw[q]];
    count_flops(p, lvl_links);
    return kOk;
  }
  MonoSet set(std::max(M, 1));
  std::vector<std::vector<int64_t>> level(static_cast<size_t>(L) + 1);
  std::vector<uint8_t> row(static_cast<size_t>(std::max(M, 1)), 0);
  std::vector<int64_t> out_raw;
  out_raw.reserve(p.out_index.size());
  for (int64_t q : p.out_index) {
    int deg = 0;
    for (int v = 0; v < M; ++v) {
      row[v] = static_cast<uint8_t>(exponents[q * n_vars + p.vars[v]]);
      deg += row[v];
    }
    bool added = false;
    const int64_t k = set.find_or_add(row.data(), &added);
    if (added) level[deg].push_back(k);
    out_raw.push_back(k);
  }
  for (int D = L; D >= 1; --D) {
    for (size_t t = 0; t < level[D].size(); ++t) {
      const int64_t k = level[D][t];
      // a copy: find_or_add may move the row storage
      std::memcpy(row.data(), &set.rows[static_cast<size_t>(k) * M], static_cast<size_t>(M));
      for (int v = 0; v < M; ++v) {
        if (row[v] == 0) continue;
        --row[v];
        bool added = false;
        const int64_t j = set.find_or_add(row.data(), &added);
        if (added) level[D - 1].push_back(j);
        ++row[v];
      }
      if (set.count > kPolyMaxMono) {
        err = "more than " + std::to_string(kPolyMaxMono) + " monomials";
        return kErrRange;
      }
    }
  }
  if (level[0].empty()) {  // the constant monomial (L = 0 with only e = 0 requested adds it above)
    std::fill(row.begin(), row.end(), 0);
    bool added = false;
    level[0].push_back(set.find_or_add(row.data(), &added));
  }
  p.n_mono = set.count;
  std::vector<int32_t> order_of(static_cast<size_t>(set.count), -1);
  p.lvl_off.assign(static_cast<size_t>(L) + 2, 0);
  int32_t next = 0;
  std::vector<int64_t> by_order;
  by_order.reserve(static_cast<size_t>(set.count));
  for (int D = 0; D <= L; ++D) {
    p.lvl_off[D] = next;
    for (int64_t k : level[D]) {
      order_of[k] = next++;
      by_order.push_back(k);
    }
  }
  p.lvl_off[L + 1] = next;
  p.link_ptr.assign(static_cast<size_t>(p.n_mono) + 1, 0);
  std::vector<int64_t> lvl_links(static_cast<size_t>(L) + 1, 0);
  for (int D = 0; D <= L; ++D) {
    for (int32_t a = p.lvl_off[D]; a < p.lvl_off[D + 1]; ++a) {
      const int64_t k = by_order[a];
      p.link_ptr[a] = static_cast<int32_t>(p.links.size());
      std::memcpy(row.data(), &set.rows[static_cast<size_t>(k) * M], static_cast<size_t>(M));
      for (int v = 0; v < M; ++v) {
        if (row[v] == 0) continue;
        --row[v];
        const int64_t j = set.find(row.data());
        p.links.push_back(order_of[j] << 8 | v);
        ++row[v];
      }
      lvl_links[D] += static_cast<int64_t>(p.links.size()) - p.link_ptr[a];
    }
  }
  p.link_ptr[p.n_mono] = static_cast<int32_t>(p.links.size());
  p.out_mono.resize(out_raw.size());
  for (size_t q = 0; q < out_raw.size(); ++q) p.out_mono[q] = order_of[out_raw[q]];
  count_flops(p, lvl_links);
  return kOk;
}

}  // namespace pqp
