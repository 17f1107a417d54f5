// Seeded blending-style LP generator for the five workloads (see lp.hpp).
//
// The reference ships no generator for these shapes; its nearest relative is
// bench/runners/generate_large_lp.py:49-206 (feasible by construction from a chosen point).
// Every random draw below comes from one SplitMix64 stream in a fixed order, so an instance
// is a pure function of its parameters.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "lp.hpp"

namespace sk {

std::string HostLp::validate() const {
  const int64_t n = num_cols, m = num_rows;
  char buf[160];
  if (n < 0 || m < 0) return "negative dimensions";
  if (static_cast<int64_t>(col_starts.size()) != n + 1) return "col_starts length != num_cols + 1";
  if (static_cast<int64_t>(cost.size()) != n) return "col_cost length != num_cols";
  if (static_cast<int64_t>(col_lower.size()) != n) return "col_lower length != num_cols";
  if (static_cast<int64_t>(col_upper.size()) != n) return "col_upper length != num_cols";
  if (static_cast<int64_t>(row_lower.size()) != m) return "row_lower length != num_rows";
  if (static_cast<int64_t>(row_upper.size()) != m) return "row_upper length != num_rows";
  if (col_starts[0] != 0) return "col_starts[0] != 0";
  const int64_t nz = col_starts[static_cast<size_t>(n)];
  if (static_cast<int64_t>(row_indices.size()) != nz || static_cast<int64_t>(values.size()) != nz)
    return "row_indices/values length != nnz";
  for (int64_t j = 0; j < n; ++j) {
    const int32_t b = col_starts[static_cast<size_t>(j)], e = col_starts[static_cast<size_t>(j) + 1];
    if (e < b) return "col_starts not monotone";
    for (int32_t p = b; p < e; ++p) {
      const int32_t r = row_indices[static_cast<size_t>(p)];
      if (r < 0 || r >= m) {
        std::snprintf(buf, sizeof buf, "row index %d out of range in column %lld", r,
                      static_cast<long long>(j));
        return buf;
      }
      if (p > b && r <= row_indices[static_cast<size_t>(p) - 1]) {
        std::snprintf(buf, sizeof buf, "row indices not strictly ascending in column %lld",
                      static_cast<long long>(j));
        return buf;
      }
      if (!std::isfinite(values[static_cast<size_t>(p)])) {
        std::snprintf(buf, sizeof buf,
                      "matrix entry (%d, %lld) is not a usable coefficient", r,
                      static_cast<long long>(j));
        return buf;
      }
    }
  }
  for (int64_t j = 0; j < n; ++j) {
    const double lo = col_lower[static_cast<size_t>(j)], hi = col_upper[static_cast<size_t>(j)];
    if (std::isnan(lo) || std::isnan(hi)) {
      std::snprintf(buf, sizeof buf, "column %lld has a NaN bound", static_cast<long long>(j));
      return buf;
    }
    if (lo > hi) {
      std::snprintf(buf, sizeof buf, "column %lld has lower bound %g above upper bound %g",
                    static_cast<long long>(j), lo, hi);
      return buf;
    }
    if (std::isnan(cost[static_cast<size_t>(j)])) {
      std::snprintf(buf, sizeof buf, "column %lld has a NaN objective coefficient",
                    static_cast<long long>(j));
      return buf;
    }
  }
  for (int64_t i = 0; i < m; ++i) {
    const double lo = row_lower[static_cast<size_t>(i)], hi = row_upper[static_cast<size_t>(i)];
    if (std::isnan(lo) || std::isnan(hi)) {
      std::snprintf(buf, sizeof buf, "row %lld has a NaN bound", static_cast<long long>(i));
      return buf;
    }
    if (lo > hi) {
      std::snprintf(buf, sizeof buf, "row %lld has lower bound %g above upper bound %g",
                    static_cast<long long>(i), lo, hi);
      return buf;
    }
  }
  if (std::isnan(objective_offset)) return "objective offset is NaN";
  return {};
}

bool preset_params(int config, uint64_t instance, GenParams* p) {
  *p = GenParams{};
  switch (config) {
    case 1:
      p->seed = instance;
      p->num_rows = 500;
      p->num_cols = 1000;
      p->num_equality_rows = 300;
      p->degree_kind = 0;
      p->degree_lo = p->degree_hi = 5;
      return true;
    case 2:
    case 5: {
      const bool large = config == 5;
      p->seed = instance;
      p->num_rows = large ? 2000000 : 200000;
      p->num_cols = large ? 5000000 : 500000;
      p->num_equality_rows = large ? 1200000 : 120000;
      p->degree_kind = 2;
      p->degree_lo = 1;
      p->degree_hi = 2000;
      p->power_exponent = 2.1;
      p->target_nnz = large ? 25000000 : 2500000;
      p->value_kind = large ? 0 : 1;
      p->value_lo = large ? 0.1 : 1e-3;
      p->value_hi = large ? 10.0 : 1e3;
      p->fix_empty_rows = 0;
      return true;
    }
    case 3: {
      SplitMix64 shape(1000 + instance);
      const int32_t m = static_cast<int32_t>(shape.integer(200, 5000));
      const double ratio = shape.uniform(2.0, 3.0);
      p->seed = 1000 + instance;
      p->num_rows = m;
      p->num_cols = static_cast<int32_t>(std::lround(m * ratio));
      p->num_equality_rows = static_cast<int32_t>(std::lround(0.6 * m));
      p->degree_kind = 1;
      p->degree_lo = 4;
      p->degree_hi = 8;
      return true;
    }
    case 4: {
      SplitMix64 shape(100000 + instance);
      const double lm = shape.uniform(std::log(500.0), std::log(50000.0));
      const int32_t m = static_cast<int32_t>(std::lround(std::exp(lm)));
      p->seed = 100000 + instance;
      p->num_rows = m;
      p->num_cols = 2 * m;
      p->num_equality_rows = static_cast<int32_t>(std::lround(0.6 * m));
      p->degree_kind = 0;
      p->degree_lo = p->degree_hi = 5;
      return true;
    }
    default:
      return false;
  }
}

namespace {

// Draw `count` distinct rows out of [0, m) in ascending order.
void distinct_rows(SplitMix64& rng, int32_t m, int32_t count, std::vector<int32_t>* out,
                   std::vector<int32_t>* pool) {
  out->clear();
  if (static_cast<int64_t>(count) * 4 < m) {
    while (static_cast<int32_t>(out->size()) < count) {
      const auto r = static_cast<int32_t>(rng.below(static_cast<uint64_t>(m)));
      if (std::find(out->begin(), out->end(), r) == out->end()) out->push_back(r);
    }
  } else {  // partial Fisher-Yates over [0, m)
    pool->resize(static_cast<size_t>(m));
    std::iota(pool->begin(), pool->end(), 0);
    for (int32_t k = 0; k < count; ++k) {
      const auto pick = k + static_cast<int32_t>(rng.below(static_cast<uint64_t>(m - k)));
      std::swap((*pool)[static_cast<size_t>(k)], (*pool)[static_cast<size_t>(pick)]);
      out->push_back((*pool)[static_cast<size_t>(k)]);
    }
  }
  std::sort(out->begin(), out->end());
}

}  // namespace

HostLp generate_lp(const GenParams& p) {
  SplitMix64 rng(p.seed);
  HostLp lp;
  const int32_t m = p.num_rows, n = p.num_cols;
  lp.num_rows = m;
  lp.num_cols = n;

  // 1. Column degrees.
  std::vector<int32_t> degree(static_cast<size_t>(n), 0);
  const int32_t cap = std::max<int32_t>(1, m);
  if (p.degree_kind == 2) {
    const int32_t lo = std::max(1, p.degree_lo), hi = std::max(lo, p.degree_hi);
    std::vector<double> cdf(static_cast<size_t>(hi - lo + 1));
    double total = 0.0;
    for (int32_t d = lo; d <= hi; ++d) {
      total += std::pow(static_cast<double>(d), -p.power_exponent);
      cdf[static_cast<size_t>(d - lo)] = total;
    }
    for (auto& v : cdf) v /= total;
    for (int32_t j = 0; j < n; ++j) {
      const double u = rng.uniform();
      const auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
      const auto d = lo + static_cast<int32_t>(std::min<std::ptrdiff_t>(
                              it - cdf.begin(), static_cast<std::ptrdiff_t>(cdf.size()) - 1));
      degree[static_cast<size_t>(j)] = std::min(d, cap);
    }
    if (p.target_nnz > 0 && n > 0) {
      int64_t sum = 0;
      for (const auto d : degree) sum += d;
      const int32_t max_degree = std::min(hi, cap);
      int64_t guard = 0;
      while (sum < p.target_nnz && guard < 100 * p.target_nnz) {
        const auto j = static_cast<size_t>(rng.below(static_cast<uint64_t>(n)));
        ++guard;
        if (degree[j] < max_degree) {
          ++degree[j];
          ++sum;
        }
      }
      while (sum > p.target_nnz && guard < 200 * p.target_nnz) {
        const auto j = static_cast<size_t>(rng.below(static_cast<uint64_t>(n)));
        ++guard;
        if (degree[j] > 1) {
          --degree[j];
          --sum;
        }
      }
    }
  } else {
    for (int32_t j = 0; j < n; ++j) {
      const int32_t d = p.degree_kind == 1
                            ? static_cast<int32_t>(rng.integer(p.degree_lo, p.degree_hi))
                            : p.degree_lo;
      degree[static_cast<size_t>(j)] = std::min(std::max(d, 0), cap);
    }
  }

  // 2. Pattern: distinct uniform rows per column, ascending.
  lp.col_starts.assign(static_cast<size_t>(n) + 1, 0);
  for (int32_t j = 0; j < n; ++j)
    lp.col_starts[static_cast<size_t>(j) + 1] =
        lp.col_starts[static_cast<size_t>(j)] + (m > 0 ? degree[static_cast<size_t>(j)] : 0);
  const auto nnz = static_cast<size_t>(lp.col_starts.back());
  lp.row_indices.resize(nnz);
  lp.values.resize(nnz);
  std::vector<int32_t> rows, pool;
  for (int32_t j = 0; j < n && m > 0; ++j) {
    distinct_rows(rng, m, degree[static_cast<size_t>(j)], &rows, &pool);
    std::copy(rows.begin(), rows.end(),
              lp.row_indices.begin() + lp.col_starts[static_cast<size_t>(j)]);
  }

  // 3. Move an entry into every empty row, from a column entry whose row has others.
  if (p.fix_empty_rows && m > 0 && nnz >= static_cast<size_t>(m)) {
    std::vector<int32_t> count(static_cast<size_t>(m), 0);
    for (const auto r : lp.row_indices) ++count[static_cast<size_t>(r)];
    for (int32_t i = 0; i < m; ++i) {
      if (count[static_cast<size_t>(i)] != 0) continue;
      for (;;) {
        const auto j = static_cast<int32_t>(rng.below(static_cast<uint64_t>(n)));
        const int32_t b = lp.col_starts[static_cast<size_t>(j)];
        const int32_t e = lp.col_starts[static_cast<size_t>(j) + 1];
        if (e == b) continue;
        const auto k = b + static_cast<int32_t>(rng.below(static_cast<uint64_t>(e - b)));
        const int32_t old = lp.row_indices[static_cast<size_t>(k)];
        if (count[static_cast<size_t>(old)] <= 1) continue;
        if (std::find(lp.row_indices.begin() + b, lp.row_indices.begin() + e, i) !=
            lp.row_indices.begin() + e)
          continue;
        lp.row_indices[static_cast<size_t>(k)] = i;
        --count[static_cast<size_t>(old)];
        ++count[static_cast<size_t>(i)];
        std::sort(lp.row_indices.begin() + b, lp.row_indices.begin() + e);
        break;
      }
    }
  }

  // 4. Values.
  const double llo = std::log(p.value_lo), lhi = std::log(p.value_hi);
  for (size_t k = 0; k < nnz; ++k) {
    double mag = p.value_kind == 1 ? std::exp(rng.uniform(llo, lhi))
                                   : rng.uniform(p.value_lo, p.value_hi);
    if (rng.uniform() < p.negative_probability) mag = -mag;
    lp.values[k] = mag;
  }

  // 5. Bounds, interior point, costs.
  lp.col_lower.assign(static_cast<size_t>(n), 0.0);
  lp.col_upper.resize(static_cast<size_t>(n));
  lp.cost.resize(static_cast<size_t>(n));
  std::vector<double> x0(static_cast<size_t>(n));
  for (int32_t j = 0; j < n; ++j) {
    const double u = rng.uniform(p.upper_lo, p.upper_hi);
    lp.col_upper[static_cast<size_t>(j)] = u;
    x0[static_cast<size_t>(j)] = rng.uniform(0.0, u);
  }
  for (int32_t j = 0; j < n; ++j) lp.cost[static_cast<size_t>(j)] = rng.uniform(p.cost_lo, p.cost_hi);

  // 6. Rows: b = A x0 (column scatter, j ascending); <= rows get slack.
  std::vector<double> b(static_cast<size_t>(m), 0.0);
  for (int32_t j = 0; j < n; ++j) {
    for (int32_t k = lp.col_starts[static_cast<size_t>(j)]; k < lp.col_starts[static_cast<size_t>(j) + 1]; ++k)
      b[static_cast<size_t>(lp.row_indices[static_cast<size_t>(k)])] +=
          lp.values[static_cast<size_t>(k)] * x0[static_cast<size_t>(j)];
  }
  lp.row_lower.resize(static_cast<size_t>(m));
  lp.row_upper.resize(static_cast<size_t>(m));
  for (int32_t i = 0; i < m; ++i) {
    if (i < p.num_equality_rows) {
      lp.row_lower[static_cast<size_t>(i)] = b[static_cast<size_t>(i)];
      lp.row_upper[static_cast<size_t>(i)] = b[static_cast<size_t>(i)];
    } else {
      lp.row_lower[static_cast<size_t>(i)] = -INFINITY;
      lp.row_upper[static_cast<size_t>(i)] = b[static_cast<size_t>(i)] + rng.uniform(p.slack_lo, p.slack_hi);
    }
  }
  return lp;
}

}  // namespace sk
