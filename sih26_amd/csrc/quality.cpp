// Solution quality measures of a reported LP point, as Solution::recompute_quality defines
// them (src/core/model.cpp:228-422, LP case): absolute and scaled primal violations, dual
// sign and consistency violations, complementarity products, objective and gaps. This is part
// of the Solution contract every engine honours before returning, not of the PDHG loop.
#include "quality.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace sk {

void recompute_quality(const HostLp& lp, const double* x, const double* row_dual,
                       const double* col_dual, double dual_bound, Quality* q) {
  const int32_t n = lp.num_cols, m = lp.num_rows;
  const int32_t* cs = lp.col_starts.data();
  const int32_t* ri = lp.row_indices.data();
  const double* av = lp.values.data();
  std::vector<double> activity(static_cast<size_t>(m), 0.0);
  for (int32_t j = 0; j < n; ++j) {
    if (x[j] == 0.0) continue;
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) activity[static_cast<size_t>(ri[p])] += av[p] * x[j];
  }
  q->primal_infeasibility = 0.0;
  q->primal_infeasibility_scaled = 0.0;
  std::vector<double> row_scale(static_cast<size_t>(m), 1.0);
  for (int32_t j = 0; j < n; ++j) {
    if (x[j] == 0.0) continue;
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) {
      double& s = row_scale[static_cast<size_t>(ri[p])];
      s = std::max(s, std::fabs(av[p] * x[j]));
    }
  }
  const auto record = [&](double violation, double scale) {
    if (violation <= 0.0) return;
    q->primal_infeasibility = std::max(q->primal_infeasibility, violation);
    q->primal_infeasibility_scaled =
        std::max(q->primal_infeasibility_scaled, violation / std::max(1.0, scale));
  };
  for (int32_t j = 0; j < n; ++j) {
    if (is_finite_bound(lp.col_lower[static_cast<size_t>(j)]))
      record(lp.col_lower[static_cast<size_t>(j)] - x[j], std::fabs(x[j]));
    if (is_finite_bound(lp.col_upper[static_cast<size_t>(j)]))
      record(x[j] - lp.col_upper[static_cast<size_t>(j)], std::fabs(x[j]));
  }
  for (int32_t i = 0; i < m; ++i) {
    const auto u = static_cast<size_t>(i);
    if (is_finite_bound(lp.row_lower[u])) record(lp.row_lower[u] - activity[u], row_scale[u]);
    if (is_finite_bound(lp.row_upper[u])) record(activity[u] - lp.row_upper[u], row_scale[u]);
  }
  double linear = lp.objective_offset;  // Model::evaluate_objective, model.cpp:99-106
  for (int32_t j = 0; j < n; ++j) linear += lp.cost[static_cast<size_t>(j)] * x[j];
  q->objective = linear;

  q->dual_infeasibility = 0.0;
  q->dual_infeasibility_scaled = 0.0;
  q->complementarity_violation = 0.0;
  const double sense = static_cast<double>(lp.sense);
  double dual_norm = 0.0;
  for (int32_t i = 0; i < m; ++i) dual_norm = std::max(dual_norm, std::fabs(row_dual[i]));
  const auto record_dual = [&](double violation, double scale) {
    if (violation <= 0.0) return;
    q->dual_infeasibility = std::max(q->dual_infeasibility, violation);
    q->dual_infeasibility_scaled = std::max(q->dual_infeasibility_scaled, violation / scale);
  };
  const auto record_compl = [&](double mult, double slack, double mscale, double pscale) {
    if (mult == 0.0 || !is_finite_bound(slack)) return;
    const double product = mult * slack;
    q->complementarity_violation = std::max(q->complementarity_violation, product);
    q->dual_infeasibility_scaled =
        std::max(q->dual_infeasibility_scaled, product / std::max(1.0, mscale * pscale));
  };
  const auto distance = [](double v, double lo, double hi) {
    double d = INFINITY;
    if (is_finite_bound(lo)) d = std::min(d, std::fabs(v - lo));
    if (is_finite_bound(hi)) d = std::min(d, std::fabs(v - hi));
    return d;
  };
  for (int32_t j = 0; j < n; ++j) {
    double implied = lp.cost[static_cast<size_t>(j)];
    double terms = std::fabs(lp.cost[static_cast<size_t>(j)]);
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) {
      const double term = av[p] * row_dual[ri[p]];
      implied -= term;
      terms = std::max(terms, std::fabs(term));
    }
    record_dual(std::fabs(implied - col_dual[j]), std::max(1.0, terms));
  }
  for (int32_t j = 0; j < n; ++j) {
    const auto u = static_cast<size_t>(j);
    const double d = sense * col_dual[j];
    const double lo = lp.col_lower[u], hi = lp.col_upper[u];
    if (lo == hi) continue;
    double scale = std::fabs(lp.cost[u]);
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) scale = std::max(scale, std::fabs(av[p] * row_dual[ri[p]]));
    scale = std::max(1.0, scale);
    if (d > 0.0 && !is_finite_bound(lo)) record_dual(d, scale);
    if (d < 0.0 && !is_finite_bound(hi)) record_dual(-d, scale);
    record_compl(std::fabs(d), distance(x[j], lo, hi), scale, std::max(1.0, std::fabs(x[j])));
  }
  for (int32_t i = 0; i < m; ++i) {
    const auto u = static_cast<size_t>(i);
    const double y = sense * row_dual[i];
    const double lo = lp.row_lower[u], hi = lp.row_upper[u];
    if (lo == hi) continue;
    const double price_scale = std::max(1.0, dual_norm);
    if (y > 0.0 && !is_finite_bound(lo)) record_dual(y, price_scale);
    if (y < 0.0 && !is_finite_bound(hi)) record_dual(-y, price_scale);
    record_compl(std::fabs(y), distance(activity[u], lo, hi), price_scale, row_scale[u]);
  }
  q->absolute_gap = std::fabs(q->objective - dual_bound);
  q->relative_gap = q->absolute_gap / std::max(1.0, std::fabs(q->objective));
}

}  // namespace sk
