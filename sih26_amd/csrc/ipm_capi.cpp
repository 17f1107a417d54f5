// C ABI of the interior-point finish: sk_ipm_solve (ipm::solve_ipm, include/sankhya/ipm.hpp)
// and sk_pdhg_polish (polish_with_the_interior_point, src/core/solve.cpp:184-243).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sk_gpu.h"
#include "device_types.hpp"
#include "engine.hpp"
#include "handles.hpp"
#include "ipm.hpp"
#include "quality.hpp"
#include "status.hpp"

namespace {

template <class F>
int32_t guarded(F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    sk::set_last_error("host out of memory");
    return SK_GPU_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    sk::set_last_error(std::string("internal error: ") + e.what());
    return SK_GPU_INVALID_STATE;
  } catch (...) {
    sk::set_last_error("internal error");
    return SK_GPU_INVALID_STATE;
  }
}

// SolveStatus strings (model.cpp:19-34)
const char* status_name(int32_t s) {
  static const char* const names[] = {"not_solved", "optimal", "infeasible", "unbounded",
                                      "infeasible_or_unbounded", "feasible", "iteration_limit",
                                      "time_limit", "node_limit", "numerical_error",
                                      "model_error"};
  return s >= 0 && s <= 10 ? names[s] : "unknown";
}

// model.hpp:99-112
bool claims_a_point(int32_t s) {
  return s == sk::kStOptimal || s == sk::kStFeasible || s == 3 || s == sk::kStIterationLimit ||
         s == sk::kStTimeLimit || s == 8;
}

void copy_text(char* dst, size_t cap, const std::string& src) {
  const size_t k = std::min(cap - 1, src.size());
  std::memcpy(dst, src.data(), k);
  dst[k] = '\0';
}

// `head` followed by `note` in a buffer of `cap` bytes; when both do not fit, `head` is cut
// (marked with "...") so that the note survives whole.
void join_text(char* dst, size_t cap, const std::string& head, const std::string& note) {
  if (head.size() + note.size() < cap || note.size() + 4 >= cap) {
    copy_text(dst, cap, head + note);
    return;
  }
  copy_text(dst, cap, head.substr(0, cap - 1 - note.size() - 3) + "..." + note);
}

sk::IpmOptions to_ipm(const sk_ipm_options* o) {
  sk::IpmOptions r;
  if (!o) return r;
  r.iteration_limit = o->iteration_limit;
  r.start_margin = o->start_margin;
  r.max_seconds = o->max_seconds;
  r.max_factor_nonzeros = o->max_factor_nonzeros;
  r.scaling = o->scaling;
  return r;
}

// The interior point's answer as a result record and arrays (ipm.cpp:597-677 finish()).
void fill(const sk::HostLp& lp, const sk::IpmOutcome& o, sk_pdhg_result* res, double* col_value,
          double* row_dual, double* col_dual, double* row_activity) {
  const auto n = static_cast<size_t>(lp.num_cols), m = static_cast<size_t>(lp.num_rows);
  const bool point = !o.col_value.empty();
  std::vector<double> activity(m, 0.0);
  if (point)
    for (size_t j = 0; j < n; ++j)
      for (int32_t q = lp.col_starts[j]; q < lp.col_starts[j + 1]; ++q)
        activity[static_cast<size_t>(lp.row_indices[static_cast<size_t>(q)])] +=
            lp.values[static_cast<size_t>(q)] * o.col_value[j];
  if (col_value) {
    if (point) std::copy(o.col_value.begin(), o.col_value.end(), col_value);
    else std::fill(col_value, col_value + n, 0.0);
  }
  if (col_dual) {
    if (point) std::copy(o.col_dual.begin(), o.col_dual.end(), col_dual);
    else std::fill(col_dual, col_dual + n, 0.0);
  }
  if (row_dual) {
    if (point) std::copy(o.row_dual.begin(), o.row_dual.end(), row_dual);
    else std::fill(row_dual, row_dual + m, 0.0);
  }
  if (row_activity) std::copy(activity.begin(), activity.end(), row_activity);
  if (!res) return;
  std::memset(res, 0, sizeof(*res));
  res->status = o.status;
  res->converged = o.status == sk::kStOptimal ? 1 : 0;
  res->iterations = o.iterations;
  res->solve_seconds = o.seconds;
  const double unbounded = lp.sense < 0 ? INFINITY : -INFINITY;
  double objective = lp.objective_offset;
  if (point)
    for (size_t j = 0; j < n; ++j) objective += lp.cost[j] * o.col_value[j];
  res->objective = objective;
  // an optimal claim bounds the objective by itself (ipm.cpp:671-675)
  res->dual_bound = o.status == sk::kStOptimal ? objective : unbounded;
  if (point) {
    sk::Quality q;
    sk::recompute_quality(lp, o.col_value.data(), o.row_dual.data(), o.col_dual.data(),
                          res->dual_bound, &q);
    res->objective = q.objective;
    res->primal_infeasibility = q.primal_infeasibility;
    res->primal_infeasibility_scaled = q.primal_infeasibility_scaled;
    res->dual_infeasibility = q.dual_infeasibility;
    res->dual_infeasibility_scaled = q.dual_infeasibility_scaled;
    res->complementarity_violation = q.complementarity_violation;
    res->absolute_gap = q.absolute_gap;
    res->relative_gap = q.relative_gap;
  }
  copy_text(res->algorithm, sizeof(res->algorithm), "ipm");
  copy_text(res->message, sizeof(res->message), o.message);
}

}  // namespace

extern "C" {

void sk_ipm_default_options(sk_ipm_options* o) {
  if (!o) return;
  const sk::IpmOptions d;
  o->iteration_limit = d.iteration_limit;
  o->start_margin = d.start_margin;
  o->max_seconds = d.max_seconds;
  o->max_factor_nonzeros = d.max_factor_nonzeros;
  o->scaling = d.scaling;
  o->reserved = 0;
}

int32_t sk_ipm_solve(const sk_lp* lp, const sk_ipm_options* options,
                     const double* warm_col_value, const double* warm_row_dual,
                     const double* warm_col_dual, sk_pdhg_result* result, double* col_value,
                     double* row_dual, double* col_dual, double* row_activity) {
  if (!lp) return SK_GPU_INVALID_ARGUMENT;
  if (options && (options->start_margin <= 0.0 || !(options->max_seconds > 0.0))) {
    sk::set_last_error("sk_ipm_solve: start_margin and max_seconds must be positive");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int32_t {
    sk::WarmPoint warm{warm_col_value, warm_row_dual, warm_col_dual};
    const bool use_warm = warm_col_value && warm_row_dual && warm_col_dual;
    sk::IpmOutcome o;
    const int st = sk::solve_ipm(lp->lp, to_ipm(options), use_warm ? &warm : nullptr, &o);
    if (st) return st;
    fill(lp->lp, o, result, col_value, row_dual, col_dual, row_activity);
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_pdhg_polish(const sk_lp* lp, const sk_ipm_options* options,
                       const sk_pdhg_result* first, const double* first_col_value,
                       const double* first_row_dual, const double* first_col_dual,
                       sk_pdhg_result* result, double* col_value, double* row_dual,
                       double* col_dual, double* row_activity) {
  if (!lp || !first || !first_col_value || !first_row_dual || !first_col_dual || !result)
    return SK_GPU_INVALID_ARGUMENT;
  if (options && (options->start_margin <= 0.0 || !(options->max_seconds > 0.0))) {
    sk::set_last_error("sk_pdhg_polish: start_margin and max_seconds must be positive");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int32_t {
    const sk::HostLp& model = lp->lp;
    const auto n = static_cast<size_t>(model.num_cols), m = static_cast<size_t>(model.num_rows);
    auto keep_first = [&](const std::string& note, int64_t polish_iterations) {
      const sk_pdhg_result copy = *first;  // `result` may alias `first`
      *result = copy;
      result->polish_iterations = polish_iterations;
      if (!note.empty()) join_text(result->message, sizeof(result->message), copy.message, note);
      if (col_value && col_value != first_col_value) std::copy(first_col_value, first_col_value + n, col_value);
      if (row_dual && row_dual != first_row_dual) std::copy(first_row_dual, first_row_dual + m, row_dual);
      if (col_dual && col_dual != first_col_dual) std::copy(first_col_dual, first_col_dual + n, col_dual);
      if (row_activity) {
        std::fill(row_activity, row_activity + m, 0.0);
        for (size_t j = 0; j < n; ++j)
          for (int32_t q = model.col_starts[j]; q < model.col_starts[j + 1]; ++q)
            row_activity[model.row_indices[static_cast<size_t>(q)]] +=
                model.values[static_cast<size_t>(q)] * first_col_value[j];
      }
    };
    // solve.cpp:187-193: nothing to polish
    if (first->status == sk::kStOptimal || !claims_a_point(first->status)) {
      keep_first("", 0);
      return SK_GPU_SUCCESS;
    }
    sk::WarmPoint warm{first_col_value, first_row_dual, first_col_dual};
    sk::IpmOutcome o;
    const int st = sk::solve_ipm(model, to_ipm(options), &warm, &o);
    if (st) return st;
    // solve.cpp:219-242: replace only with an optimal answer, or a feasible one that is less
    // violated (scaled primal and dual violations) than the first
    sk::Quality qf;
    sk::recompute_quality(model, first_col_value, first_row_dual, first_col_dual,
                          first->dual_bound, &qf);
    const double worst_first = std::max(qf.primal_infeasibility_scaled, qf.dual_infeasibility_scaled);
    sk_pdhg_result polished;
    std::vector<double> x(n), yd(m), d(n), act(m);
    fill(model, o, &polished, x.data(), yd.data(), d.data(), act.data());
    const double worst_polished = std::max(polished.primal_infeasibility_scaled,
                                           polished.dual_infeasibility_scaled);
    const bool better = o.status == sk::kStOptimal ||
                        (o.status == sk::kStFeasible && worst_polished < worst_first);
    if (!better) {
      keep_first("; the interior-point polish did not improve it (" +
                     std::string(status_name(o.status)) + " after " +
                     std::to_string(o.iterations) + " iterations: " + o.message + ")",
                 o.iterations);
      return SK_GPU_SUCCESS;
    }
    const std::string first_algorithm(first->algorithm);
    const std::string first_message(first->message);
    const int64_t first_iterations = first->iterations;
    *result = polished;
    result->polish_iterations = o.iterations;
    result->iterations = o.iterations + first_iterations;
    copy_text(result->algorithm, sizeof(result->algorithm), first_algorithm + "+ipm");
    join_text(result->message, sizeof(result->message), first_message,
              "; polished by the interior point in " + std::to_string(o.iterations) +
                  " iterations" + (o.message.empty() ? "" : " (" + o.message + ")"));
    if (col_value) std::copy(x.begin(), x.end(), col_value);
    if (row_dual) std::copy(yd.begin(), yd.end(), row_dual);
    if (col_dual) std::copy(d.begin(), d.end(), col_dual);
    if (row_activity) std::copy(act.begin(), act.end(), row_activity);
    return SK_GPU_SUCCESS;
  });
}

}  // extern "C"
