// Reference-side adapter: the MI355X restarted-PDHG engine (libsk_gpu.so, include/sk_gpu.h)
// behind SANKHYA's PDHG engine seam.
//
// This file is compiled INTO the reference (as src/gpu/pdhg_mi355x.cpp, against its
// include/sankhya headers and this repository's include/sk_gpu.h); it is not part of
// libsk_gpu.so. run_lp_engine (src/core/solve.cpp:385-399) calls it in place of
// pdhg::solve_pdhg (include/sankhya/pdhg.hpp:15) when the backend decision is GPU, with the
// same arguments, and gets a Solution with the same contract:
//   - never throws: every failure, a C++ exception included, comes back as a status;
//   - Solution::allocate_for sizes (model.cpp:217-226); col_dual = sense (c_min + A^T y) and
//     row_dual = -sense y (pdhg.cpp:652-661); statuses numbered as SolveStatus
//     (model.hpp:62-80), which sk_gpu.h's SK_SOLVE_* values follow;
//   - dual_bound +-inf unless the answer meets the project standard (pdhg.cpp:710-716);
//   - recompute_quality before returning (model.cpp:228);
//   - algorithm "pdhg-gpu", which the audit reports as gpu_executed (verify.cpp:86-89).
#include <cstdio>
#include <exception>
#include <string>

#include "sankhya/logging.hpp"
#include "sankhya/model.hpp"
#include "sankhya/options.hpp"
#include "sk_gpu.h"

namespace sankhya::gpu {

namespace {

// Option values as text that sk_pdhg_set_option parses back to the same value. The typed
// getters are used because each key has one registered type (options.cpp:83-418):
// get_string asserts (and in a release build throws std::bad_variant_access) on a key whose
// value is not a string (options.cpp:604-615), and value_as_string prints doubles with six
// significant digits (options.cpp:640), which would round a tolerance such as 1.2345678e-5.
std::string double_text(double v) {
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  return buf;
}
std::string int_text(std::int64_t v) { return std::to_string(v); }
std::string bool_text(bool v) { return v ? "true" : "false"; }

// The options solve_pdhg reads (pdhg.cpp:304-315), in the registry's types.
int32_t forward_options(const Options& options, sk_pdhg_options* o) {
  struct Key {
    const char* name;
    char type;  // d double, i integer, b boolean
  };
  static const Key keys[] = {{"pdhg_tolerance", 'd'},       {"iteration_limit", 'i'},
                             {"time_limit", 'd'},           {"pdhg_restart", 'b'},
                             {"pdhg_stop_at_request", 'b'}, {"random_seed", 'i'}};
  for (const Key& k : keys) {
    const std::string v = k.type == 'd'   ? double_text(options.get_double(k.name))
                          : k.type == 'i' ? int_text(options.get_int(k.name))
                                          : bool_text(options.get_bool(k.name));
    const int32_t rc = sk_pdhg_set_option(o, k.name, v.c_str());
    if (rc != SK_GPU_SUCCESS) return rc;
  }
  return SK_GPU_SUCCESS;
}

// An LP handle freed on every path.
struct LpHandle {
  sk_lp* lp = nullptr;
  ~LpHandle() { sk_lp_free(lp); }
};

Solution failed(const Model& model, SolveStatus status, const std::string& message) {
  Solution s;
  s.status = status;
  s.algorithm = "pdhg-gpu";
  s.message = message;
  s.dual_bound = model.sense == ObjSense::kMaximize ? kInfinity : -kInfinity;
  return s;
}

std::string library_error(const char* what, int32_t rc) {
  return std::string("GPU PDHG failed: ") + what + ": " + sk_gpu_status_string(rc) + ": " +
         sk_gpu_last_error();
}

}  // namespace

Solution solve_pdhg_mi355x(const Model& model, const Options& options, Logger& logger) noexcept {
  try {
    if (const std::string problem = model.validate(); !problem.empty()) {
      return failed(model, SolveStatus::kModelError, problem);
    }
    const SparseMatrix& a = model.matrix;  // frozen CSC: column_starts, row_indices, values
    LpHandle h;
    int32_t rc = sk_lp_create(model.num_rows(), model.num_cols(),
                              model.sense == ObjSense::kMaximize ? -1 : 1, model.objective_offset,
                              a.column_starts().data(), a.row_indices().data(),
                              a.values().data(), model.col_cost.data(), model.col_lower.data(),
                              model.col_upper.data(), model.row_lower.data(),
                              model.row_upper.data(), &h.lp);
    if (rc != SK_GPU_SUCCESS) {
      // the library validates the layout as Model::validate does; a refusal is a model error
      return failed(model, rc == SK_GPU_INVALID_ARGUMENT ? SolveStatus::kModelError
                                                         : SolveStatus::kNotSolved,
                    library_error("sk_lp_create", rc));
    }
    sk_pdhg_options o;
    sk_pdhg_default_options(&o);
    if ((rc = forward_options(options, &o)) != SK_GPU_SUCCESS) {
      return failed(model, SolveStatus::kNotSolved, library_error("sk_pdhg_set_option", rc));
    }
    Solution solution;
    solution.allocate_for(model);
    solution.algorithm = "pdhg-gpu";
    sk_pdhg_result r{};
    rc = sk_pdhg_solve(h.lp, &o, &r, solution.col_value.data(), solution.row_dual.data(),
                       solution.col_dual.data(), solution.row_activity.data());
    if (rc != SK_GPU_SUCCESS) {
      // an explicit GPU request fails clearly rather than falling back (adaptive.cpp:144-158)
      return failed(model, SolveStatus::kNotSolved, library_error("sk_pdhg_solve", rc));
    }
    solution.status = static_cast<SolveStatus>(r.status);
    solution.objective = r.objective;
    solution.iterations = r.iterations;
    solution.solve_seconds = r.solve_seconds;
    solution.dual_bound = r.dual_bound;
    solution.message = r.message;
    solution.recompute_quality(model);
    logger.info("Status: {}   objective {:.10e}   iterations {}   restarts {}   time {:.3f}s",
                to_string(solution.status), solution.objective, solution.iterations, r.restarts,
                solution.solve_seconds);
    logger.info("Relative residuals: primal {:.3e}, dual {:.3e}, gap {:.3e}", r.rel_primal,
                r.rel_dual, r.rel_gap);
    if (!solution.message.empty()) logger.info("{}", solution.message);
    return solution;
  } catch (const std::exception& e) {
    try {
      return failed(model, SolveStatus::kNotSolved, std::string("GPU PDHG failed: ") + e.what());
    } catch (...) {
      return Solution{};  // kNotSolved with nothing attached
    }
  } catch (...) {
    try {
      return failed(model, SolveStatus::kNotSolved, "GPU PDHG failed: unknown exception");
    } catch (...) {
      return Solution{};
    }
  }
}

}  // namespace sankhya::gpu
