// extern "C" entry points of include/sk_gpu.h for models, options, single solves and batches.
// No C++ exception crosses this boundary: every entry point catches and reports a status, as
// the reference's C API does with its `guarded` firewall (src/api/c_api.cpp:129-140).
#include <algorithm>
#include <cctype>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sk_gpu.h"
#include "engine.hpp"
#include "lp.hpp"
#include "lpfile.hpp"
#include "presolve.hpp"
#include "quality.hpp"
#include "server.hpp"
#include "status.hpp"

#include "handles.hpp"

struct sk_batch {
  sk_pdhg_options options;
  std::unique_ptr<sk::Engine> engine;
  std::vector<const sk_lp*> lps;
  std::vector<sk::WarmStart> warm;  // per instance (engine sign), replayed into every engine
  int32_t tile_nnz = 0;  // the row-block capacity the uploaded batch runs with
  bool uploaded = false;
  bool solved = false;
};

struct sk_presolved {
  sk::presolve::Result result;
  double seconds = 0.0;
  std::vector<std::string> col_names, row_names;  // of the reduced LP, when the original had names
};

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

sk::EngineOptions to_engine(const sk_pdhg_options& o) {
  sk::EngineOptions e;
  e.tolerance = o.tolerance;
  e.iteration_limit = o.iteration_limit;
  e.time_limit = o.time_limit;
  e.restart = o.restart;
  e.stop_at_request = o.stop_at_request;
  e.absolute_clause = o.absolute_primal;
  e.random_seed = o.random_seed;
  e.precision = o.precision;
  e.steps_per_block = std::max(1, o.steps_per_block);
  e.eval_stride = std::max(1, std::min(o.eval_stride > 0 ? o.eval_stride : 20, e.steps_per_block));
  e.use_graph = o.use_graph;
  e.stage = o.spmv_stage;
  e.tile_nnz = o.tile_nnz;
  e.streams = o.streams;
  e.column_order = o.column_order == 1 || o.column_order == 2 ? o.column_order : 0;
  e.step_fusion = o.step_fusion & 3;
  e.column_slices = o.column_slices == 1 || o.column_slices == 2 ? o.column_slices : 0;
  return e;
}

// A solve, batch or server of this library runs the GPU engine: backend=cpu selects the
// reference's CPU engine and backend=hybrid is unsupported (adaptive.cpp:141-166).
int32_t refuse_backend(const sk_pdhg_options& o) {
  if (o.backend == SK_BACKEND_CPU) {
    sk::set_last_error("backend=cpu selects the CPU engine, which is not part of this library "
                       "(it is the GPU engine; see sk_select_backend)");
    return SK_GPU_INVALID_ARGUMENT;
  }
  if (o.backend == SK_BACKEND_HYBRID) {
    sk::set_last_error("backend=hybrid is unsupported in this version");
    return SK_GPU_INVALID_ARGUMENT;
  }
  if (o.backend != SK_BACKEND_AUTO && o.backend != SK_BACKEND_GPU) {
    sk::set_last_error("backend must be one of SK_BACKEND_*");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return SK_GPU_SUCCESS;
}

std::string lower(const char* s) {
  std::string r(s ? s : "");
  for (auto& c : r) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return r;
}

bool parse_bool(const std::string& v, int32_t* out) {
  if (v == "true" || v == "1" || v == "on" || v == "yes") return *out = 1, true;
  if (v == "false" || v == "0" || v == "off" || v == "no") return *out = 0, true;
  return false;
}

bool parse_double(const std::string& v, double* out) {
  if (v.empty()) return false;
  char* end = nullptr;
  const double d = std::strtod(v.c_str(), &end);
  if (*end != '\0' || std::isnan(d)) return false;
  *out = d;
  return true;
}

bool parse_int(const std::string& v, int64_t* out) {
  if (v.empty()) return false;
  char* end = nullptr;
  const long long d = std::strtoll(v.c_str(), &end, 10);
  if (*end != '\0') return false;
  *out = d;
  return true;
}

void fill_message(const sk::InstanceResult& r, const sk_pdhg_options& o, char* out, size_t size) {
  const sk::Residuals& f = r.final_residuals;
  if (r.status == SK_SOLVE_OPTIMAL) {
    std::snprintf(out, size,
                  "converged after %lld iterations and %lld restarts; absolute primal %.3e, "
                  "dual %.3e, relative gap %.3e",
                  static_cast<long long>(r.iterations), static_cast<long long>(r.restarts),
                  f.absolute_primal, f.absolute_dual, f.gap_as_verified);
  } else if (r.status == SK_SOLVE_FEASIBLE) {
    // pdhg.cpp:684-696, word for word
    std::snprintf(out, size,
                  "met the requested relative tolerance %.1e after %lld iterations, but NOT the "
                  "absolute standard this project verifies against (primal %.3e vs %.1e, dual "
                  "%.3e vs %.1e, relative gap %.3e vs %.1e). Reported as feasible, not optimal. %s",
                  o.tolerance, static_cast<long long>(r.iterations), f.absolute_primal, 1e-7,
                  f.absolute_dual, 1e-7, f.gap_as_verified, 1e-9,
                  o.stop_at_request ? "pdhg_stop_at_request is on, so this is the cheap answer "
                                      "that was asked for; turn it off to run on to the standard"
                                    : "Tighten --option pdhg_tolerance to close it");
  } else if (r.status == SK_SOLVE_NUMERICAL_ERROR) {
    std::snprintf(out, size, "a PDHG step produced a non-finite value after %lld iterations",
                  static_cast<long long>(r.iterations));
  } else {
    std::snprintf(out, size,
                  "stopped at relative primal %.3e, dual %.3e, gap %.3e after %lld iterations "
                  "and %lld restarts (target %.1e)",
                  f.primal, f.dual, f.gap, static_cast<long long>(r.iterations),
                  static_cast<long long>(r.restarts), o.tolerance);
  }
}

int32_t fill_result(const sk::HostLp& lp, const sk::InstanceResult& r, const sk_pdhg_options& o,
                    double seconds, sk_pdhg_result* res, double* col_value, double* row_dual,
                    double* col_dual, double* row_activity) {
  const auto n = static_cast<size_t>(lp.num_cols), m = static_cast<size_t>(lp.num_rows);
  if (col_value) std::copy(r.col_value.begin(), r.col_value.end(), col_value);
  if (col_dual) std::copy(r.col_dual.begin(), r.col_dual.end(), col_dual);
  if (row_dual) std::copy(r.row_dual.begin(), r.row_dual.end(), row_dual);
  if (row_activity) std::copy(r.row_activity.begin(), r.row_activity.end(), row_activity);
  if (!res) return SK_GPU_SUCCESS;
  std::memset(res, 0, sizeof(*res));
  res->status = r.status;
  res->converged = r.converged;
  res->iterations = r.iterations;
  res->restarts = r.restarts;
  res->evaluations = r.evaluations;
  res->attempts = r.attempts;
  res->dual_bound = r.dual_bound;
  res->spectral_norm = r.spectral_norm;
  res->solve_seconds = seconds;
  const sk::Residuals& f = r.final_residuals;
  res->rel_primal = f.primal;
  res->rel_dual = f.dual;
  res->rel_gap = f.gap;
  res->abs_primal = f.absolute_primal;
  res->abs_dual = f.absolute_dual;
  res->gap_as_verified = f.gap_as_verified;
  res->complementarity = f.complementarity;
  res->primal_objective = f.primal_objective;
  res->dual_objective = f.dual_objective;
  // objective in the model's own sense, with its offset (model.cpp:99-106)
  double objective = lp.objective_offset;
  for (size_t j = 0; j < n; ++j) objective += lp.cost[j] * r.col_value[j];
  res->objective = objective;
  res->absolute_gap = std::fabs(objective - r.dual_bound);
  res->relative_gap = res->absolute_gap / std::max(1.0, std::fabs(objective));
  if (o.compute_quality && r.col_value.size() == n && r.row_dual.size() == m) {
    sk::Quality q;
    sk::recompute_quality(lp, r.col_value.data(), r.row_dual.data(), r.col_dual.data(),
                          r.dual_bound, &q);
    res->objective = q.objective;
    res->primal_infeasibility = q.primal_infeasibility;
    res->primal_infeasibility_scaled = q.primal_infeasibility_scaled;
    res->dual_infeasibility = q.dual_infeasibility;
    res->dual_infeasibility_scaled = q.dual_infeasibility_scaled;
    res->complementarity_violation = q.complementarity_violation;
    res->absolute_gap = q.absolute_gap;
    res->relative_gap = q.relative_gap;
  }
  res->primal_weight = r.omega;
  res->step_size = r.eta;
  std::snprintf(res->algorithm, sizeof(res->algorithm), "pdhg-gpu");
  fill_message(r, o, res->message, sizeof(res->message));
  return SK_GPU_SUCCESS;
}

// solve.cpp:423-443 around the engine: presolve, the engine on the reduced LP, postsolve, and
// the answer measured on the original LP (fill_result recomputes its objective and quality
// there; the residuals reported are the engine's, on the reduced LP it solved).
int32_t solve_presolved(const sk::HostLp& lp, const sk_pdhg_options& o, sk_pdhg_result* res,
                        double* col_value, double* row_dual, double* col_dual, double* row_activity) {
  const auto t0 = std::chrono::steady_clock::now();
  const sk::presolve::Result P = sk::presolve::run(lp, {}, 1e-7);
  const auto n = static_cast<size_t>(lp.num_cols), m = static_cast<size_t>(lp.num_rows);
  auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  if (P.proved_infeasible) {  // solve.cpp:425-442: no engine runs, and there is no certificate
    if (col_value) std::fill(col_value, col_value + n, 0.0);
    if (col_dual) std::fill(col_dual, col_dual + n, 0.0);
    if (row_dual) std::fill(row_dual, row_dual + m, 0.0);
    if (row_activity) std::fill(row_activity, row_activity + m, 0.0);
    if (!res) return SK_GPU_SUCCESS;
    std::memset(res, 0, sizeof(*res));
    res->status = SK_SOLVE_INFEASIBLE;
    res->dual_bound = std::numeric_limits<double>::infinity();
    res->solve_seconds = elapsed();
    std::snprintf(res->algorithm, sizeof(res->algorithm), "presolve");
    std::snprintf(res->message, sizeof(res->message),
                  "%s; proved by presolve, which carries no Farkas certificate - re-run with "
                  "--option presolve=false for one", P.message.c_str());
    return SK_GPU_SUCCESS;
  }
  const sk::HostLp& red = P.reduced;
  sk::InstanceResult rr;
  if (red.num_cols > 0) {
    auto engine = sk::make_engine(to_engine(o));
    engine->add(&red);
    int32_t st = engine->upload();
    if (!st) st = engine->solve();
    if (!st) st = engine->result(0, &rr);
    if (st) return st;
  } else {
    // Presolve removed every column: what is left is a set of rows with no entries, which
    // hold exactly when zero lies within their bounds. Nothing remains to iterate on.
    bool holds = true;
    for (int32_t i = 0; i < red.num_rows; ++i)
      holds = holds && !(red.row_lower[static_cast<size_t>(i)] > 1e-7) && !(red.row_upper[static_cast<size_t>(i)] < -1e-7);
    rr.status = holds ? SK_SOLVE_OPTIMAL : SK_SOLVE_INFEASIBLE;
    rr.converged = holds ? 1 : 0;
    rr.dual_bound = holds ? red.objective_offset : std::numeric_limits<double>::infinity();
    rr.row_dual.assign(static_cast<size_t>(red.num_rows), 0.0);
  }
  sk::InstanceResult full = rr;
  full.col_value.assign(n, 0.0);
  full.row_dual.assign(m, 0.0);
  full.col_dual.assign(n, 0.0);
  sk::presolve::postsolve(P, lp, rr.col_value.data(), rr.row_dual.data(), rr.col_dual.data(),
                          full.col_value.data(), full.row_dual.data(), full.col_dual.data());
  full.row_activity.assign(m, 0.0);
  for (size_t j = 0; j < n; ++j)
    for (int32_t p = lp.col_starts[j]; p < lp.col_starts[j + 1]; ++p)
      full.row_activity[static_cast<size_t>(lp.row_indices[static_cast<size_t>(p)])] += lp.values[static_cast<size_t>(p)] * full.col_value[j];
  const int32_t st = fill_result(lp, full, o, elapsed(), res, col_value, row_dual, col_dual, row_activity);
  if (st || !res) return st;
  char head[96];
  std::snprintf(head, sizeof head, "presolve %d x %d -> %d x %d; ", lp.num_rows, lp.num_cols,
                red.num_rows, red.num_cols);
  const std::string msg = std::string(head) + res->message;
  std::snprintf(res->message, sizeof(res->message), "%s", msg.c_str());
  return SK_GPU_SUCCESS;
}

}  // namespace

extern "C" {

int32_t sk_lp_create(int32_t num_rows, int32_t num_cols, int32_t sense, double objective_offset,
                     const int32_t* col_starts, const int32_t* row_indices, const double* values,
                     const double* col_cost, const double* col_lower, const double* col_upper,
                     const double* row_lower, const double* row_upper, sk_lp** out) {
  return guarded([&]() -> int32_t {
    if (!out) return SK_GPU_INVALID_ARGUMENT;
    *out = nullptr;
    if (num_rows < 0 || num_cols < 0 || (sense != 1 && sense != -1) || !col_starts ||
        (num_cols > 0 && (!col_cost || !col_lower || !col_upper)) ||
        (num_rows > 0 && (!row_lower || !row_upper))) {
      sk::set_last_error("sk_lp_create: bad dimensions, sense or null array");
      return SK_GPU_INVALID_ARGUMENT;
    }
    auto h = std::make_unique<sk_lp>();
    sk::HostLp& lp = h->lp;
    lp.num_rows = num_rows;
    lp.num_cols = num_cols;
    lp.sense = sense;
    lp.objective_offset = objective_offset;
    lp.col_starts.assign(col_starts, col_starts + num_cols + 1);
    const int64_t nnz = lp.col_starts.back();
    if (nnz < 0 || (nnz > 0 && (!row_indices || !values))) {
      sk::set_last_error("sk_lp_create: bad nonzero count or null matrix arrays");
      return SK_GPU_INVALID_ARGUMENT;
    }
    lp.row_indices.assign(row_indices, row_indices + nnz);
    lp.values.assign(values, values + nnz);
    lp.cost.assign(col_cost, col_cost + num_cols);
    lp.col_lower.assign(col_lower, col_lower + num_cols);
    lp.col_upper.assign(col_upper, col_upper + num_cols);
    lp.row_lower.assign(row_lower, row_lower + num_rows);
    lp.row_upper.assign(row_upper, row_upper + num_rows);
    const std::string problem = lp.validate();
    if (!problem.empty()) {
      sk::set_last_error(problem);
      return SK_GPU_INVALID_ARGUMENT;
    }
    *out = h.release();
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_gen_preset(int32_t config, uint64_t instance, sk_gen_params* out) {
  if (!out) return SK_GPU_INVALID_ARGUMENT;
  sk::GenParams p;
  if (!sk::preset_params(config, instance, &p)) {
    sk::set_last_error("unknown workload; expected 1..5");
    return SK_GPU_INVALID_ARGUMENT;
  }
  out->seed = p.seed;
  out->num_rows = p.num_rows;
  out->num_cols = p.num_cols;
  out->num_equality_rows = p.num_equality_rows;
  out->degree_kind = p.degree_kind;
  out->degree_lo = p.degree_lo;
  out->degree_hi = p.degree_hi;
  out->power_exponent = p.power_exponent;
  out->target_nnz = p.target_nnz;
  out->value_kind = p.value_kind;
  out->value_lo = p.value_lo;
  out->value_hi = p.value_hi;
  out->negative_probability = p.negative_probability;
  out->upper_lo = p.upper_lo;
  out->upper_hi = p.upper_hi;
  out->slack_lo = p.slack_lo;
  out->slack_hi = p.slack_hi;
  out->cost_lo = p.cost_lo;
  out->cost_hi = p.cost_hi;
  out->fix_empty_rows = p.fix_empty_rows;
  return SK_GPU_SUCCESS;
}

int32_t sk_lp_generate_params(const sk_gen_params* in, sk_lp** out) {
  return guarded([&]() -> int32_t {
    if (!in || !out || in->num_rows < 0 || in->num_cols < 0) return SK_GPU_INVALID_ARGUMENT;
    sk::GenParams p;
    p.seed = in->seed;
    p.num_rows = in->num_rows;
    p.num_cols = in->num_cols;
    p.num_equality_rows = in->num_equality_rows;
    p.degree_kind = in->degree_kind;
    p.degree_lo = in->degree_lo;
    p.degree_hi = in->degree_hi;
    p.power_exponent = in->power_exponent;
    p.target_nnz = in->target_nnz;
    p.value_kind = in->value_kind;
    p.value_lo = in->value_lo;
    p.value_hi = in->value_hi;
    p.negative_probability = in->negative_probability;
    p.upper_lo = in->upper_lo;
    p.upper_hi = in->upper_hi;
    p.slack_lo = in->slack_lo;
    p.slack_hi = in->slack_hi;
    p.cost_lo = in->cost_lo;
    p.cost_hi = in->cost_hi;
    p.fix_empty_rows = in->fix_empty_rows;
    auto h = std::make_unique<sk_lp>();
    h->lp = sk::generate_lp(p);
    *out = h.release();
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_lp_generate(int32_t config, uint64_t instance, sk_lp** out) {
  sk_gen_params p;
  const int32_t st = sk_gen_preset(config, instance, &p);
  if (st) return st;
  return sk_lp_generate_params(&p, out);
}

int32_t sk_lp_dims(const sk_lp* lp, int32_t* num_rows, int32_t* num_cols, int64_t* nnz) {
  if (!lp) return SK_GPU_INVALID_ARGUMENT;
  if (num_rows) *num_rows = lp->lp.num_rows;
  if (num_cols) *num_cols = lp->lp.num_cols;
  if (nnz) *nnz = lp->lp.nnz();
  return SK_GPU_SUCCESS;
}

int32_t sk_lp_export(const sk_lp* h, int32_t* sense, double* offset, int32_t* col_starts,
                     int32_t* row_indices, double* values, double* col_cost, double* col_lower,
                     double* col_upper, double* row_lower, double* row_upper) {
  if (!h) return SK_GPU_INVALID_ARGUMENT;
  const sk::HostLp& lp = h->lp;
  if (sense) *sense = lp.sense;
  if (offset) *offset = lp.objective_offset;
  auto put = [](const auto& v, auto* dst) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  put(lp.col_starts, col_starts);
  put(lp.row_indices, row_indices);
  put(lp.values, values);
  put(lp.cost, col_cost);
  put(lp.col_lower, col_lower);
  put(lp.col_upper, col_upper);
  put(lp.row_lower, row_lower);
  put(lp.row_upper, row_upper);
  return SK_GPU_SUCCESS;
}

void sk_lp_free(sk_lp* lp) { delete lp; }

int32_t sk_lp_read_mps(const char* path, sk_lp** out, int32_t* integer_columns) {
  if (!path || !out) return SK_GPU_INVALID_ARGUMENT;
  *out = nullptr;
  return guarded([&]() -> int32_t {
    auto h = std::make_unique<sk_lp>();
    std::string error;
    if (!sk::read_mps(path, &h->lp, &h->file, &error)) {
      sk::set_last_error(error);
      return SK_GPU_INVALID_ARGUMENT;
    }
    if (integer_columns) *integer_columns = h->file.integer_columns;
    *out = h.release();
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_lp_write_mps(const sk_lp* lp, const char* path) {
  if (!lp || !path) return SK_GPU_INVALID_ARGUMENT;
  return guarded([&]() -> int32_t {
    std::string error;
    if (!sk::write_mps(lp->lp, &lp->file, path, &error)) {
      sk::set_last_error(error);
      return SK_GPU_INVALID_ARGUMENT;
    }
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_lp_save(const sk_lp* lp, const char* path) {
  if (!lp || !path) return SK_GPU_INVALID_ARGUMENT;
  return guarded([&]() -> int32_t {
    std::string error;
    if (!sk::save_lp(lp->lp, path, &error)) {
      sk::set_last_error(error);
      return SK_GPU_INVALID_ARGUMENT;
    }
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_lp_load(const char* path, sk_lp** out) {
  if (!path || !out) return SK_GPU_INVALID_ARGUMENT;
  *out = nullptr;
  return guarded([&]() -> int32_t {
    auto h = std::make_unique<sk_lp>();
    std::string error;
    if (!sk::load_lp(path, &h->lp, &error)) {
      sk::set_last_error(error);
      return SK_GPU_INVALID_ARGUMENT;
    }
    *out = h.release();
    return SK_GPU_SUCCESS;
  });
}

const char* sk_lp_name(const sk_lp* lp, int32_t kind, int32_t index) {
  if (!lp) return nullptr;
  if (kind == 2) return lp->file.name.c_str();
  const auto& names = kind == 0 ? lp->file.row_names : lp->file.col_names;
  if ((kind != 0 && kind != 1) || index < 0 || static_cast<size_t>(index) >= names.size()) return nullptr;
  return names[static_cast<size_t>(index)].c_str();
}

void sk_pdhg_default_options(sk_pdhg_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->tolerance = 1e-4;
  o->iteration_limit = -1;
  o->time_limit = DBL_MAX;
  o->restart = 1;
  o->stop_at_request = 0;
  o->absolute_primal = 1;
  o->random_seed = 0;
  o->precision = 0;
  o->steps_per_block = 40;
  o->use_graph = 1;
  o->spmv_stage = -1;
  o->tile_nnz = 0;
  o->compute_quality = 1;
  o->eval_stride = 20;
  o->streams = 0;
  o->column_order = 0;
  o->column_slices = 0;
}

int32_t sk_pdhg_set_option(sk_pdhg_options* o, const char* key_in, const char* value_in) {
  if (!o || !key_in || !value_in) return SK_GPU_INVALID_ARGUMENT;
  const std::string key = lower(key_in);
  const std::string v = lower(value_in);
  int32_t b = 0;
  double d = 0.0;
  int64_t i = 0;
  auto bad = [&](const char* what) {
    sk::set_last_error("option " + key + ": " + what + " '" + value_in + "'");
    return SK_GPU_INVALID_ARGUMENT;
  };
  if (key == "algorithm") {
    if (v != "pdhg") return bad("this library runs only algorithm=pdhg, not");
    return SK_GPU_SUCCESS;
  }
  if (key == "backend") {
    // parse_backend_mode (adaptive.hpp:43-62); the registry's values (options.cpp:228-236)
    if (v == "auto") o->backend = SK_BACKEND_AUTO;
    else if (v == "cpu") o->backend = SK_BACKEND_CPU;
    else if (v == "gpu" || v == "cuda") o->backend = SK_BACKEND_GPU;
    else if (v == "hybrid") o->backend = SK_BACKEND_HYBRID;
    else return bad("expected auto, cpu, gpu or hybrid, got");
    return SK_GPU_SUCCESS;
  }
  if (key == "gpu") {
    if (!parse_bool(v, &b)) return bad("expected a boolean, got");
    o->gpu = b;
    return SK_GPU_SUCCESS;
  }
  if (key == "presolve") {
    if (!parse_bool(v, &b)) return bad("expected a boolean, got");
    o->presolve = b;
    return SK_GPU_SUCCESS;
  }
  if (key == "pdhg_polish") {
    // read by the caller, which then hands the answer to sk_pdhg_polish (solve.cpp:184-243);
    // the first-order solve itself is the same either way
    if (!parse_bool(v, &b)) return bad("expected a boolean, got");
    return SK_GPU_SUCCESS;
  }
  if (key == "pdhg_tolerance") {
    if (!parse_double(v, &d) || d < 1e-14 || d > 1e-1) return bad("expected a number in [1e-14, 0.1], got");
    o->tolerance = d;
    return SK_GPU_SUCCESS;
  }
  if (key == "iteration_limit") {
    if (!parse_int(v, &i) || i < -1) return bad("expected an integer >= -1, got");
    o->iteration_limit = i;
    return SK_GPU_SUCCESS;
  }
  if (key == "time_limit") {
    if (!parse_double(v, &d) || d < 0.0) return bad("expected a non-negative number, got");
    o->time_limit = d;
    return SK_GPU_SUCCESS;
  }
  if (key == "random_seed") {
    if (!parse_int(v, &i) || i < 0) return bad("expected a non-negative integer, got");
    o->random_seed = i;
    return SK_GPU_SUCCESS;
  }
  if (key == "pdhg_restart" || key == "pdhg_stop_at_request" || key == "pdhg_absolute_primal") {
    if (!parse_bool(v, &b)) return bad("expected a boolean, got");
    if (key == "pdhg_restart") o->restart = b;
    else if (key == "pdhg_stop_at_request") o->stop_at_request = b;
    else o->absolute_primal = b;
    return SK_GPU_SUCCESS;
  }
  if (key == "pdhg_precision") {
    if (v == "fp32") o->precision = 0;
    else if (v == "fp64") o->precision = 1;
    else return bad("expected fp32 or fp64, got");
    return SK_GPU_SUCCESS;
  }
  if (key == "pdhg_steps_per_block") {
    if (!parse_int(v, &i) || i < 1 || i > 400) return bad("expected an integer in [1, 400], got");
    o->steps_per_block = static_cast<int32_t>(i);
    return SK_GPU_SUCCESS;
  }
  sk::set_last_error("unknown option '" + std::string(key_in) + "'");
  return SK_GPU_INVALID_ARGUMENT;
}

int32_t sk_pdhg_solve(const sk_lp* lp, const sk_pdhg_options* options, sk_pdhg_result* result,
                      double* col_value, double* row_dual, double* col_dual, double* row_activity) {
  return guarded([&]() -> int32_t {
    if (!lp) return SK_GPU_INVALID_ARGUMENT;
    sk_pdhg_options o;
    if (options) o = *options;
    else sk_pdhg_default_options(&o);
    if (const int32_t refused = refuse_backend(o)) return refused;
    if (o.presolve) return solve_presolved(lp->lp, o, result, col_value, row_dual, col_dual, row_activity);
    const auto t0 = std::chrono::steady_clock::now();
    auto engine = sk::make_engine(to_engine(o));
    engine->add(&lp->lp);
    int32_t st = engine->upload();
    if (st) return st;
    if ((st = engine->solve())) return st;
    sk::InstanceResult r;
    if ((st = engine->result(0, &r))) return st;
    const double seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return fill_result(lp->lp, r, o, seconds, result, col_value, row_dual, col_dual, row_activity);
  });
}

int32_t sk_pdhg_refine(const sk_lp* lp, const sk_pdhg_options* options, const sk_pdhg_result* first,
                       const double* first_col_value, const double* first_row_dual,
                       sk_pdhg_result* result, double* col_value, double* row_dual,
                       double* col_dual, double* row_activity) {
  return guarded([&]() -> int32_t {
    if (!lp || !first || !first_col_value || !first_row_dual) return SK_GPU_INVALID_ARGUMENT;
    sk_pdhg_options o;
    if (options) o = *options;
    else sk_pdhg_default_options(&o);
    if (const int32_t refused = refuse_backend(o)) return refused;
    o.precision = 1;
    const sk::HostLp& m = lp->lp;
    // the engine's dual sign: row_dual = -sense y (pdhg.cpp:657-661)
    std::vector<double> y(static_cast<size_t>(m.num_rows));
    for (int32_t i = 0; i < m.num_rows; ++i) y[static_cast<size_t>(i)] = -static_cast<double>(m.sense) * first_row_dual[i];
    const auto t0 = std::chrono::steady_clock::now();
    auto engine = sk::make_engine(to_engine(o));
    engine->add(&m);
    int32_t st = engine->warm_start(0, first_col_value, y.data(), first->primal_weight, first->step_size);
    if (!st) st = engine->upload();
    if (!st) st = engine->solve();
    sk::InstanceResult r;
    if (!st) st = engine->result(0, &r);
    if (st) return st;
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st = fill_result(m, r, o, seconds, result, col_value, row_dual, col_dual, row_activity);
    if (st || !result) return st;
    result->polish_iterations = result->iterations;
    result->iterations += first->iterations;
    result->solve_seconds += first->solve_seconds;
    std::snprintf(result->algorithm, sizeof(result->algorithm), "pdhg-gpu+fp64");
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_presolve(const sk_lp* lp, const uint8_t* integer_columns, double feasibility,
                    sk_presolved** out) {
  return guarded([&]() -> int32_t {
    if (!lp || !out) return SK_GPU_INVALID_ARGUMENT;
    *out = nullptr;
    const std::string problem = lp->lp.validate();
    if (!problem.empty()) {
      sk::set_last_error("sk_presolve: " + problem);
      return SK_GPU_INVALID_ARGUMENT;
    }
    if (!(feasibility > 0.0)) feasibility = 1e-7;
    if (feasibility < 1e-12 || feasibility > 1e-3) {  // the option's range (options.cpp:293-299)
      sk::set_last_error("sk_presolve: primal_feasibility_tolerance outside [1e-12, 1e-3]");
      return SK_GPU_INVALID_ARGUMENT;
    }
    std::vector<uint8_t> integer;
    if (integer_columns) integer.assign(integer_columns, integer_columns + lp->lp.num_cols);
    auto h = std::make_unique<sk_presolved>();
    const auto t0 = std::chrono::steady_clock::now();
    h->result = sk::presolve::run(lp->lp, integer, feasibility);
    h->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const sk::MpsInfo& f = lp->file;
    if (f.col_names.size() == static_cast<size_t>(lp->lp.num_cols))
      for (const int32_t j : h->result.col_to_original) h->col_names.push_back(f.col_names[static_cast<size_t>(j)]);
    if (f.row_names.size() == static_cast<size_t>(lp->lp.num_rows))
      for (const int32_t i : h->result.row_to_original) h->row_names.push_back(f.row_names[static_cast<size_t>(i)]);
    *out = h.release();
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_presolve_get_info(const sk_presolved* p, sk_presolve_info* info) {
  if (!p || !info) return SK_GPU_INVALID_ARGUMENT;
  const sk::presolve::Result& r = p->result;
  std::memset(info, 0, sizeof(*info));
  info->original_rows = r.original_rows;
  info->original_cols = r.original_cols;
  info->original_nonzeros = r.original_nonzeros;
  info->reduced_rows = r.reduced.num_rows;
  info->reduced_cols = r.reduced.num_cols;
  info->reduced_nonzeros = r.proved_infeasible ? 0 : r.reduced.nnz();
  info->records = static_cast<int32_t>(r.records.size());
  info->passes = r.passes;
  info->proved_infeasible = r.proved_infeasible ? 1 : 0;
  for (const sk::presolve::Record& rec : r.records) ++info->by_kind[static_cast<int>(rec.kind)];
  info->seconds = p->seconds;
  return SK_GPU_SUCCESS;
}

const char* sk_presolve_message(const sk_presolved* p) { return p ? p->result.message.c_str() : ""; }

int32_t sk_presolve_reduced_lp(const sk_presolved* p, sk_lp** out) {
  return guarded([&]() -> int32_t {
    if (!p || !out) return SK_GPU_INVALID_ARGUMENT;
    *out = nullptr;
    if (p->result.proved_infeasible) {
      sk::set_last_error("sk_presolve_reduced_lp: presolve proved the LP infeasible; there is no reduced LP");
      return SK_GPU_INVALID_STATE;
    }
    auto h = std::make_unique<sk_lp>();
    h->lp = p->result.reduced;
    h->file.col_names = p->col_names;
    h->file.row_names = p->row_names;
    *out = h.release();
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_presolve_maps(const sk_presolved* p, int32_t* col_to_original, int32_t* row_to_original) {
  if (!p) return SK_GPU_INVALID_ARGUMENT;
  const sk::presolve::Result& r = p->result;
  if (col_to_original) std::copy(r.col_to_original.begin(), r.col_to_original.end(), col_to_original);
  if (row_to_original) std::copy(r.row_to_original.begin(), r.row_to_original.end(), row_to_original);
  return SK_GPU_SUCCESS;
}

int32_t sk_presolve_get_record(const sk_presolved* p, int32_t k, sk_presolve_record* out) {
  if (!p || !out || k < 0 || static_cast<size_t>(k) >= p->result.records.size()) return SK_GPU_INVALID_ARGUMENT;
  const sk::presolve::Record& rec = p->result.records[static_cast<size_t>(k)];
  out->kind = static_cast<int32_t>(rec.kind);
  out->index = rec.index;
  out->value = rec.value;
  out->coefficient = rec.coefficient;
  out->row_lower = rec.row_lower;
  out->row_upper = rec.row_upper;
  out->column = rec.column;
  out->partner_column = rec.partner_column;
  out->substituted_rhs = rec.substituted_rhs;
  out->partner_coefficient = rec.partner_coefficient;
  out->eliminated_cost = rec.eliminated_cost;
  return SK_GPU_SUCCESS;
}

int32_t sk_postsolve(const sk_presolved* p, const sk_lp* original, const double* col_value,
                     const double* row_dual, const double* col_dual, double* out_col_value,
                     double* out_row_dual, double* out_col_dual) {
  return guarded([&]() -> int32_t {
    if (!p || !original || !out_col_value || !out_row_dual || !out_col_dual) return SK_GPU_INVALID_ARGUMENT;
    const sk::presolve::Result& r = p->result;
    if (r.proved_infeasible) {
      sk::set_last_error("sk_postsolve: presolve proved the LP infeasible; there is nothing to postsolve");
      return SK_GPU_INVALID_STATE;
    }
    if (original->lp.num_rows != r.original_rows || original->lp.num_cols != r.original_cols ||
        original->lp.nnz() != r.original_nonzeros) {
      sk::set_last_error("sk_postsolve: `original` is not the LP this presolve reduced");
      return SK_GPU_INVALID_ARGUMENT;
    }
    if ((r.reduced.num_cols > 0 && (!col_value || !col_dual)) || (r.reduced.num_rows > 0 && !row_dual))
      return SK_GPU_INVALID_ARGUMENT;
    sk::presolve::postsolve(r, original->lp, col_value, row_dual, col_dual, out_col_value,
                            out_row_dual, out_col_dual);
    return SK_GPU_SUCCESS;
  });
}

void sk_presolve_free(sk_presolved* p) { delete p; }

int32_t sk_batch_warm_start(sk_batch* b, int32_t index, const double* col_value,
                            const double* row_dual, double primal_weight, double step_size) {
  if (!b || !col_value || !row_dual || index < 0 || index >= static_cast<int32_t>(b->lps.size())) {
    sk::set_last_error("sk_batch_warm_start: no such instance, or a null start");
    return SK_GPU_INVALID_ARGUMENT;
  }
  if (b->uploaded) {
    sk::set_last_error("sk_batch_warm_start after sk_batch_upload");
    return SK_GPU_INVALID_STATE;
  }
  return guarded([&]() -> int32_t {
    const sk::HostLp& m = b->lps[static_cast<size_t>(index)]->lp;
    std::vector<double> y(static_cast<size_t>(m.num_rows));
    for (int32_t i = 0; i < m.num_rows; ++i) y[static_cast<size_t>(i)] = -static_cast<double>(m.sense) * row_dual[i];
    const int32_t st = b->engine->warm_start(index, col_value, y.data(), primal_weight, step_size);
    if (st) return st;
    if (b->warm.size() < b->lps.size()) b->warm.resize(b->lps.size());
    sk::WarmStart& w = b->warm[static_cast<size_t>(index)];
    w.set = true;
    w.x.assign(col_value, col_value + m.num_cols);
    w.y = std::move(y);
    w.omega = primal_weight;
    w.eta = step_size;
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_batch_create(const sk_pdhg_options* options, sk_batch** out) {
  return guarded([&]() -> int32_t {
    if (!out) return SK_GPU_INVALID_ARGUMENT;
    auto b = std::make_unique<sk_batch>();
    if (options) b->options = *options;
    else sk_pdhg_default_options(&b->options);
    if (const int32_t refused = refuse_backend(b->options)) return refused;
    b->engine = sk::make_engine(to_engine(b->options));
    *out = b.release();
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_batch_add(sk_batch* b, const sk_lp* lp, int32_t* index) {
  if (!b || !lp) return SK_GPU_INVALID_ARGUMENT;
  if (b->uploaded) {
    sk::set_last_error("sk_batch_add after sk_batch_upload");
    return SK_GPU_INVALID_STATE;
  }
  return guarded([&]() -> int32_t {
    const int32_t k = b->engine->add(&lp->lp);
    b->lps.push_back(lp);
    if (index) *index = k;
    return SK_GPU_SUCCESS;
  });
}

namespace {

// tile_nnz = -1: the batch is uploaded with each candidate capacity and the two SpMV step
// kernels are timed with every instance stepping, as the solve schedules them (all sub-batch
// streams at once): the median of three timings of 10 launches each. The default capacity is
// timed first and another one replaces it only when it is more than 3% faster, since
// concurrent launches time with a few percent of noise. The engine kept is the chosen one;
// the others are freed. A capacity changes no row sum, only which tiles share a workgroup and
// a partial sum.
int32_t autotune_upload(sk_batch* b) {
  static const int32_t kCandidates[] = {sk::kDefaultTileNnz, 1536, 2048, 3072, 4096};
  constexpr int kReps = 10, kTimings = 3;
  constexpr double kMargin = 0.97;
  std::unique_ptr<sk::Engine> best;
  double best_ms = INFINITY;
  for (const int32_t cap : kCandidates) {
    sk_pdhg_options o = b->options;
    o.tile_nnz = cap;
    auto e = sk::make_engine(to_engine(o));
    for (const sk_lp* lp : b->lps) e->add(&lp->lp);
    int32_t st = SK_GPU_SUCCESS;
    for (size_t k = 0; k < b->warm.size() && !st; ++k) {
      const sk::WarmStart& w = b->warm[k];
      if (w.set) st = e->warm_start(static_cast<int32_t>(k), w.x.data(), w.y.data(), w.omega, w.eta);
    }
    if (!st) st = e->upload();
    double t[kTimings];
    for (int k = 0; k < kTimings && !st; ++k) {
      double dual_ms = 0.0, aty_ms = 0.0;
      st = e->time_kernel(1, kReps, &dual_ms);
      if (!st) st = e->time_kernel(2, kReps, &aty_ms);
      t[k] = dual_ms + aty_ms;
    }
    if (st) return st;
    std::sort(t, t + kTimings);
    const double ms = t[kTimings / 2];
    if (!best || ms < kMargin * best_ms) {
      best_ms = ms;
      best = std::move(e);
      b->tile_nnz = cap;
    }
  }
  b->engine = std::move(best);
  return SK_GPU_SUCCESS;
}

}  // namespace

int32_t sk_batch_upload(sk_batch* b) {
  if (!b) return SK_GPU_INVALID_ARGUMENT;
  return guarded([&]() -> int32_t {
    int32_t st;
    if (b->options.tile_nnz == -1 && !b->lps.empty()) {
      st = autotune_upload(b);
    } else {
      st = b->engine->upload();
      b->tile_nnz = b->engine->tile_capacity();
    }
    if (!st) b->uploaded = true;
    return st;
  });
}

int32_t sk_batch_solve(sk_batch* b) {
  if (!b) return SK_GPU_INVALID_ARGUMENT;
  if (!b->uploaded) {
    sk::set_last_error("sk_batch_solve before sk_batch_upload");
    return SK_GPU_INVALID_STATE;
  }
  return guarded([&]() -> int32_t {
    const int32_t st = b->engine->solve();
    b->solved = st == SK_GPU_SUCCESS;
    return st;
  });
}

int32_t sk_batch_result(sk_batch* b, int32_t index, sk_pdhg_result* result, double* col_value,
                        double* row_dual, double* col_dual, double* row_activity) {
  if (!b || !b->solved || index < 0 || index >= static_cast<int32_t>(b->lps.size())) {
    sk::set_last_error("sk_batch_result: no solved instance with that index");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int32_t {
    sk::InstanceResult r;
    const int32_t st = b->engine->result(index, &r);
    if (st) return st;
    const auto& t = b->engine->timings();
    const double seconds = (t.setup_ms + t.loop_ms + t.report_ms) * 1e-3;
    return fill_result(b->lps[static_cast<size_t>(index)]->lp, r, b->options, seconds, result,
                       col_value, row_dual, col_dual, row_activity);
  });
}

int32_t sk_batch_stats_get(const sk_batch* b, sk_batch_stats* out) {
  if (!b || !out) return SK_GPU_INVALID_ARGUMENT;
  const auto& t = b->engine->timings();
  out->setup_ms = t.setup_ms;
  out->loop_ms = t.loop_ms;
  out->report_ms = t.report_ms;
  out->graph_launches = t.graph_launches;
  out->total_rows = b->engine->total_rows();
  out->total_cols = b->engine->total_cols();
  out->total_nnz = b->engine->total_nnz();
  out->instances = b->engine->size();
  out->spmv_stage = b->engine->stage_choice();
  out->streams = b->engine->streams();
  out->tile_nnz = b->tile_nnz;
  out->fused = b->engine->fused();
  out->column_slices = b->engine->column_slices();
  return SK_GPU_SUCCESS;
}

int32_t sk_batch_time_kernel(sk_batch* b, int32_t kernel, int32_t reps, double* mean_ms,
                             double* bytes) {
  if (!b || !mean_ms || kernel < 0 || kernel > 5 || !b->uploaded) return SK_GPU_INVALID_ARGUMENT;
  return guarded([&]() -> int32_t {
    const int32_t st = b->engine->time_kernel(kernel, reps, mean_ms);
    if (bytes) *bytes = b->engine->kernel_bytes(kernel);
    return st;
  });
}

int32_t sk_batch_probe_spmv(sk_batch* b, int32_t index, int32_t which, const double* in,
                            double* out) {
  if (!b || !b->solved || !in || !out) {
    sk::set_last_error("sk_batch_probe_spmv: needs a solved batch and both arrays");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int32_t { return b->engine->probe_spmv(index, which, in, out); });
}

int32_t sk_batch_probe_scaling(sk_batch* b, int32_t index, double* row_scale, double* col_scale,
                               double* scaled_values) {
  if (!b || !b->solved) {
    sk::set_last_error("sk_batch_probe_scaling: needs a solved batch");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int32_t {
    return b->engine->probe_scaling(index, row_scale, col_scale, scaled_values);
  });
}

int32_t sk_batch_column_order(const sk_batch* b, int32_t index, int32_t* device_to_lp) {
  if (!b || !b->uploaded || !device_to_lp) {
    sk::set_last_error("sk_batch_column_order: needs an uploaded batch and an array");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int32_t { return b->engine->column_order(index, device_to_lp); });
}

void sk_batch_destroy(sk_batch* b) { delete b; }

}  // extern "C"

// A completed request as the C ABI hands it out.
struct ServedAnswer {
  sk_server_record record{};
  sk_pdhg_result result{};
  std::vector<double> col_value, row_dual, col_dual, row_activity;
};

struct sk_server {
  sk_pdhg_options options{};
  sk_server_options server_options{};
  std::mutex mu;  // answers and completed ids (written by the slot threads)
  std::unordered_map<int64_t, ServedAnswer> answers;
  std::deque<int64_t> completed;
  std::unique_ptr<sk::Server> server;  // last member: its slot threads stop first
};

namespace {

// Runs on a slot thread while the request's LP is still alive.
void store_answer(sk_server* s, sk::Request& r) {
  ServedAnswer a;
  const sk::HostLp& lp = *r.lp;
  a.record.id = r.id;
  a.record.arrival = r.arrival;
  a.record.admitted = r.admitted;
  a.record.finished = r.finished;
  a.record.latency = r.finished - r.arrival;
  a.record.slot = r.slot;
  a.record.batch_size = r.batch_size;
  a.record.error = r.error;
  a.record.num_rows = lp.num_rows;
  a.record.num_cols = lp.num_cols;
  if (r.error) {
    a.result.status = SK_SOLVE_NOT_SOLVED;
    std::snprintf(a.result.algorithm, sizeof(a.result.algorithm), "pdhg-gpu");
    std::snprintf(a.result.message, sizeof(a.result.message), "GPU PDHG failed: %s: %s",
                  sk_gpu_status_string(r.error), r.error_message.c_str());
  } else {
    if (s->server_options.keep_solutions) {
      a.col_value.resize(static_cast<size_t>(lp.num_cols));
      a.col_dual.resize(static_cast<size_t>(lp.num_cols));
      a.row_dual.resize(static_cast<size_t>(lp.num_rows));
      a.row_activity.resize(static_cast<size_t>(lp.num_rows));
    }
    const bool keep = s->server_options.keep_solutions != 0;
    const int32_t st = fill_result(lp, r.result, s->options, r.finished - r.admitted, &a.result,
                                   keep ? a.col_value.data() : nullptr,
                                   keep ? a.row_dual.data() : nullptr,
                                   keep ? a.col_dual.data() : nullptr,
                                   keep ? a.row_activity.data() : nullptr);
    if (st) a.record.error = st;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->answers.emplace(r.id, std::move(a));
  s->completed.push_back(r.id);
}

}  // namespace

extern "C" {

void sk_server_default_options(sk_server_options* out) {
  if (!out) return;
  std::memset(out, 0, sizeof(*out));
  // two slots: more concurrent batches shorten the queue wait but slow every batch, and the
  // largest requests' latency with them (p99 of a 2,000-request stream: 2 slots ~150 ms,
  // 4 slots 180-350 ms)
  out->slots = 2;
  out->max_batch = 256;
  out->memory_budget = 0;
  out->keep_solutions = 1;
  // one lane: with two slots the device is already saturated, and a batch started beside a
  // tail slows every batch more than it shortens the queue (2,000-request stream: p50 48 ms and
  // p99 141 ms with one lane, 57 ms and 171 ms with two); with one slot two lanes help
  // (p50 82 -> 68 ms, p99 205 -> 168 ms)
  out->lanes = 1;
  out->tail_fraction = 0.25;
}

int32_t sk_server_create(const sk_pdhg_options* options, const sk_server_options* server_options,
                         sk_server** out) {
  if (!out) return SK_GPU_INVALID_ARGUMENT;
  *out = nullptr;
  return guarded([&]() -> int32_t {
    auto s = std::make_unique<sk_server>();
    if (options) s->options = *options;
    else sk_pdhg_default_options(&s->options);
    if (const int32_t refused = refuse_backend(s->options)) return refused;
    if (server_options) s->server_options = *server_options;
    else sk_server_default_options(&s->server_options);
    const sk_server_options& so = s->server_options;
    if (so.slots < 1 || so.slots > 64 || so.max_batch < 1 || so.memory_budget < 0 ||
        so.lanes < 1 || so.lanes > 8 || !(so.tail_fraction >= 0.0 && so.tail_fraction <= 1.0)) {
      sk::set_last_error("sk_server_create: slots must be in [1, 64], max_batch >= 1, memory_budget "
                         ">= 0, lanes in [1, 8], tail_fraction in [0, 1]");
      return SK_GPU_INVALID_ARGUMENT;
    }
    sk::ServerOptions o;
    o.engine = to_engine(s->options);
    o.slots = so.slots;
    o.max_batch = so.max_batch;
    o.memory_budget = so.memory_budget;
    o.lanes = so.lanes;
    o.tail_fraction = so.tail_fraction;
    sk_server* raw = s.get();
    s->server = std::make_unique<sk::Server>(o, [raw](sk::Request& r) { store_answer(raw, r); });
    const int32_t st = s->server->open();
    if (st) return st;
    *out = s.release();
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_server_now(const sk_server* s, double* seconds) {
  if (!s || !seconds) return SK_GPU_INVALID_ARGUMENT;
  *seconds = s->server->now();
  return SK_GPU_SUCCESS;
}

int32_t sk_server_submit(sk_server* s, sk_lp* lp, double arrival, int64_t* id) {
  if (!s || !lp) return SK_GPU_INVALID_ARGUMENT;
  return guarded([&]() -> int32_t {
    // the server takes the model out of the handle only once it accepts the request
    auto owned = std::make_unique<sk::HostLp>(std::move(lp->lp));
    const int32_t st = s->server->submit(owned, arrival, id);
    if (st) {
      lp->lp = std::move(*owned);  // refused: the caller keeps the model
      return st;
    }
    delete lp;
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_server_completed(sk_server* s, int64_t* ids, int32_t capacity, int32_t* count) {
  if (!s || !count || capacity < 0 || (capacity > 0 && !ids)) return SK_GPU_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  int32_t n = 0;
  while (n < capacity && !s->completed.empty()) {
    ids[n++] = s->completed.front();
    s->completed.pop_front();
  }
  *count = n;
  return SK_GPU_SUCCESS;
}

int32_t sk_server_wait(sk_server* s, double timeout, int32_t* idle) {
  if (!s) return SK_GPU_INVALID_ARGUMENT;
  return guarded([&]() -> int32_t {
    const bool done = s->server->wait_idle(timeout);
    if (idle) *idle = done ? 1 : 0;
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_server_result(sk_server* s, int64_t id, sk_server_record* record, sk_pdhg_result* result,
                         double* col_value, double* row_dual, double* col_dual,
                         double* row_activity) {
  if (!s) return SK_GPU_INVALID_ARGUMENT;
  if (!s->server_options.keep_solutions && (col_value || row_dual || col_dual || row_activity)) {
    sk::set_last_error("sk_server_result: the server keeps no solution vectors (keep_solutions = 0)");
    return SK_GPU_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int32_t {
    ServedAnswer a;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->answers.find(id);
      if (it == s->answers.end()) {
        sk::set_last_error("sk_server_result: no completed request with id " + std::to_string(id));
        return SK_GPU_INVALID_ARGUMENT;
      }
      a = std::move(it->second);
      s->answers.erase(it);
    }
    if (record) *record = a.record;
    if (result) *result = a.result;
    if (col_value) std::copy(a.col_value.begin(), a.col_value.end(), col_value);
    if (row_dual) std::copy(a.row_dual.begin(), a.row_dual.end(), row_dual);
    if (col_dual) std::copy(a.col_dual.begin(), a.col_dual.end(), col_dual);
    if (row_activity) std::copy(a.row_activity.begin(), a.row_activity.end(), row_activity);
    return SK_GPU_SUCCESS;
  });
}

int32_t sk_server_stats_get(const sk_server* s, sk_server_stats* out) {
  if (!s || !out) return SK_GPU_INVALID_ARGUMENT;
  const sk::ServerStats t = s->server->stats();
  out->submitted = t.submitted;
  out->completed = t.completed;
  out->rejected = t.rejected;
  out->failed = t.failed;
  out->batches = t.batches;
  out->max_queue = t.max_queue;
  out->max_batch_seen = t.max_batch_seen;
  out->memory_budget = t.memory_budget;
  out->peak_held = t.peak_held;
  out->releases = t.releases;
  return SK_GPU_SUCCESS;
}

void sk_server_destroy(sk_server* s) {
  if (!s) return;
  s->server.reset();  // joins the slot threads before the answers go
  delete s;
}

}  // extern "C"
