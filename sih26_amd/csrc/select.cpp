// Backend selection: the reference's adaptive CPU/GPU decision (src/core/adaptive.cpp:32-232,
// include/sankhya/adaptive.hpp:22-121) restated behind the C ABI. Host code only; the same
// object goes into the HIP build and the CPU-only build, where no device is ever reported.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sk_gpu.h"
#include "engine.hpp"
#include "handles.hpp"
#include "status.hpp"

namespace {

struct Decision {
  int32_t selected = SK_BACKEND_CPU;
  int32_t requested = SK_BACKEND_AUTO;
  bool valid = true;
  std::string primary;
  std::vector<std::string> details;
};

const char* backend_name(int32_t b) {
  switch (b) {
    case SK_BACKEND_AUTO: return "auto";
    case SK_BACKEND_CPU: return "cpu";
    case SK_BACKEND_GPU: return "gpu";
    case SK_BACKEND_HYBRID: return "hybrid";
    default: return "unknown";
  }
}

std::string format(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
std::string format(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return buf;
}

// select_backend (adaptive.cpp:98-232): explicit modes first, then the legacy flag, then the
// adaptive tests in the reference's order (device, problem class, size, memory headroom).
Decision decide(const sk_problem_characteristics& c, const sk_pdhg_options& o,
                const sk_adaptive_config& cfg) {
  Decision d;
  int32_t mode = SK_BACKEND_AUTO;
  const bool explicit_backend = o.backend != SK_BACKEND_AUTO;
  if (explicit_backend) mode = o.backend;
  // the legacy gpu flag (adaptive.cpp:117-134): a GPU request that degrades to the CPU
  if (!explicit_backend && o.gpu) {
    if (c.gpu_available) {
      mode = SK_BACKEND_GPU;
    } else {
      d.selected = SK_BACKEND_CPU;
      d.requested = SK_BACKEND_AUTO;
      d.valid = true;
      d.primary = "--gpu requested but no HIP device is available in this environment/build; "
                  "running on CPU";
      d.details.push_back("Legacy --gpu option permits graceful fallback to CPU");
      return d;
    }
  }
  d.requested = mode;
  if (mode == SK_BACKEND_CPU) {
    d.selected = SK_BACKEND_CPU;
    d.primary = "Explicitly requested CPU backend (--backend=cpu)";
    d.details.push_back("User override forces CPU execution");
    return d;
  }
  if (mode == SK_BACKEND_GPU) {
    d.selected = SK_BACKEND_GPU;
    if (c.gpu_available) {
      d.primary = "Explicitly requested GPU backend (--backend=gpu); HIP device ready";
      d.details.push_back(std::string("Device 0: ") + (c.gpu_name[0] ? c.gpu_name : "HIP device"));
    } else {
      // an explicit GPU request never falls back silently
      d.valid = false;
      d.primary = "Explicitly requested GPU backend (--backend=gpu), but no GPU device is available";
      d.details.push_back("No HIP device detected, or the HIP runtime is unavailable on this system");
      d.details.push_back("Explicit GPU mode refuses silent fallback to ensure predictable optimization");
    }
    return d;
  }
  if (mode == SK_BACKEND_HYBRID) {
    d.selected = SK_BACKEND_HYBRID;
    d.valid = false;
    d.primary = "Explicitly requested Hybrid backend (--backend=hybrid), which is unsupported in "
                "this version";
    d.details.push_back("Coordinated CPU/GPU hybrid execution is not implemented");
    return d;
  }
  d.requested = SK_BACKEND_AUTO;
  if (!c.gpu_available) {
    d.selected = SK_BACKEND_CPU;
    d.primary = "Adaptive selector chose CPU: no GPU device is available in this environment";
    d.details.push_back("Host system does not report an available HIP device");
    d.details.push_back("Safe fallback to deterministic CPU solver");
    return d;
  }
  if (c.num_integer_cols > 0) {
    d.selected = SK_BACKEND_CPU;
    d.primary = format("Adaptive selector chose CPU: problem is %s, which is optimized for CPU "
                       "branch-and-cut", c.problem_class);
    d.details.push_back("MIP branch-and-cut tree exploration runs on host CPU");
    return d;
  }
  const int64_t max_dim = std::max<int64_t>(c.num_rows, c.num_cols);
  if (c.num_nonzeros < cfg.min_gpu_nonzeros || max_dim < cfg.min_gpu_dimension) {
    d.selected = SK_BACKEND_CPU;
    d.primary = format("Adaptive selector chose CPU: problem size (%lld nonzeros, %dx%d) is below "
                       "GPU threshold (%lld nonzeros, %lld min dimension)",
                       static_cast<long long>(c.num_nonzeros), c.num_rows, c.num_cols,
                       static_cast<long long>(cfg.min_gpu_nonzeros),
                       static_cast<long long>(cfg.min_gpu_dimension));
    d.details.push_back("Small problem workload does not justify the host-to-device transfer");
    d.details.push_back("CPU caches offer superior throughput on small workloads");
    return d;
  }
  if (c.gpu_memory_bytes > 0 &&
      c.estimated_memory_bytes > static_cast<double>(c.gpu_memory_bytes) * cfg.max_gpu_memory_headroom) {
    d.selected = SK_BACKEND_CPU;
    d.primary = "Adaptive selector chose CPU: estimated model memory exceeds safe GPU memory headroom";
    d.details.push_back(format("Model requires ~%.1f MB, exceeding %.0f%% of device capacity (%.1f MB)",
                               c.estimated_memory_bytes / (1024.0 * 1024.0),
                               cfg.max_gpu_memory_headroom * 100.0,
                               static_cast<double>(c.gpu_memory_bytes) / (1024.0 * 1024.0)));
    return d;
  }
  d.selected = SK_BACKEND_GPU;
  d.primary = format("Adaptive selector chose GPU: problem size (%lld nonzeros, %dx%d) exceeds GPU "
                     "acceleration threshold",
                     static_cast<long long>(c.num_nonzeros), c.num_rows, c.num_cols);
  d.details.push_back(std::string("HIP device '") + (c.gpu_name[0] ? c.gpu_name : "default") +
                      "' has sufficient memory");
  d.details.push_back("Sparse matrix-vector products parallelize across the GPU's compute units");
  return d;
}

void copy_text(char* dst, size_t size, const std::string& src) {
  std::snprintf(dst, size, "%s", src.c_str());
}

}  // namespace

extern "C" {

void sk_adaptive_default_config(sk_adaptive_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->min_gpu_nonzeros = 50000;
  c->min_gpu_dimension = 1000;
  c->min_gpu_density = 1e-6;
  c->max_gpu_memory_headroom = 0.8;
  c->explain_decision = 1;
}

int32_t sk_adaptive_set_option(sk_adaptive_config* c, const char* key_in, const char* value_in) {
  if (!c || !key_in || !value_in) return SK_GPU_INVALID_ARGUMENT;
  const std::string key(key_in), v(value_in);
  auto bad = [&](const char* what) {
    sk::set_last_error("option " + key + ": " + what + " '" + v + "'");
    return SK_GPU_INVALID_ARGUMENT;
  };
  char* end = nullptr;
  if (key == "adaptive_min_nnz" || key == "adaptive_min_dim") {
    const long long i = std::strtoll(v.c_str(), &end, 10);
    if (v.empty() || *end != '\0' || i < 0 || i > 1000000000LL) return bad("expected an integer in [0, 1e9], got");
    (key == "adaptive_min_nnz" ? c->min_gpu_nonzeros : c->min_gpu_dimension) = i;
    return SK_GPU_SUCCESS;
  }
  if (key == "adaptive_min_density") {
    const double d = std::strtod(v.c_str(), &end);
    if (v.empty() || *end != '\0' || !(d >= 0.0 && d <= 1.0)) return bad("expected a number in [0, 1], got");
    c->min_gpu_density = d;
    return SK_GPU_SUCCESS;
  }
  if (key == "adaptive_explain") {
    if (v == "true" || v == "1" || v == "on" || v == "yes") c->explain_decision = 1;
    else if (v == "false" || v == "0" || v == "off" || v == "no") c->explain_decision = 0;
    else return bad("expected a boolean, got");
    return SK_GPU_SUCCESS;
  }
  sk::set_last_error("unknown option '" + key + "'");
  return SK_GPU_INVALID_ARGUMENT;
}

int32_t sk_analyze_problem(const sk_lp* lp, int32_t integer_columns,
                           int32_t has_quadratic_objective, sk_problem_characteristics* out) {
  if (!lp || !out || integer_columns < 0) return SK_GPU_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  const sk::HostLp& m = lp->lp;
  out->num_rows = m.num_rows;
  out->num_cols = m.num_cols;
  out->num_nonzeros = m.nnz();
  out->num_integer_cols = integer_columns;
  out->has_quadratic_objective = has_quadratic_objective ? 1 : 0;
  const char* cls = integer_columns > 0 && has_quadratic_objective ? "MIQP"
                    : integer_columns > 0                          ? "MILP"
                    : has_quadratic_objective                      ? "QP"
                                                                   : "LP";
  std::snprintf(out->problem_class, sizeof out->problem_class, "%s", cls);
  const double entries = static_cast<double>(m.num_rows) * static_cast<double>(m.num_cols);
  out->density = entries > 0.0 ? static_cast<double>(out->num_nonzeros) / entries : 0.0;
  out->sparsity_ratio = entries > 0.0 ? 1.0 - out->density : 1.0;
  out->estimated_matvec_flops = 2.0 * static_cast<double>(out->num_nonzeros);
  // CSC (index + FP64 value per nonzero, column pointers) plus four FP64 vectors of rows and
  // columns, as analyze_problem estimates (adaptive.cpp:60-65)
  out->estimated_memory_bytes =
      static_cast<double>(out->num_nonzeros) * (sizeof(int32_t) + sizeof(double)) +
      static_cast<double>(m.num_cols + 1) * sizeof(int32_t) +
      static_cast<double>(m.num_rows + m.num_cols) * sizeof(double) * 4.0;
  out->gpu_available = sk_gpu_is_available();
  out->gpu_device_count = sk_gpu_device_count();
  if (out->gpu_available && out->gpu_device_count > 0) {
    sk_gpu_device_info info;
    if (sk_gpu_get_device_info(0, &info) == SK_GPU_SUCCESS) {
      std::snprintf(out->gpu_name, sizeof out->gpu_name, "%s", info.name);
      out->gpu_memory_bytes = info.total_memory_bytes;
    }
  }
  return SK_GPU_SUCCESS;
}

int32_t sk_select_backend(const sk_problem_characteristics* chars, const sk_pdhg_options* options,
                          const sk_adaptive_config* config, sk_backend_decision* out) {
  if (!chars || !out) return SK_GPU_INVALID_ARGUMENT;
  sk_pdhg_options o;
  if (options) o = *options;
  else sk_pdhg_default_options(&o);
  if (o.backend < SK_BACKEND_AUTO || o.backend > SK_BACKEND_HYBRID) {
    sk::set_last_error("sk_select_backend: backend must be one of SK_BACKEND_*");
    return SK_GPU_INVALID_ARGUMENT;
  }
  sk_adaptive_config cfg;
  if (config) cfg = *config;
  else sk_adaptive_default_config(&cfg);
  const Decision d = decide(*chars, o, cfg);
  std::memset(out, 0, sizeof(*out));
  out->selected_backend = d.selected;
  out->requested_mode = d.requested;
  out->is_valid = d.valid ? 1 : 0;
  out->detail_count = static_cast<int32_t>(d.details.size());
  copy_text(out->primary_reason, sizeof out->primary_reason, d.primary);
  // BackendDecision::format_explanation (adaptive.cpp:15-29)
  std::string e = std::string("Backend decision: ") + backend_name(d.selected);
  if (!d.valid) e += " [REJECTED: requested backend unavailable]";
  e += "\nPrimary reason: " + d.primary;
  if (!d.details.empty()) {
    e += "\nDetailed factors:";
    for (const auto& r : d.details) e += "\n  - " + r;
  }
  copy_text(out->explanation, sizeof out->explanation, e);
  return SK_GPU_SUCCESS;
}

}  // extern "C"
