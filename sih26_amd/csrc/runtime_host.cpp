// Host-side parts of the library shared by the HIP build (libsk_gpu.so) and the CPU-only stub
// build (libsk_gpu_stub.so): the thread-local last error, the status strings, the device
// footprint estimate used for admission, and the host vector comparison.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/sk_gpu.h"
#include "engine.hpp"
#include "status.hpp"

namespace sk {

namespace {
thread_local std::string g_last_error;
}  // namespace

void set_last_error(const std::string& message) { g_last_error = message; }
const char* last_error() { return g_last_error.c_str(); }

// ---- fault injection (sk_fault_inject) ---------------------------------------------------
namespace {
std::atomic<int64_t> g_alloc_countdown{0};
std::atomic<int64_t> g_launch_countdown{0};
std::atomic<int64_t> g_nan_instance{-1};
std::atomic<int64_t> g_solve_giveups{0};
std::atomic<int64_t> g_fired[5];  // per kind: how many times an injection took effect

// Counts one event down; true exactly when the countdown reaches zero from one.
bool count_down(std::atomic<int64_t>& c) {
  int64_t v = c.load();
  while (v > 0 && !c.compare_exchange_weak(v, v - 1)) {
  }
  return v == 1;
}
}  // namespace

bool fault_allocation_fails() {
  const bool f = count_down(g_alloc_countdown);
  if (f) ++g_fired[1];
  return f;
}
bool fault_launch_fails() {
  const bool f = count_down(g_launch_countdown);
  if (f) ++g_fired[2];
  return f;
}
int64_t fault_take_nan_instance() {
  const int64_t k = g_nan_instance.exchange(-1);
  if (k >= 0) ++g_fired[3];
  return k;
}
bool fault_solve_gives_up() {
  int64_t v = g_solve_giveups.load();
  while (v > 0 && !g_solve_giveups.compare_exchange_weak(v, v - 1)) {
  }
  if (v > 0) ++g_fired[4];
  return v > 0;
}

// Per-element byte counts of the buffers upload() allocates (w = iterate bytes):
//   nonzero: in both layouts, the CSR-order index, the FP64 original and FP64 scaled value,
//            and the staging-order index, value (w) and 16-bit slot; about 1 B of tile tables
//   row:     pointer, bounds, scaling, iterates (y, y', restart point: w each, 2 scaled
//            bounds), sums and candidates, best, activity, power-iteration image
//   column:  pointer, cost, bounds, scaling, 3 scaled vectors + FP64 bounds, iterates (x, x',
//            A^T y twice, xbar, restart point), sums, candidates, best, reduced costs, power
//            iteration vectors
//   instance: two state records (a decision reads one and writes the other), descriptor and
//            minimum tile entries
// plus a fixed allowance for the allocation granularity of some 60 buffers.
int64_t estimate_device_bytes(int64_t rows, int64_t cols, int64_t nnz, int64_t instances,
                              int32_t precision) {
  const double w = precision == 1 ? 8.0 : 4.0;
  const double per_nnz = 56.0 + 2.0 * w + 12.0 + 1.1;
  // each includes the iterate-precision running sum and its FP64 total
  const double per_row = 84.0 + 6.0 * w + 0.5;
  const double per_col = 116.0 + 10.0 * w + 0.5;
  double bytes = per_nnz * static_cast<double>(nnz) + per_row * static_cast<double>(rows) +
                 per_col * static_cast<double>(cols) + 1664.0 * static_cast<double>(instances) +
                 262144.0;
  if (instances == 1 && cols >= kSliceMinCols) {
    // the stacked column and row slices of an LP that runs sliced by default: per table and
    // nonzero the gather index, the map, the staging index and slot and the staged value; row
    // pointers and partial sums per stacked row
    bytes += 2.0 * (14.0 + w) * static_cast<double>(nnz) +
             (4.0 + w) * (kColumnSlices * static_cast<double>(rows) + kRowSlices * static_cast<double>(cols));
  }
  return static_cast<int64_t>(std::ceil(bytes));
}



}  // namespace sk

using namespace sk;

extern "C" {

const char* sk_gpu_status_string(int32_t status) {
  switch (status) {
    case SK_GPU_SUCCESS: return "success";
    case SK_GPU_UNAVAILABLE: return "hip unavailable";
    case SK_GPU_NO_DEVICE: return "no compatible hip device";
    case SK_GPU_INIT_FAILED: return "initialization failed";
    case SK_GPU_OUT_OF_MEMORY: return "out of device memory";
    case SK_GPU_LAUNCH_FAILED: return "kernel launch failed";
    case SK_GPU_COPY_FAILED: return "memory copy failed";
    case SK_GPU_SYNC_FAILED: return "device synchronization failed";
    case SK_GPU_INVALID_STATE: return "invalid device state";
    case SK_GPU_NON_FINITE: return "non-finite numerical output detected";
    case SK_GPU_INVALID_ARGUMENT: return "invalid argument";
    default: return "unknown status";
  }
}

const char* sk_gpu_last_error(void) { return last_error(); }

int32_t sk_fault_inject(int32_t kind, int64_t n) {
  switch (kind) {
    case 0:
      g_alloc_countdown = 0;
      g_launch_countdown = 0;
      g_nan_instance = -1;
      g_solve_giveups = 0;
      return kStatusSuccess;
    case 1:
    case 2:
      if (n < 1) break;
      (kind == 1 ? g_alloc_countdown : g_launch_countdown) = n;
      return kStatusSuccess;
    case 3:
      if (n < 0) break;
      g_nan_instance = n;
      return kStatusSuccess;
    case 4:
      if (n < 1) break;
      g_solve_giveups = n;
      return kStatusSuccess;
    default:
      break;
  }
  set_last_error("sk_fault_inject: kind must be 0..4, with n >= 1 (kinds 1, 2, 4) or n >= 0 (kind 3)");
  return kStatusInvalidArgument;
}

int64_t sk_fault_fired(int32_t kind) { return kind >= 1 && kind <= 4 ? g_fired[kind].load() : 0; }

/* gpu_ops_cpu.cpp:85-124 compare_vectors (a host function in the reference too) */
int32_t sk_gpu_compare_vectors(int32_t n, const double* t, const double* r, double tol,
                               sk_gpu_comparison* out) {
  if (!out) return kStatusInvalidArgument;
  std::memset(out, 0, sizeof(*out));
  out->passed = 1;
  if (n <= 0) return kStatusSuccess;
  if (!t || !r) {
    out->passed = 0;
    return kStatusSuccess;
  }
  double max_abs = 0.0, max_rel = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    if (!std::isfinite(t[i]) || !std::isfinite(r[i])) {
      out->contains_nan_or_inf = 1;
      out->passed = 0;
      return kStatusSuccess;
    }
    const double abs_diff = std::fabs(t[i] - r[i]);
    max_abs = std::max(max_abs, abs_diff);
    const double rel = abs_diff / std::max(1.0, std::max(std::fabs(t[i]), std::fabs(r[i])));
    max_rel = std::max(max_rel, rel);
    if (abs_diff > tol && rel > tol) out->passed = 0;
  }
  out->max_abs_diff = max_abs;
  out->max_rel_diff = max_rel;
  return kStatusSuccess;
}

}  // extern "C"
