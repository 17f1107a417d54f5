// The no-device half of the CPU-only build (libsk_gpu_stub.so), the counterpart of the
// reference's src/gpu/gpu_runtime_stub.cpp:11-82, which SANKHYA links when CUDA is off
// (CMakeLists.txt:212-222). It replaces the HIP translation units (engine.hip, ops.hip,
// normal.hip): every device entry point returns SK_GPU_UNAVAILABLE with a message, the device
// queries report no device, and nothing is faked. The host parts (models, the MPS reader and
// writer, the instance cache, the .sol writer, the option parser and the backend selection)
// are the same objects as in the HIP build.
#include <memory>
#include <string>
#include <vector>

#include "../../../include/sk_gpu.h"
#include "../engine.hpp"
#include "../normal.hpp"
#include "../status.hpp"

namespace sk {

namespace {

constexpr const char* kNoDevice =
    "this is the CPU-only build of the library (libsk_gpu_stub.so): it has no HIP device code";

int unavailable() {
  set_last_error(kNoDevice);
  return kStatusUnavailable;
}

class StubEngine final : public Engine {
 public:
  int32_t add(const HostLp*) override { return count_++; }
  int warm_start(int32_t, const double*, const double*, double, double) override { return unavailable(); }
  int upload() override { return unavailable(); }
  int solve() override { return unavailable(); }
  int result(int32_t, InstanceResult*) const override { return unavailable(); }
  void clear() override { count_ = 0; }
  int start() override { return unavailable(); }
  int poll(bool, bool* finished, std::vector<int32_t>*) override {
    *finished = true;
    return unavailable();
  }
  int64_t device_bytes() const override { return 0; }
  void release() override {}
  const EngineTimings& timings() const override { return timings_; }
  int32_t size() const override { return count_; }
  int time_kernel(int, int, double*) override { return unavailable(); }
  double kernel_bytes(int) const override { return 0.0; }
  int64_t total_nnz() const override { return 0; }
  int64_t total_rows() const override { return 0; }
  int64_t total_cols() const override { return 0; }
  int stage_choice() const override { return 1; }
  int32_t streams() const override { return 1; }
  int32_t fused() const override { return 0; }
  int32_t column_slices() const override { return 0; }
  int32_t tile_capacity() const override { return 0; }
  int probe_spmv(int32_t, int, const double*, double*) override { return unavailable(); }
  int probe_scaling(int32_t, double*, double*, double*) override { return unavailable(); }
  int column_order(int32_t, int32_t*) const override { return unavailable(); }

 private:
  int32_t count_ = 0;
  EngineTimings timings_;
};

}  // namespace

std::unique_ptr<Engine> make_single_engine(const EngineOptions&) {
  return std::make_unique<StubEngine>();
}
int current_device(int*) { return unavailable(); }
int use_device(int) { return unavailable(); }
int query_device_memory(int64_t*, int64_t*) { return unavailable(); }

struct NormalEquations::Impl {};
NormalEquations::NormalEquations() : impl_(new Impl) {}
NormalEquations::~NormalEquations() = default;
int64_t NormalEquations::factor_bytes(int32_t m) {
  return static_cast<int64_t>(m) * m * static_cast<int64_t>(sizeof(double));
}
int NormalEquations::init(int32_t, int32_t, const std::vector<int32_t>&, const std::vector<int32_t>&,
                          const std::vector<double>&, const std::vector<int32_t>&,
                          const std::vector<int32_t>&, const std::vector<double>&) {
  return unavailable();
}
int NormalEquations::factor(const double*, const double*, double, int64_t* floored) {
  if (floored) *floored = 0;
  return unavailable();
}
int NormalEquations::solve(double*) { return unavailable(); }

}  // namespace sk

using sk::unavailable;

extern "C" {

/* gpu_runtime_stub.cpp:11-17: no device */
int32_t sk_gpu_is_available(void) { return 0; }
int32_t sk_gpu_device_count(void) { return 0; }
int32_t sk_gpu_get_device_info(int32_t, sk_gpu_device_info*) { return unavailable(); }
int32_t sk_gpu_set_device(int32_t) { return unavailable(); }
int32_t sk_gpu_synchronize(void) { return unavailable(); }
int32_t sk_gpu_malloc(void** dev_ptr, size_t) {
  if (dev_ptr) *dev_ptr = nullptr;
  return unavailable();
}
int32_t sk_gpu_free(void*) { return unavailable(); }
int32_t sk_gpu_copy_h2d(void*, const void*, size_t) { return unavailable(); }
int32_t sk_gpu_copy_d2h(void*, const void*, size_t) { return unavailable(); }
int32_t sk_gpu_memset(void*, int32_t, size_t) { return unavailable(); }
int32_t sk_gpu_axpy(int32_t, double, const double*, double*) { return unavailable(); }
int32_t sk_gpu_dot(int32_t, const double*, const double*, double* host_result) {
  if (host_result) *host_result = 0.0;
  return unavailable();
}
int32_t sk_gpu_dot_noalloc(int32_t, const double*, const double*, double* host_result) {
  if (host_result) *host_result = 0.0;
  return unavailable();
}
int32_t sk_gpu_pdhg_primal_step(int32_t, double, const double*, const double*, const double*,
                                const double*, const double*, double*, double*) {
  return unavailable();
}
int32_t sk_gpu_execute_pdhg_primal_step(int32_t n, double, const double*, const double*,
                                        const double*, const double*, const double*, double*,
                                        double*) {
  return n <= 0 ? static_cast<int32_t>(sk::kStatusSuccess) : unavailable();
}
int32_t sk_gpu_spmv_csc(int32_t, int32_t, const int32_t*, const int32_t*, const double*,
                        const double*, double*, double, int32_t) {
  return unavailable();
}
int32_t sk_gpu_spmv_csr(int32_t, int32_t, const int32_t*, const int32_t*, const double*,
                        const double*, double*, double) {
  return unavailable();
}

}  // extern "C"
