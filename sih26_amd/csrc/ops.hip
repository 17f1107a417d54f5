// HIP runtime wrapper and standalone device operators behind the C ABI (include/sk_gpu.h):
// the counterparts of the reference's gpu_runtime.hpp and gpu_ops.hpp launch_gpu_* functions
// (src/gpu/gpu_runtime_cuda.cu:15-121, src/gpu/cuda_kernels.cu:25-196).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/sk_gpu.h"
#include "engine.hpp"
#include "device_util.hpp"
#include "status.hpp"

namespace sk {
namespace {

// products are rounded before they are added, as in the reference's CPU operators
// (gpu_ops_cpu.cpp:13-124): no FMA contraction in these kernels
#pragma clang fp contract(off)

int fail(hipError_t e, const char* what, int code) {
  set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  return code;
}

int launch_status(const char* what) {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? kStatusSuccess : fail(e, what, kStatusLaunchFailed);
}

unsigned grid_for(int64_t n, int per_block) {
  return static_cast<unsigned>(std::max<int64_t>(1, (n + per_block - 1) / per_block));
}

// Grow-only device scratch for the two-stage dot product.
double* g_dot_scratch = nullptr;
size_t g_dot_scratch_count = 0;

// Stage-1 workgroups of the dot product cover this many elements each.
constexpr int32_t kDotPerBlock = 8 * kBlock;

__global__ __launch_bounds__(kBlock) void k_axpy(int32_t n, double alpha, const double* __restrict__ x,
                                                  double* __restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] += alpha * x[i];
}

// Stage 1 of the dot product: each workgroup reduces a contiguous slice in a fixed order.
__global__ __launch_bounds__(kBlock) void k_dot_partial(int32_t n, int32_t per_block,
                                                         const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         double* __restrict__ part) {
  __shared__ double red[kBlock / kWave];
  const int begin = blockIdx.x * per_block;
  const int end = min(n, begin + per_block);
  double acc = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) acc += x[i] * y[i];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Stage 2: one workgroup reduces the partials.
__global__ __launch_bounds__(kBlock) void k_dot_final(int32_t count, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  __shared__ double red[kBlock / kWave];
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += kBlock) acc += part[i];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) *out = acc;
}

// cuda_kernels.cu:60-75 counterpart, FP64.
__global__ __launch_bounds__(kBlock) void k_primal_step(int32_t n, double tau, const double* __restrict__ x,
                                                         const double* __restrict__ cost,
                                                         const double* __restrict__ at_y,
                                                         const double* __restrict__ lo,
                                                         const double* __restrict__ hi,
                                                         double* __restrict__ x_next,
                                                         double* __restrict__ xbar) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const double g = cost[j] + at_y[j];
  double v = x[j] - tau * g;
  if (dev_finite_bound(lo[j]) && v < lo[j]) v = lo[j];
  else if (dev_finite_bound(hi[j]) && v > hi[j]) v = hi[j];
  x_next[j] = v;
  xbar[j] = 2.0 * v - x[j];
}

// One wavefront per compressed row: lanes stride the row, then a fixed butterfly.
__global__ __launch_bounds__(kBlock) void k_wave_rows(int32_t rows, const int32_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ idx,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ y, double alpha) {
  const int row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= rows) return;  // whole wavefronts exit together
  double acc = 0.0;
  for (int p = ptr[row] + lane; p < ptr[row + 1]; p += kWave) acc += val[p] * x[idx[p]];
  acc = group_sum(acc, kWave);
  if (lane == 0) y[row] = alpha * acc;
}

__global__ __launch_bounds__(kBlock) void k_csc_scatter(int32_t cols, const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ idx,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y, double alpha) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= cols) return;
  const double xj = alpha * x[j];
  if (xj == 0.0) return;
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) atomicAdd(&y[idx[p]], val[p] * xj);
}

// ---- the dot product without allocation (sk_gpu_dot_noalloc) ----------------------------
// Its scratch is a fixed device array of the library: slot 0 the result, then one partial per
// stage-1 workgroup. Each device holds its own copy until the process ends; every call on a
// device shares that device's copy (calls that overlap on one device race on it). Nothing
// needs clearing: stage 1 writes every partial stage 2 reads, and stage 2 writes the result.
constexpr int32_t kNoAllocDotBlocks = 131072;  // n <= 131,072 * 2,048 = 268,435,456
static __device__ __align__(256) double scratch[kNoAllocDotBlocks + 1];

// `scratch` in host code names the host-side stand-in of the array, not device memory: its
// device address is looked up once per device (with that device current) and kept here.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The address of the current device's `scratch` (set-up on the first call on each device).
int scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(e, "hipGetDevice", kStatusNoDevice);
  if (dev < 0 || dev >= kMaxDevices) {
    set_last_error("sk_gpu_dot_noalloc: device index beyond the library's table");
    return kStatusInvalidState;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(e, "hipGetSymbolAddress(scratch)", kStatusInitFailed);
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return kStatusSuccess;
}

}  // namespace
}  // namespace sk

using namespace sk;

extern "C" {

int32_t sk_gpu_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

int32_t sk_gpu_is_available(void) { return sk_gpu_device_count() > 0 ? 1 : 0; }

int32_t sk_gpu_get_device_info(int32_t device_id, sk_gpu_device_info* out) {
  if (!out) {
    set_last_error("sk_gpu_get_device_info: null output");
    return kStatusInvalidArgument;
  }
  const int count = sk_gpu_device_count();
  if (count == 0) {
    set_last_error("no HIP device is available");
    return kStatusNoDevice;
  }
  if (device_id < 0 || device_id >= count) {
    set_last_error("device id out of range");
    return kStatusInvalidState;
  }
  hipDeviceProp_t p;
  const hipError_t e = hipGetDeviceProperties(&p, device_id);
  if (e != hipSuccess) return fail(e, "hipGetDeviceProperties", kStatusInitFailed);
  std::memset(out, 0, sizeof(*out));
  out->device_id = device_id;
  std::strncpy(out->name, p.name, sizeof(out->name) - 1);
  std::strncpy(out->gcn_arch_name, p.gcnArchName, sizeof(out->gcn_arch_name) - 1);
  out->total_memory_bytes = p.totalGlobalMem;
  out->gcn_arch_major = p.major;
  out->gcn_arch_minor = p.minor;
  out->multiprocessor_count = p.multiProcessorCount;
  out->warp_size = p.warpSize;
  out->max_threads_per_block = p.maxThreadsPerBlock;
  return kStatusSuccess;
}

int32_t sk_gpu_set_device(int32_t device_id) {
  if (sk_gpu_device_count() == 0) {
    set_last_error("no HIP device is available");
    return kStatusNoDevice;
  }
  const hipError_t e = hipSetDevice(device_id);
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipSetDevice", kStatusInvalidState);
}

int32_t sk_gpu_synchronize(void) {
  const hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipDeviceSynchronize", kStatusSyncFailed);
}

int32_t sk_gpu_malloc(void** dev_ptr, size_t bytes) {
  if (!dev_ptr) return kStatusInvalidArgument;
  *dev_ptr = nullptr;
  const hipError_t e = hipMalloc(dev_ptr, bytes);
  if (e != hipSuccess) {
    *dev_ptr = nullptr;
    return fail(e, "hipMalloc", kStatusOutOfMemory);
  }
  return kStatusSuccess;
}

int32_t sk_gpu_free(void* dev_ptr) {
  const hipError_t e = hipFree(dev_ptr);
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipFree", kStatusInvalidState);
}

int32_t sk_gpu_copy_h2d(void* dst, const void* src, size_t bytes) {
  const hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipMemcpy H2D", kStatusCopyFailed);
}

int32_t sk_gpu_copy_d2h(void* dst, const void* src, size_t bytes) {
  const hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipMemcpy D2H", kStatusCopyFailed);
}

int32_t sk_gpu_memset(void* dev_ptr, int32_t value, size_t bytes) {
  const hipError_t e = hipMemset(dev_ptr, value, bytes);
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipMemset", kStatusCopyFailed);
}

int32_t sk_gpu_axpy(int32_t n, double alpha, const double* d_x, double* d_y) {
  if (n <= 0) return kStatusSuccess;
  if (!d_x || !d_y) return kStatusInvalidState;
  k_axpy<<<std::min<unsigned>(grid_for(n, kBlock), 4096u), kBlock>>>(n, alpha, d_x, d_y);
  return launch_status("k_axpy");
}

int32_t sk_gpu_dot(int32_t n, const double* d_x, const double* d_y, double* host_result) {
  if (!host_result) return kStatusInvalidArgument;
  *host_result = 0.0;
  if (n <= 0) return kStatusSuccess;
  if (!d_x || !d_y) return kStatusInvalidState;
  const int32_t per_block = kDotPerBlock;
  const auto blocks = static_cast<int32_t>(grid_for(n, per_block));
  if (static_cast<size_t>(blocks) + 1 > g_dot_scratch_count) {
    if (g_dot_scratch) (void)hipFree(g_dot_scratch);
    g_dot_scratch = nullptr;
    g_dot_scratch_count = 0;
    const hipError_t e = hipMalloc(&g_dot_scratch, (static_cast<size_t>(blocks) + 1) * sizeof(double));
    if (e != hipSuccess) return fail(e, "hipMalloc", kStatusOutOfMemory);
    g_dot_scratch_count = static_cast<size_t>(blocks) + 1;
  }
  k_dot_partial<<<blocks, kBlock>>>(n, per_block, d_x, d_y, g_dot_scratch + 1);
  int st = launch_status("k_dot_partial");
  if (st) return st;
  k_dot_final<<<1, kBlock>>>(blocks, g_dot_scratch + 1, g_dot_scratch);
  if ((st = launch_status("k_dot_final"))) return st;
  const hipError_t e = hipMemcpy(host_result, g_dot_scratch, sizeof(double), hipMemcpyDeviceToHost);
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipMemcpy D2H", kStatusCopyFailed);
}

int32_t sk_gpu_dot_noalloc(int32_t n, const double* d_x, const double* d_y, double* host_result) {
  if (!host_result) return kStatusInvalidArgument;
  *host_result = 0.0;
  if (n <= 0) return kStatusSuccess;
  if (!d_x || !d_y) return kStatusInvalidState;
  const auto blocks = static_cast<int32_t>(grid_for(n, kDotPerBlock));
  if (blocks > kNoAllocDotBlocks) {
    set_last_error("sk_gpu_dot_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                   std::to_string(static_cast<int64_t>(kNoAllocDotBlocks) * kDotPerBlock) +
                   " elements (its fixed scratch holds one partial per 2,048 elements)");
    return kStatusInvalidArgument;
  }
  double* part = nullptr;
  int st = scratch_address(&part);
  if (st) return st;
  // the same two kernels and partition as sk_gpu_dot, so the same bits
  k_dot_partial<<<blocks, kBlock>>>(n, kDotPerBlock, d_x, d_y, part + 1);
  if ((st = launch_status("k_dot_partial"))) return st;
  k_dot_final<<<1, kBlock>>>(blocks, part + 1, part);
  if ((st = launch_status("k_dot_final"))) return st;
  const hipError_t e = hipMemcpy(host_result, part, sizeof(double), hipMemcpyDeviceToHost);
  return e == hipSuccess ? kStatusSuccess : fail(e, "hipMemcpy D2H", kStatusCopyFailed);
}

int32_t sk_gpu_pdhg_primal_step(int32_t n, double tau, const double* d_x, const double* d_cost,
                                const double* d_at_y, const double* d_lo, const double* d_hi,
                                double* d_x_next, double* d_xbar) {
  if (n <= 0) return kStatusSuccess;
  if (!d_x || !d_cost || !d_at_y || !d_lo || !d_hi || !d_x_next || !d_xbar) return kStatusInvalidState;
  k_primal_step<<<grid_for(n, kBlock), kBlock>>>(n, tau, d_x, d_cost, d_at_y, d_lo, d_hi, d_x_next, d_xbar);
  return launch_status("k_primal_step");
}

int32_t sk_gpu_execute_pdhg_primal_step(int32_t n, double tau, const double* x,
                                        const double* cost, const double* at_y,
                                        const double* col_lower, const double* col_upper,
                                        double* x_next, double* extrapolated) {
  // execute_pdhg_primal_step (gpu_ops_cpu.cpp:126-207) without its CPU fallback: the same
  // sequence of failure classes, each returned as soon as it happens; the caller that wants
  // the fallback runs its own CPU step on any status other than success.
  if (n <= 0) return kStatusSuccess;
  if (!x || !cost || !at_y || !col_lower || !col_upper || !x_next || !extrapolated) {
    set_last_error("sk_gpu_execute_pdhg_primal_step: null host array");
    return kStatusInvalidArgument;
  }
  if (sk_gpu_device_count() <= 0) {
    set_last_error("sk_gpu_execute_pdhg_primal_step: no HIP device is available");
    return kStatusUnavailable;
  }
  const size_t bytes = static_cast<size_t>(n) * sizeof(double);
  struct Buffers {
    double* p[7] = {};
    ~Buffers() {
      for (double* q : p)
        if (q) (void)hipFree(q);
    }
  } d;
  for (double*& q : d.p) {
    const hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
      q = nullptr;
      return fail(e, "hipMalloc", kStatusOutOfMemory);
    }
  }
  const double* host_in[5] = {x, cost, at_y, col_lower, col_upper};
  for (int k = 0; k < 5; ++k) {
    const hipError_t e = hipMemcpy(d.p[k], host_in[k], bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(e, "hipMemcpy H2D", kStatusCopyFailed);
  }
  k_primal_step<<<grid_for(n, kBlock), kBlock>>>(n, tau, d.p[0], d.p[1], d.p[2], d.p[3], d.p[4],
                                                 d.p[5], d.p[6]);
  int st = launch_status("k_primal_step");
  if (st) return st;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(e, "hipDeviceSynchronize", kStatusSyncFailed);
  e = hipMemcpy(x_next, d.p[5], bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(extrapolated, d.p[6], bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(e, "hipMemcpy D2H", kStatusCopyFailed);
  // the reference's numerical sanity check of the GPU output (gpu_ops_cpu.cpp:187-193)
  for (int32_t i = 0; i < n; ++i) {
    if (!std::isfinite(x_next[i]) || !std::isfinite(extrapolated[i])) {
      set_last_error("sk_gpu_execute_pdhg_primal_step: non-finite output at column " + std::to_string(i));
      return kStatusNonFinite;
    }
  }
  return kStatusSuccess;
}

int32_t sk_gpu_spmv_csc(int32_t num_rows, int32_t num_cols, const int32_t* d_col_starts,
                        const int32_t* d_row_indices, const double* d_values, const double* d_x,
                        double* d_y, double alpha, int32_t transpose) {
  if (num_rows <= 0 || num_cols <= 0) return kStatusSuccess;
  if (!d_col_starts || !d_row_indices || !d_values || !d_x || !d_y) return kStatusInvalidState;
  if (transpose) {
    k_wave_rows<<<grid_for(num_cols, kBlock / kWave), kBlock>>>(num_cols, d_col_starts, d_row_indices,
                                                                d_values, d_x, d_y, alpha);
    return launch_status("k_wave_rows");
  }
  const hipError_t e = hipMemsetAsync(d_y, 0, static_cast<size_t>(num_rows) * sizeof(double), nullptr);
  if (e != hipSuccess) return fail(e, "hipMemsetAsync", kStatusCopyFailed);
  k_csc_scatter<<<grid_for(num_cols, kBlock), kBlock>>>(num_cols, d_col_starts, d_row_indices,
                                                        d_values, d_x, d_y, alpha);
  return launch_status("k_csc_scatter");
}

int32_t sk_gpu_spmv_csr(int32_t num_rows, int32_t num_cols, const int32_t* d_row_starts,
                        const int32_t* d_col_indices, const double* d_values, const double* d_x,
                        double* d_y, double alpha) {
  if (num_rows <= 0 || num_cols <= 0) return kStatusSuccess;
  if (!d_row_starts || !d_col_indices || !d_values || !d_x || !d_y) return kStatusInvalidState;
  k_wave_rows<<<grid_for(num_rows, kBlock / kWave), kBlock>>>(num_rows, d_row_starts, d_col_indices,
                                                              d_values, d_x, d_y, alpha);
  return launch_status("k_wave_rows");
}

}  // extern "C"
