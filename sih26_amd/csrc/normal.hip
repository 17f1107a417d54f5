// Dense normal equations of the interior-point finish (normal.hpp): assembly and the
// diagonal-block Cholesky are HIP kernels; the panel solves and trailing updates, which hold
// all but O(m * 64^2) of the m^3 / 3 flops, are rocBLAS dtrsm / dsyrk (FP64 matrix cores).
#include "normal.hpp"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <string>

#include "engine.hpp"
#include "status.hpp"

namespace sk {
namespace {

constexpr int kLanes = 256;
// Width of the diagonal blocks factorised by k_chol_block.
constexpr int kPanel = 64;
// Rows of one column of M accumulated in LDS per assembly pass (64 KB of FP64).
constexpr int kWindow = 8192;

#define NE_HIP(call, code)                                                              \
  do {                                                                                  \
    const hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                             \
      set_last_error(std::string(#call) + ": " + hipGetErrorString(e_));                \
      return (code);                                                                    \
    }                                                                                   \
  } while (0)

#define NE_BLAS(call)                                                                   \
  do {                                                                                  \
    const rocblas_status s_ = (call);                                                   \
    if (s_ != rocblas_status_success) {                                                 \
      set_last_error(std::string(#call) + ": " + rocblas_status_to_string(s_));         \
      return kStatusLaunchFailed;                                                       \
    }                                                                                   \
  } while (0)

// Column j of M's lower triangle, rows [max(j, w0), w1):
//   M_ij = sum over the nonzeros A_jk of row j, in CSR order, of (A_jk theta_k) A_ik,
// plus shift_j on the diagonal. Each column k of A adds to distinct rows i, so the lanes of
// one k never collide; a barrier separates consecutive k, which fixes the order of additions.
__global__ __launch_bounds__(kLanes) void k_assemble(int32_t m, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ ri,
                                                      const double* __restrict__ rv,
                                                      const int32_t* __restrict__ cp,
                                                      const int32_t* __restrict__ ci,
                                                      const double* __restrict__ cv,
                                                      const double* __restrict__ theta,
                                                      const double* __restrict__ shift,
                                                      int32_t w0, int32_t w1,
                                                      double* __restrict__ M) {
  extern __shared__ double col[];
  const int j = blockIdx.x, tid = threadIdx.x;
  const int lo = max(w0, j);
  if (lo >= w1) return;  // the same in every lane
  for (int i = lo + tid; i < w1; i += kLanes) col[i - w0] = 0.0;
  __syncthreads();
  for (int p = rp[j]; p < rp[j + 1]; ++p) {
    const int k = ri[p];
    const double f = rv[p] * theta[k];
    for (int q = cp[k] + tid; q < cp[k + 1]; q += kLanes) {
      const int i = ci[q];
      if (i >= lo && i < w1) col[i - w0] += f * cv[q];
    }
    __syncthreads();
  }
  if (tid == 0 && j >= w0 && j < w1) col[j - w0] += shift[j];
  __syncthreads();
  double* out = M + static_cast<size_t>(j) * m;
  for (int i = lo + tid; i < w1; i += kLanes) out[i] = col[i - w0];
}

// In-place Cholesky of the b x b diagonal block at (kb, kb) of the column-major M (b <= 64),
// right-looking, in LDS. A pivot below `floor` (zero and negative ones included) is raised
// to `floor` and counted; a NaN pivot stays NaN, so a broken system shows in the solution.
__global__ __launch_bounds__(kLanes) void k_chol_block(double* __restrict__ M, int32_t ld,
                                                        int32_t kb, int32_t b, double floor,
                                                        int64_t* __restrict__ floored) {
  __shared__ double L[kPanel][kPanel + 1];
  const int tid = threadIdx.x;
  double* A = M + static_cast<size_t>(kb) * ld + kb;
  for (int idx = tid; idx < b * b; idx += kLanes) {
    const int i = idx % b, c = idx / b;
    L[i][c] = i >= c ? A[static_cast<size_t>(c) * ld + i] : 0.0;
  }
  __syncthreads();
  int64_t count = 0;
  for (int k = 0; k < b; ++k) {
    double d = L[k][k];
    const bool low = d < floor;
    if (low) d = floor;
    const double s = sqrt(d);
    __syncthreads();  // every lane has read the pivot before it is overwritten
    if (tid == 0) {
      L[k][k] = s;
      count += low ? 1 : 0;
    }
    for (int i = k + 1 + tid; i < b; i += kLanes) L[i][k] /= s;
    __syncthreads();
    const int r = b - k - 1;
    for (int idx = tid; idx < r * r; idx += kLanes) {
      const int i = k + 1 + idx % r, c = k + 1 + idx / r;
      if (i >= c) L[i][c] -= L[i][k] * L[c][k];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < b * b; idx += kLanes) {
    const int i = idx % b, c = idx / b;
    if (i >= c) A[static_cast<size_t>(c) * ld + i] = L[i][c];
  }
  if (tid == 0) *floored += count;
}

// Waits for a stream when it goes out of scope, so host sources of queued asynchronous copies
// outlive the copies on every return path.
struct StreamDrain {
  hipStream_t stream;
  ~StreamDrain() {
    if (stream) (void)hipStreamSynchronize(stream);
  }
};

template <class U>
int upload(const std::vector<U>& host, U** dev, hipStream_t s) {
  NE_HIP(hipMalloc(reinterpret_cast<void**>(dev), std::max<size_t>(1, host.size()) * sizeof(U)),
         kStatusOutOfMemory);
  if (!host.empty())
    NE_HIP(hipMemcpyAsync(*dev, host.data(), host.size() * sizeof(U), hipMemcpyHostToDevice, s),
           kStatusCopyFailed);
  return kStatusSuccess;
}

}  // namespace

struct NormalEquations::Impl {
  int32_t m = 0, n = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  int32_t *rp = nullptr, *ri = nullptr, *cp = nullptr, *ci = nullptr;
  double *rv = nullptr, *cv = nullptr, *theta = nullptr, *shift = nullptr, *M = nullptr;
  double* x = nullptr;
  int64_t* floored = nullptr;

  ~Impl() {
    for (void* p : {static_cast<void*>(rp), static_cast<void*>(ri), static_cast<void*>(cp),
                    static_cast<void*>(ci), static_cast<void*>(rv), static_cast<void*>(cv),
                    static_cast<void*>(theta), static_cast<void*>(shift), static_cast<void*>(M),
                    static_cast<void*>(x), static_cast<void*>(floored)})
      if (p) (void)hipFree(p);
    if (blas) (void)rocblas_destroy_handle(blas);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

NormalEquations::NormalEquations() : impl_(new Impl) {}
NormalEquations::~NormalEquations() = default;

int64_t NormalEquations::factor_bytes(int32_t m) {
  return static_cast<int64_t>(m) * m * static_cast<int64_t>(sizeof(double));
}

int NormalEquations::init(int32_t m, int32_t n, const std::vector<int32_t>& row_ptr,
                          const std::vector<int32_t>& row_idx, const std::vector<double>& row_val,
                          const std::vector<int32_t>& col_ptr, const std::vector<int32_t>& col_idx,
                          const std::vector<double>& col_val) {
  Impl& d = *impl_;
  d.m = m;
  d.n = n;
  NE_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking), kStatusInitFailed);
  if (rocblas_create_handle(&d.blas) != rocblas_status_success) {
    set_last_error("rocblas_create_handle failed");
    return kStatusInitFailed;
  }
  NE_BLAS(rocblas_set_stream(d.blas, d.stream));
  // the caller's vectors are the sources of the queued copies: every return waits for them
  const StreamDrain drain{d.stream};
  int st = 0;
  if ((st = upload(row_ptr, &d.rp, d.stream)) || (st = upload(row_idx, &d.ri, d.stream)) ||
      (st = upload(row_val, &d.rv, d.stream)) || (st = upload(col_ptr, &d.cp, d.stream)) ||
      (st = upload(col_idx, &d.ci, d.stream)) || (st = upload(col_val, &d.cv, d.stream)))
    return st;
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.theta), std::max<size_t>(1, n) * sizeof(double)),
         kStatusOutOfMemory);
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.shift), std::max<size_t>(1, m) * sizeof(double)),
         kStatusOutOfMemory);
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.x), std::max<size_t>(1, m) * sizeof(double)),
         kStatusOutOfMemory);
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.M),
                   static_cast<size_t>(std::max<int64_t>(8, factor_bytes(m)))),
         kStatusOutOfMemory);
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.floored), sizeof(int64_t)), kStatusOutOfMemory);
  NE_HIP(hipStreamSynchronize(d.stream), kStatusSyncFailed);
  return kStatusSuccess;
}

int NormalEquations::factor(const double* theta, const double* shift, double floor,
                            int64_t* floored) {
  Impl& d = *impl_;
  const int32_t m = d.m;
  *floored = 0;
  if (m == 0) return kStatusSuccess;
  hipStream_t s = d.stream;
  const StreamDrain drain{s};  // theta and shift are copied asynchronously from the caller
  NE_HIP(hipMemcpyAsync(d.theta, theta, static_cast<size_t>(d.n) * sizeof(double),
                        hipMemcpyHostToDevice, s), kStatusCopyFailed);
  NE_HIP(hipMemcpyAsync(d.shift, shift, static_cast<size_t>(m) * sizeof(double),
                        hipMemcpyHostToDevice, s), kStatusCopyFailed);
  NE_HIP(hipMemsetAsync(d.floored, 0, sizeof(int64_t), s), kStatusCopyFailed);
  for (int32_t w0 = 0; w0 < m; w0 += kWindow) {
    const int32_t w1 = std::min(m, w0 + kWindow);
    const size_t lds = static_cast<size_t>(w1 - w0) * sizeof(double);
    k_assemble<<<static_cast<unsigned>(std::min(m, w1)), kLanes, lds, s>>>(
        m, d.rp, d.ri, d.rv, d.cp, d.ci, d.cv, d.theta, d.shift, w0, w1, d.M);
    NE_HIP(hipGetLastError(), kStatusLaunchFailed);
  }
  const double one = 1.0, minus_one = -1.0;
  for (int32_t kb = 0; kb < m; kb += kPanel) {
    const int32_t b = std::min(kPanel, m - kb);
    k_chol_block<<<1, kLanes, 0, s>>>(d.M, m, kb, b, floor, d.floored);
    NE_HIP(hipGetLastError(), kStatusLaunchFailed);
    const int32_t rest = m - kb - b;
    if (rest == 0) continue;
    double* l11 = d.M + static_cast<size_t>(kb) * m + kb;
    double* l21 = l11 + b;
    double* a22 = d.M + static_cast<size_t>(kb + b) * m + kb + b;
    NE_BLAS(rocblas_dtrsm(d.blas, rocblas_side_right, rocblas_fill_lower,
                          rocblas_operation_transpose, rocblas_diagonal_non_unit, rest, b, &one,
                          l11, m, l21, m));
    NE_BLAS(rocblas_dsyrk(d.blas, rocblas_fill_lower, rocblas_operation_none, rest, b,
                          &minus_one, l21, m, &one, a22, m));
  }
  NE_HIP(hipMemcpyAsync(floored, d.floored, sizeof(int64_t), hipMemcpyDeviceToHost, s),
         kStatusCopyFailed);
  NE_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  return kStatusSuccess;
}

int NormalEquations::solve(double* rhs) {
  Impl& d = *impl_;
  const int32_t m = d.m;
  if (m == 0) return kStatusSuccess;
  hipStream_t s = d.stream;
  const StreamDrain drain{s};  // rhs is the source and the target of asynchronous copies
  NE_HIP(hipMemcpyAsync(d.x, rhs, static_cast<size_t>(m) * sizeof(double), hipMemcpyHostToDevice, s),
         kStatusCopyFailed);
  NE_BLAS(rocblas_dtrsv(d.blas, rocblas_fill_lower, rocblas_operation_none,
                        rocblas_diagonal_non_unit, m, d.M, m, d.x, 1));
  NE_BLAS(rocblas_dtrsv(d.blas, rocblas_fill_lower, rocblas_operation_transpose,
                        rocblas_diagonal_non_unit, m, d.M, m, d.x, 1));
  NE_HIP(hipMemcpyAsync(rhs, d.x, static_cast<size_t>(m) * sizeof(double), hipMemcpyDeviceToHost, s),
         kStatusCopyFailed);
  NE_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  return kStatusSuccess;
}

}  // namespace sk
