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

// Lane `src` of a wavefront's double, in every lane (a uniform source: two readlanes).
__device__ __forceinline__ double lane_value(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// In-place Cholesky of the b x b diagonal block at (kb, kb) of the column-major M (b <= 64),
// right-looking, by one wavefront: lane i keeps row i of the block in registers, and column k's
// pivot and entries reach the other lanes by readlane, so no step needs a barrier. Each entry
// sees the same operations in the same order as the textbook loop (its updates in column
// order, then the division by its column's pivot). A pivot below `floor` (zero and negative
// ones included) is raised to `floor` and counted; a NaN pivot stays NaN, so a broken system
// shows in the solution. Rows and columns past b are zero and never stored.
__global__ __launch_bounds__(kPanel) void k_chol_block(double* __restrict__ M, int32_t ld,
                                                        int32_t kb, int32_t b, double floor,
                                                        int64_t* __restrict__ floored) {
  static_assert(kPanel == 64, "one lane per row of the panel");
  const int i = threadIdx.x;
  double* A = M + static_cast<size_t>(kb) * ld + kb;
  double a[kPanel];
#pragma unroll
  for (int c = 0; c < kPanel; ++c) a[c] = (i < b && c < b && i >= c) ? A[static_cast<size_t>(c) * ld + i] : 0.0;
  int64_t count = 0;
#pragma unroll
  for (int k = 0; k < kPanel; ++k) {
    double d = lane_value(a[k], k);
    const bool low = d < floor;
    if (low) d = floor;
    const double s = sqrt(d);
    count += (k < b && low) ? 1 : 0;
    a[k] = i == k ? s : a[k] / s;
    // rows below k: a[c] -= L[i][k] * L[c][k] for c in (k, i]; entries above a lane's
    // diagonal are updated too and never stored
#pragma unroll
    for (int c = k + 1; c < kPanel; ++c) a[c] -= a[k] * lane_value(a[k], c);
  }
#pragma unroll
  for (int c = 0; c < kPanel; ++c)
    if (i < b && c < b && i >= c) A[static_cast<size_t>(c) * ld + i] = a[c];
  if (i == 0) *floored += count;
}

// Triangular solves with the factor (column-major lower-triangular L, leading dimension m):
// trans 0 solves L x = b, trans 1 solves L^T x = b, in place in x. One 256-lane workgroup per
// block of kTriRows rows, dispatched in the order the blocks depend on each other (first to
// last for L, last to first for L^T). A block folds in the contribution of each block it
// depends on, in a fixed order, as soon as that block publishes its part of the solution
// (flags[j] == epoch, set with an agent-scope release after the values); the factor's block
// for it is loaded before the wait. It then multiplies by the inverse of its diagonal block
// (k_tri_inverse, once per factorization) and publishes. A block only waits for blocks
// dispatched before it, which never wait for it, and each wait is bounded by spin_limit
// polls (a bound of zero: at its first wait, published or not); a block that runs out sets
// *failed and stops, and the host then redoes the solve with rocBLAS. Every launch uses a new
// epoch, so the flags are never cleared between solves.
constexpr int kTriRows = 64;

// Inverse of each kTriRows x kTriRows diagonal block of L (lower triangular, pivots already
// floored positive), by forward substitution of the identity, one column per lane and one
// workgroup per block; block k's inverse is column-major at inv + k * kTriRows^2.
__global__ __launch_bounds__(kLanes) void k_tri_inverse(const double* __restrict__ L, int32_t m,
                                                        int32_t first, double* __restrict__ inv) {
  __shared__ double D[kTriRows][kTriRows + 1];
  __shared__ double X[kTriRows][kTriRows + 1];  // X[i][j]: row i, column j of the inverse
  const int k = first + static_cast<int>(blockIdx.x), r0 = k * kTriRows;
  const int rows = min(kTriRows, m - r0), tid = threadIdx.x;
  for (int idx = tid; idx < kTriRows * kTriRows; idx += kLanes) {
    const int i = idx % kTriRows, c = idx / kTriRows;
    D[i][c] = (i < rows && c < rows && i >= c) ? L[static_cast<size_t>(r0 + c) * m + r0 + i] : 0.0;
    X[i][c] = 0.0;
  }
  __syncthreads();
  if (tid < rows) {
    const int j = tid;  // each lane its own column: no lane reads another's
    for (int i = j; i < rows; ++i) {
      double v = i == j ? 1.0 : 0.0;
      for (int c = j; c < i; ++c) v -= D[i][c] * X[c][j];
      X[i][j] = v / D[i][i];
    }
  }
  __syncthreads();
  double* out = inv + static_cast<size_t>(k) * kTriRows * kTriRows;
  for (int idx = tid; idx < kTriRows * kTriRows; idx += kLanes) {
    const int i = idx % kTriRows, c = idx / kTriRows;
    out[idx] = X[i][c];
  }
}

__global__ __launch_bounds__(kLanes) void k_trsv_blocks(const double* __restrict__ L,
                                                        const double* __restrict__ inv, int32_t m,
                                                        int32_t trans, double* x,
                                                        uint32_t* flags, uint32_t epoch,
                                                        uint32_t spin_limit, int32_t* failed) {
  constexpr int W = kLanes / 64;
  constexpr int CPW = kTriRows / W;  // columns (or rows of L^T) of a block per wavefront
  __shared__ double part[W][kTriRows];
  __shared__ double xs[kTriRows];
  __shared__ double rs[kTriRows];
  __shared__ double Di[kTriRows][kTriRows + 1];  // the diagonal block's inverse, Di[row][col]
  __shared__ int stop;
  const int nb = (m + kTriRows - 1) / kTriRows;
  const int k = trans ? nb - 1 - static_cast<int>(blockIdx.x) : static_cast<int>(blockIdx.x);
  const int r0 = k * kTriRows, rows = min(kTriRows, m - r0);
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  if (tid == 0) stop = 0;
  // what the end of the solve needs and no other block writes, loaded before any wait: the
  // inverse of the diagonal block and this block's part of b
  {
    const double* src = inv + static_cast<size_t>(k) * kTriRows * kTriRows;
    for (int idx = tid; idx < kTriRows * kTriRows; idx += kLanes) Di[idx % kTriRows][idx / kTriRows] = src[idx];
  }
  const double b = tid < rows ? x[r0 + tid] : 0.0;
  double acc = 0.0;
  const int jn = trans ? nb - 1 - k : k;
  for (int t = 0; t < jn; ++t) {
    const int j = trans ? nb - 1 - t : t;
    const int c0 = j * kTriRows, cols = min(kTriRows, m - c0);
    // the factor's block for j, before the wait: forward, lane = row of this block and the
    // wavefront's columns c = wave + W u of block j; backward, lane = row of L in block j
    // and the wavefront's rows c = wave + W u of this block (columns of L)
    double lv[CPW];
#pragma unroll
    for (int u = 0; u < CPW; ++u) {
      const int c = wave + W * u;
      lv[u] = 0.0;
      if (!trans) {
        if (lane < rows && c < cols) lv[u] = L[static_cast<size_t>(c0 + c) * m + r0 + lane];
      } else {
        if (lane < cols && c < rows) lv[u] = L[static_cast<size_t>(r0 + c) * m + c0 + lane];
      }
    }
    if (tid == 0) {
      uint32_t spins = 0;
      while (spin_limit == 0 ||
           __hip_atomic_load(flags + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (spin_limit == 0 || ++spins > spin_limit ||
          __hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          stop = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (stop) return;
    if (tid < cols) xs[tid] = __hip_atomic_load(x + c0 + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (!trans) {
#pragma unroll
      for (int u = 0; u < CPW; ++u) acc += lv[u] * xs[wave + W * u];
    } else {
      const double xl = xs[lane];
#pragma unroll
      for (int u = 0; u < CPW; ++u) {
        double v = lv[u] * xl;
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == wave + W * u) acc += v;
      }
    }
  }
  // combine the wavefronts' partial sums in wavefront order: r = b - sum
  part[wave][lane] = acc;
  __syncthreads();
  if (tid < kTriRows) {
    double sum = 0.0;
    for (int w = 0; w < W; ++w) sum += part[w][tid];
    rs[tid] = tid < rows ? b - sum : 0.0;
  }
  __syncthreads();
  // x_k = D^-1 r (forward) or D^-T r (backward): wavefront w takes the terms
  // c in [w CPW, (w + 1) CPW), then wavefront 0 adds the four partial sums in order
  {
    double v = 0.0;
    if (lane < rows) {
#pragma unroll
      for (int u = 0; u < CPW; ++u) {
        const int c = wave * CPW + u;
        v += (trans ? Di[c][lane] : Di[lane][c]) * rs[c];
      }
    }
    part[wave][lane] = v;
  }
  __syncthreads();
  if (wave == 0) {
    double v = 0.0;
    for (int w = 0; w < W; ++w) v += part[w][lane];
    // the block's solution, then its flag: the same wavefront stores both, so the release
    // orders every lane's value before the flag
    if (lane < rows) __hip_atomic_store(x + r0 + lane, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lane == 0) __hip_atomic_store(flags + k, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
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
  // the blocked triangular solves: per-block flags, the failure word, the last epoch used
  uint32_t* flags = nullptr;
  int32_t* failed = nullptr;
  double* inv = nullptr;  // inverses of the factor's diagonal blocks
  double* panel = nullptr;  // m x kPanel: a panel solve's product before it replaces the panel
  uint32_t epoch = 0;
  int32_t* failed_host = nullptr;  // pinned

  ~Impl() {
    for (void* p : {static_cast<void*>(rp), static_cast<void*>(ri), static_cast<void*>(cp),
                    static_cast<void*>(ci), static_cast<void*>(rv), static_cast<void*>(cv),
                    static_cast<void*>(theta), static_cast<void*>(shift), static_cast<void*>(M),
                    static_cast<void*>(x), static_cast<void*>(floored), static_cast<void*>(flags),
                    static_cast<void*>(failed), static_cast<void*>(inv),
                    static_cast<void*>(panel)})
      if (p) (void)hipFree(p);
    if (failed_host) (void)hipHostFree(failed_host);
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
  const int32_t nb = (m + kTriRows - 1) / kTriRows;
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.flags), std::max<size_t>(1, nb) * sizeof(uint32_t)),
         kStatusOutOfMemory);
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.failed), sizeof(int32_t)), kStatusOutOfMemory);
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.inv),
                   std::max<size_t>(1, nb) * kTriRows * kTriRows * sizeof(double)),
         kStatusOutOfMemory);
  NE_HIP(hipMalloc(reinterpret_cast<void**>(&d.panel),
                   std::max<size_t>(1, m) * kPanel * sizeof(double)),
         kStatusOutOfMemory);
  NE_HIP(hipHostMalloc(reinterpret_cast<void**>(&d.failed_host), sizeof(int32_t)), kStatusOutOfMemory);
  // epochs start at 1, so no flag holds one before its block publishes it
  NE_HIP(hipMemsetAsync(d.flags, 0, std::max<size_t>(1, nb) * sizeof(uint32_t), d.stream), kStatusCopyFailed);
  NE_HIP(hipMemsetAsync(d.failed, 0, sizeof(int32_t), d.stream), kStatusCopyFailed);
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
  static_assert(kPanel == kTriRows, "a panel is one block of the triangular solves");
  const double one = 1.0, zero = 0.0, minus_one = -1.0;
  for (int32_t kb = 0; kb < m; kb += kPanel) {
    const int32_t b = std::min(kPanel, m - kb);
    k_chol_block<<<1, kPanel, 0, s>>>(d.M, m, kb, b, floor, d.floored);
    NE_HIP(hipGetLastError(), kStatusLaunchFailed);
    // the inverse of the diagonal block: the panel solve below and the triangular solves use it
    k_tri_inverse<<<1, kLanes, 0, s>>>(d.M, m, kb / kPanel, d.inv);
    NE_HIP(hipGetLastError(), kStatusLaunchFailed);
    const int32_t rest = m - kb - b;
    if (rest == 0) continue;
    double* l21 = d.M + static_cast<size_t>(kb) * m + kb + b;
    double* a22 = d.M + static_cast<size_t>(kb + b) * m + kb + b;
    const double* inv11 = d.inv + static_cast<size_t>(kb / kPanel) * kPanel * kPanel;
    // L21 = A21 L11^-T (the panel solve as one product with the block's inverse), then the
    // product replaces the panel
    NE_BLAS(rocblas_dgemm(d.blas, rocblas_operation_none, rocblas_operation_transpose, rest, b, b,
                          &one, l21, m, inv11, kPanel, &zero, d.panel, rest));
    NE_HIP(hipMemcpy2DAsync(l21, static_cast<size_t>(m) * sizeof(double), d.panel,
                            static_cast<size_t>(rest) * sizeof(double),
                            static_cast<size_t>(rest) * sizeof(double), static_cast<size_t>(b),
                            hipMemcpyDeviceToDevice, s),
           kStatusCopyFailed);
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
  // forward then backward substitution with the blocked look-back kernel (k_trsv_blocks)
  const unsigned nb = static_cast<unsigned>((m + kTriRows - 1) / kTriRows);
  constexpr uint32_t kSpinLimit = 1u << 20;  // polls of one flag (~1 s), far above any wait
  const bool give_up = fault_solve_gives_up();  // injected (sk_fault_inject kind 4)
  for (int32_t trans = 0; trans < 2; ++trans) {
    if (++d.epoch == 0) ++d.epoch;
    k_trsv_blocks<<<nb, kLanes, 0, s>>>(d.M, d.inv, m, trans, d.x, d.flags, d.epoch,
                                        give_up ? 0u : kSpinLimit, d.failed);
    NE_HIP(hipGetLastError(), kStatusLaunchFailed);
  }
  NE_HIP(hipMemcpyAsync(d.failed_host, d.failed, sizeof(int32_t), hipMemcpyDeviceToHost, s), kStatusCopyFailed);
  NE_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  if (*d.failed_host) {
    // a block ran out of polls: the solve is redone with rocBLAS from the caller's rhs
    NE_HIP(hipMemsetAsync(d.failed, 0, sizeof(int32_t), s), kStatusCopyFailed);
    NE_HIP(hipMemcpyAsync(d.x, rhs, static_cast<size_t>(m) * sizeof(double), hipMemcpyHostToDevice, s),
           kStatusCopyFailed);
    NE_BLAS(rocblas_dtrsv(d.blas, rocblas_fill_lower, rocblas_operation_none,
                          rocblas_diagonal_non_unit, m, d.M, m, d.x, 1));
    NE_BLAS(rocblas_dtrsv(d.blas, rocblas_fill_lower, rocblas_operation_transpose,
                          rocblas_diagonal_non_unit, m, d.M, m, d.x, 1));
  }
  NE_HIP(hipMemcpyAsync(rhs, d.x, static_cast<size_t>(m) * sizeof(double), hipMemcpyDeviceToHost, s),
         kStatusCopyFailed);
  NE_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  return kStatusSuccess;
}

}  // namespace sk
