// Device helpers shared by every HIP translation unit of the library: bound tests with the
// reference's semantics, projection, and deterministic wavefront/workgroup reductions, and
// the row-block sparse product used by all CSR kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device_types.hpp"

namespace sk {

// ---- small device helpers ------------------------------------------------------------------

__device__ __forceinline__ bool dev_finite_bound(double v) { return !(isinf(v) || fabs(v) >= 1e30); }
__device__ __forceinline__ double dev_max(double a, double b) { return a < b ? b : a; }  // std::max
__device__ __forceinline__ double dev_min(double a, double b) { return b < a ? b : a; }  // std::min

// Bounds reach the kernels normalised: an absent side is +-infinity, so a plain comparison
// projects (pdhg.cpp:63-67).
template <class T>
__device__ __forceinline__ T dev_project(T v, T lo, T hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}

struct D2 {
  double a, b;
};
__device__ __forceinline__ D2 operator+(D2 x, D2 y) { return D2{x.a + y.a, x.b + y.b}; }

__device__ __forceinline__ float shfl_xor(float v, int m) { return __shfl_xor(v, m, kWave); }
__device__ __forceinline__ double shfl_xor(double v, int m) { return __shfl_xor(v, m, kWave); }
__device__ __forceinline__ D2 shfl_xor(D2 v, int m) {
  return D2{__shfl_xor(v.a, m, kWave), __shfl_xor(v.b, m, kWave)};
}
template <class A>
__device__ __forceinline__ A zero_of() {
  return A{};
}

// Butterfly over the `width` lanes of a group (width a power of two <= 64). Every lane of
// the group ends with the same value; the order of additions is fixed.
template <class A>
__device__ __forceinline__ A group_sum(A v, int width) {
  for (int o = width >> 1; o >= 1; o >>= 1) v = v + shfl_xor(v, o);
  return v;
}

// Deterministic workgroup reductions; the result is valid in every thread.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  v = group_sum(v, kWave);
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < kBlock / kWave; ++w) r += lds[w];
  __syncthreads();
  return r;
}

// N workgroup reductions at once: sums (slots whose bit is clear in `max_mask`), each bit for
// bit a block_sum, and maxima (bit set: the same butterfly and wavefront order with max).
// Every wavefront reduces all N values, then one LDS exchange and one barrier instead of two
// barriers per value. The results are valid in every thread. `lds` holds N * (kBlock / kWave)
// doubles.
template <int N, uint32_t MAX_MASK>
__device__ __forceinline__ void block_reduce_n(double (&v)[N], double* lds) {
  constexpr uint32_t max_mask = MAX_MASK;
  constexpr int W = kBlock / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int o = kWave >> 1; o >= 1; o >>= 1) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double u = shfl_xor(v[i], o);
      v[i] = ((max_mask >> i) & 1u) ? dev_max(v[i], u) : v[i] + u;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) lds[i * W + wave] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if ((max_mask >> i) & 1u) {
      double r = lds[i * W];
#pragma unroll
      for (int w = 1; w < W; ++w) r = dev_max(r, lds[i * W + w]);
      v[i] = r;
    } else {
      double r = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) r += lds[i * W + w];
      v[i] = r;
    }
  }
  __syncthreads();
}

// Workgroups are dispatched round-robin over the 8 XCDs of the chip (workgroup b runs on
// XCD b % 8), and each XCD has its own L2. Tile tables are ordered by instance, so giving
// XCD x the x-th contiguous eighth of a table keeps each instance's tiles, and the slices of
// the vectors they gather from, in one L2. The map is a bijection of [0, n).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int x = b % kXcds, k = b / kXcds;
  const int per = n / kXcds, extra = n % kXcds;
  return x * per + min(x, extra) + k;
}

// Packed index words: (gather index - tile base) << 12 | CSR slot. Usable when every stream
// tile holds at most 4096 nonzeros and every instance has fewer than 2^20 rows and columns.
constexpr int kPackShift = 12;
constexpr int32_t kPackMaxLocal = 1 << 20;

// ---- the row-block product ---------------------------------------------------------------
// Computes acc_r = sum_p prod(p) for every row r of tile t of a CSR matrix with a workgroup of
// NT lanes (NT = kWave or kBlock) and calls epi(r, k, acc_r) exactly once per row, on one
// lane; that lane calls pre(r, k) first, before any product is loaded. k in
// [0, kMaxTileRows / NT) numbers the rows a lane finalises, so the caller can keep per-row
// operands in registers.
//
// Stream tile (kTileShort, kTileLong): all lanes first write the tile's products into LDS with
// coalesced loads (CSR-Stream; Greathouse & Daga, SC'14) through stager(t, stage), then each
// row is reduced from LDS in its fixed order (see kLongRow): one lane per short row (lane l
// takes rows l, l + NT, ...), one wavefront per long row. The stager writes the product of
// every staging position q in [t.p0, t.p1) to the CSR-order slot it belongs to within the
// tile (stage[0 .. p1 - p0)). The step kernels store a tile's nonzeros sorted by gather index,
// so the lanes of one load instruction gather from a few cache lines; the reduction reads the
// slots in CSR order.
// Vector tile (kTileVector, one row longer than the staging capacity): the row's staging
// order is its CSR order; wavefront 0 reduces prod(q), q in [p0, p1), straight from memory in
// the long-row order.
template <int NT, class A, class Prod, class Stager, class Pre, class Epi>
__device__ __forceinline__ void rowblock_product(const Tile t, const int32_t* __restrict__ ptr,
                                                 Prod prod, Stager stager, Pre pre, Epi epi,
                                                 A* stage) {
  static_assert(NT % kWave == 0 && kMaxTileRows % NT == 0, "NT must be a multiple of the wavefront");
  constexpr int R = kMaxTileRows / NT;  // short rows per lane; long rows per wavefront
  constexpr int W = NT / kWave;   // wavefronts
  const int tid = threadIdx.x;
  const int r0 = t.begin, nrows = t.end - t.begin;
  const int p0 = t.p0;
  const int p1 = t.p1;
  if (t.kind == kTileVector) {
    // Each product is rounded before it is added, as in the staged paths (products go through
    // LDS there): without this the compiler fuses a product into the running sum (one FMA per
    // unrolled batch), and a row would sum differently as a vector tile than as a long row.
#pragma clang fp contract(off)
    if (tid < kWave) {
      if (tid == 0) pre(r0, 0);
      A acc = zero_of<A>();
      constexpr int U = 4;
      for (int base = p0 + tid; base < p1; base += U * kWave) {
        // the U products' loads are issued together: a position past the row reads its last
        // nonzero and the product is replaced by zero
        A part[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int p = base + u * kWave;
          const A v = prod(p < p1 ? p : p1 - 1);
          part[u] = p < p1 ? v : zero_of<A>();
        }
        // in order: lane l adds nonzeros l, l + 64, l + 128, ... (adding a zero is exact)
#pragma unroll
        for (int u = 0; u < U; ++u) acc = acc + part[u];
      }
      acc = group_sum(acc, kWave);
      if (tid == 0) epi(r0, 0, acc);
    }
    return;
  }
  if (nrows <= 0) return;  // an instance without rows has one empty tile (same for every lane)
  const bool long_rows = t.kind == kTileLong;
  const int wave = tid / kWave, lane = tid % kWave;
  // the lane that finalises a row loads its epilogue operands first, so those loads are in
  // flight together with the product loads below; the row pointers are loaded whether or not
  // the lane has a row (a missing row reads the tile's last one), so no branch waits for them
  int b[R], e[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int row = long_rows ? wave + W * k : tid + NT * k;
    if (row < nrows && (!long_rows || lane == 0)) pre(r0 + row, k);
    const int rr = r0 + (row < nrows ? row : nrows - 1);
    const int pb = ptr[rr], pe = ptr[rr + 1];
    b[k] = row < nrows ? pb - p0 : 0;
    e[k] = row < nrows ? pe - p0 : 0;
  }
  stager(t, stage);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int row = long_rows ? wave + W * k : tid + NT * k;
    A acc = zero_of<A>();
    if (long_rows) {
      // `row` is the same in every lane of the wavefront
      if (row < nrows)
        for (int q = b[k] + lane; q < e[k]; q += kWave) acc = acc + stage[q];
      acc = group_sum(acc, kWave);
      if (row < nrows && lane == 0) epi(r0 + row, k, acc);
    } else if (row < nrows) {
      // in CSR order; the reads of a batch of 8 are all issued before the first add
      int q = b[k];
      for (; q + 8 <= e[k]; q += 8) {
        A v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = stage[q + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + v[u];
      }
      for (; q < e[k]; ++q) acc = acc + stage[q];
      epi(r0 + row, k, acc);
    }
  }
}

}  // namespace sk
