// HIP kernels of the batched restarted-PDHG engine for CDNA4 (gfx950).
//
// One iteration of the reference loop (src/pdhg/pdhg.cpp:383-496) is three launches:
//   k_step_primal  columns:   x' = proj(x - tau (c + A^T y)), xbar = 2x' - x       (:392-400)
//   k_step_dual    rows of A: y' = prox(y + sigma A xbar)                           (:402-408)
//   k_step_aty     rows of A^T: A^T y', interaction (x'-x)^T (A^T y' - A^T y)      (:410-438)
// and the adaptive step-size decision (:440-503), which the next step's primal launch makes
// first (every chunk of an instance makes it, the first one records it), or, before an
// evaluation stage and at the end of a captured block, its own launch:
//   k_step_finalize one workgroup per instance
// The carried A^T y replaces the reference's third product A (x'-x) (pdhg.cpp:432):
// (y'-y)^T A (x'-x) = (x'-x)^T (A^T y' - A^T y), and A^T y' is the next primal step's input.
//
// Every 40 accepted iterations (or on a step without information) an instance runs the
// FP64 evaluation stage (pdhg.cpp:502-643):
//   k_eval_prepare  unscale current and average candidates; ||x - x_restart||, ||y - y_restart||
//   k_eval_products eval_rows_tile: A x per candidate on the unscaled A (FP64): primal
//                   residual, row terms; eval_cols_tile: c + A^T y per candidate: dual
//                   residual, objectives (one launch for both)
//   k_eval_finalize one workgroup per instance: picks the candidate, tests termination,
//                   restarts
//   k_eval_commit   copy the chosen candidate into `best`, reset sums on restart
//
// Reductions are deterministic: per-tile partials are reduced in a fixed order by one
// workgroup per instance, so an instance gives bitwise the same trajectory alone or in a batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "device_types.hpp"
#include "device_util.hpp"

namespace sk {

// Every kernel of the engine rounds each product before it is added, as the reference's
// arithmetic does (a multiply, then an add: pdhg.cpp, scaling.cpp, sparse.cpp): the compiler
// may not fuse them into an FMA. It matters beyond the last bit: in the dual step
// y' = v - sigma proj(v / sigma) an FMA keeps the division's rounding residual, so a row whose
// bound is inactive gets a tiny nonzero y' where the reference gets exactly zero, and the
// step-size rule then sees information where the reference sees none.
#pragma clang fp contract(off)

// ---- kernel arguments ----------------------------------------------------------------------

// Everything a kernel of the engine may touch, passed by value. T is the iterate precision
// (float for the FP32/FP64 split engine, double for the plain FP64 engine).
template <class T>
struct Args {
  // CSR(A) and CSR(A^T): row pointers and CSR-order column indices
  const int32_t* __restrict__ a_ptr;
  const int32_t* __restrict__ a_idx;
  const int32_t* __restrict__ t_ptr;
  const int32_t* __restrict__ t_idx;
  // the scaled matrix A-hat for the step kernels, in each tile's staging order (sorted by
  // gather index; see rowblock_product): gather index, value in the iterate precision, and
  // the entry's CSR-order slot within its tile
  const int32_t* __restrict__ a_sidx;
  const T* __restrict__ a_val;
  const uint16_t* __restrict__ a_pos;
  const int32_t* __restrict__ t_sidx;
  const T* __restrict__ t_val;
  const uint16_t* __restrict__ t_pos;
  // the original A in the same two layouts, FP64 (evaluation): CSR order, and the staging
  // order of the step kernels' tiles
  const double* __restrict__ a_orig;
  const double* __restrict__ t_orig;
  const double* __restrict__ a_orig_st;
  const double* __restrict__ t_orig_st;
  // scaled problem data in the iterate precision; absent bounds are +-inf
  const T* __restrict__ chat;
  const T* __restrict__ lhat;
  const T* __restrict__ uhat;
  const T* __restrict__ rlhat;
  const T* __restrict__ ruhat;
  // original problem data, FP64
  const double* __restrict__ cmin;  // sense * cost
  const double* __restrict__ cl;
  const double* __restrict__ cu;
  const double* __restrict__ rl;
  const double* __restrict__ ru;
  const double* __restrict__ dc;  // column scaling D_c
  const double* __restrict__ dr;  // row scaling D_r
  // iterates (scaled), double-buffered by instance parity
  T* x0;
  T* x1;
  T* y0;
  T* y1;
  T* aty0;
  T* aty1;
  T* xbar;
  T* xrst;  // restart point
  T* yrst;
  // running sums of the accepted iterates since the last restart. The step kernels add in the
  // iterate precision; with FP32 iterates every evaluation folds that sum into the FP64 total
  // (xtot, ytot) and clears it, so FP32 only ever sums the iterates between two evaluations
  // and the average of a long restart window keeps FP64 accuracy. With FP64 iterates the sum
  // is the FP64 running sum itself and the totals are unused.
  T* xsum;
  T* ysum;
  double* xtot;
  double* ytot;
  // unscaled candidates: *u0 current, *u1 running average; best; report outputs
  double* xu0;
  double* xu1;
  double* yu0;
  double* yu1;
  double* bestx;
  double* besty;
  double* activity;
  double* reduced;
  // tile tables and per-instance records
  const Tile* __restrict__ tc;  // column chunks
  const Tile* __restrict__ tr;  // row chunks
  const Tile* __restrict__ ta;  // row blocks of CSR(A)
  const Tile* __restrict__ tt;  // row blocks of CSR(A^T)
  const InstDesc* __restrict__ desc;
  // per column chunk, its instance's descriptor: a deciding primal launch loads it beside the
  // chunk's tile record instead of after it
  const InstDesc* __restrict__ cdesc;
  InstState* state;
  // where a decision writes the instance records: the primal launch that decides the previous
  // step (decide = 1) reads `state` and writes here; the decision launch (k_step_finalize)
  // reads `state` and writes here (the same buffer or the other one of the pair)
  InstState* state_next;
  // per-tile flags (kFlag*) of ta, tt, tc and tr
  uint32_t* fa;
  uint32_t* ft;
  uint32_t* fc;
  uint32_t* fr;
  // per-tile partials
  double* ps_c;  // step: ||x'-x||^2 per column chunk (this step's primal launch writes it)
  const double* ps_c_prev;  // the previous step's, read by a deciding primal launch
  double* ps_a;  // step: ||y'-y||^2 per CSR(A) block
  double* ps_t;  // step: interaction per CSR(A^T) block
  double* pe_c;  // eval: 2 per column chunk
  double* pe_r;  // eval: 2 per row chunk
  double* pe_a;  // eval: 8 per CSR(A) block
  double* pe_t;  // eval: 8 per CSR(A^T) block
  uint32_t* done_count;
  // the step-size rule's factors for iterations [0, rule_len), two per iteration (shrink,
  // grow), computed once on the host with the reference's formula
  const double* __restrict__ rule;
  int32_t rule_len;
  int32_t n_tc;    // number of column-chunk tiles (eval kernels cover tc then tr)
  int32_t n_ta;    // row blocks of CSR(A)
  int32_t n_tt;    // row blocks of CSR(A^T)
  int32_t force;   // timing mode: every instance steps, nothing is decided
  int32_t report;  // evaluation kernels: 0 in-loop candidates, 1 final report of best
  int32_t decide;  // primal launch: first make the previous step's decision (see k_step_primal)
  // Column slices of A (a batch of one instance): A x-bar as `slices` partial products, one
  // per slice of the columns, then their sum and the dual step per row chunk (k_slice_product,
  // k_slice_dual). S row s * slice_rows + i holds row i's entries in slice s; ts is the row-block
  // table of S, laid out so that XCD s runs slice s; xpart holds S's row sums; ps_r the row
  // chunks' ||y'-y||^2 partials, which replace ps_a for such a batch.
  const Tile* __restrict__ ts;
  const int32_t* __restrict__ s_ptr;
  const int32_t* __restrict__ s_sidx;
  const T* __restrict__ s_val;
  const uint16_t* __restrict__ s_pos;
  T* xpart;
  double* ps_r;
  int32_t n_ts;
  int32_t slices;      // 0: A x-bar by k_step_dual
  int32_t slice_rows;
  // Row slices of A (the same batches): A^T y' as `row_slices` partial products, one per range
  // of A's rows (y' entries), then their sum and the interaction per column chunk
  // (k_rslice_product, k_rslice_aty). Q row k * slice_cols + j holds column j's entries in row
  // slice k; tq is its row-block table, laid out so that XCDs 2k and 2k + 1 run slice k; ypart
  // holds Q's row sums; ps_q the column chunks' interaction partials, which replace ps_t.
  const Tile* __restrict__ tq;
  const int32_t* __restrict__ q_ptr;
  const int32_t* __restrict__ q_sidx;
  const T* __restrict__ q_val;
  const uint16_t* __restrict__ q_pos;
  T* ypart;
  double* ps_q;
  int32_t n_tq;
  int32_t row_slices;  // 0: A^T y' by k_step_aty
  int32_t slice_cols;
};

__device__ __forceinline__ bool stepping(const InstState& s) { return !s.done && !s.eval_pending; }

__device__ __forceinline__ uint32_t step_flags(const InstState& s) {
  return (stepping(s) ? kFlagStepping : 0u) | (s.parity ? kFlagParity : 0u) |
         (s.pending_accept ? kFlagAccept : 0u) | (s.eval_pending ? kFlagEval : 0u) |
         (s.report_pending ? kFlagReport : 0u) |
         ((s.copy_best || s.do_restart) ? kFlagCommit : 0u);
}

// Writes `f` into the step flags of every tile of the instance (whole workgroup).
template <class T>
__device__ __forceinline__ void publish_flags(const Args<T>& a, const InstDesc& d, uint32_t f) {
  for (int k = d.ta0 + threadIdx.x; k < d.ta1; k += kBlock) a.fa[k] = f;
  for (int k = d.tt0 + threadIdx.x; k < d.tt1; k += kBlock) a.ft[k] = f;
  for (int k = d.tc0 + threadIdx.x; k < d.tc1; k += kBlock) a.fc[k] = f;
  for (int k = d.tr0 + threadIdx.x; k < d.tr1; k += kBlock) a.fr[k] = f;
}

// Set-up: the first flags of every instance, one workgroup per instance.
template <class T>
__global__ __launch_bounds__(kBlock) void k_publish_flags(Args<T> a) {
  publish_flags(a, a.desc[blockIdx.x], step_flags(a.state[blockIdx.x]));
}

// ---- step kernels ---------------------------------------------------------------------------

// Columns per lane of a column chunk: lane l of the workgroup takes columns
// t.begin + u * kBlock + l, u in [0, kChunkCols).
constexpr int kChunkCols = kChunk / kBlock;

// The primal update of column chunk t from operands already loaded (pdhg.cpp:393-400): x' and
// x-bar from buffer p's iterate, the running sum when the last step was accepted, and the
// chunk's ||x'-x||^2 partial.
template <class T>
__device__ __forceinline__ void primal_update(const Args<T>& a, const Tile& t, const int tile,
                                              const int p, const bool acc, const T tau,
                                              const T (&xj)[kChunkCols], const T (&c)[kChunkCols],
                                              const T (&at)[kChunkCols], const T (&lo)[kChunkCols],
                                              const T (&hi)[kChunkCols], const T (&sm)[kChunkCols],
                                              double* red) {
  T* __restrict__ xn = p ? a.x0 : a.x1;
  double d2 = 0.0;
#pragma unroll
  for (int u = 0; u < kChunkCols; ++u) {
    const int j = t.begin + u * kBlock + threadIdx.x;
    if (j >= t.end) continue;
    if (acc) a.xsum[j] = sm[u] + xj[u];
    const T g = c[u] + at[u];
    const T v = dev_project<T>(xj[u] - tau * g, lo[u], hi[u]);
    xn[j] = v;
    a.xbar[j] = static_cast<T>(2) * v - xj[u];
    const double d = static_cast<double>(v - xj[u]);
    d2 += d * d;
  }
  d2 = block_sum(d2, red);
  if (threadIdx.x == 0) a.ps_c[tile] = d2;
}

// The primal step of column chunk `tile` (pdhg.cpp:393-400) and its ||x'-x||^2 partial, with
// the step control the last decision published (the chunk's flags and the instance record).
template <class T>
__device__ __forceinline__ void primal_chunk(const Args<T>& a, const int tile, double* red) {
  const Tile t = a.tc[tile];
  const uint32_t f = a.fc[tile];
  if (!a.force && !(f & kFlagStepping)) return;
  const int p = (f & kFlagParity) ? 1 : 0;
  const bool acc = a.force || (f & kFlagAccept);
  const InstState& s = a.state[t.inst];
  const double eta = s.eta, omega = s.omega;
  const T* __restrict__ x = p ? a.x1 : a.x0;
  const T* __restrict__ aty = p ? a.aty1 : a.aty0;
  // every operand load is issued before the step size (which waits for the instance record)
  // is used
  constexpr int U = kChunkCols;
  T xj[U], c[U], at[U], lo[U], hi[U], sm[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = t.begin + u * kBlock + threadIdx.x;
    if (j < t.end) {
      xj[u] = x[j];
      c[u] = a.chat[j];
      at[u] = aty[j];
      lo[u] = a.lhat[j];
      hi[u] = a.uhat[j];
      if (acc) sm[u] = a.xsum[j];
    }
  }
  primal_update(a, t, tile, p, acc, static_cast<T>(eta / omega), xj, c, at, lo, hi, sm, red);
}

// Gather index of staging position q: a packed word holds (local index << 12 | CSR slot),
// relative to the tile's base; otherwise separate index and 16-bit slot arrays.
template <bool PACK>
__device__ __forceinline__ int gather_index(const Tile& t, const int32_t* __restrict__ sidx, int q) {
  return PACK ? t.base + static_cast<int>(static_cast<uint32_t>(sidx[q]) >> kPackShift) : sidx[q];
}

// Four consecutive elements from a 16-byte aligned address (two 16-byte loads for FP64).
__device__ __forceinline__ void load4(const float* __restrict__ p, float v[4]) {
  const float4 w = *reinterpret_cast<const float4*>(p);
  v[0] = w.x, v[1] = w.y, v[2] = w.z, v[3] = w.w;
}
__device__ __forceinline__ void load4(const double* __restrict__ p, double v[4]) {
  const double2 w0 = *reinterpret_cast<const double2*>(p);
  const double2 w1 = *reinterpret_cast<const double2*>(p + 2);
  v[0] = w0.x, v[1] = w0.y, v[2] = w1.x, v[3] = w1.y;
}
__device__ __forceinline__ void load4(const int32_t* __restrict__ p, int v[4]) {
  const int4 w = *reinterpret_cast<const int4*>(p);
  v[0] = w.x, v[1] = w.y, v[2] = w.z, v[3] = w.w;
}
__device__ __forceinline__ void load4(const uint16_t* __restrict__ p, int v[4]) {
  const ushort4 w = *reinterpret_cast<const ushort4*>(p);
  v[0] = w.x, v[1] = w.y, v[2] = w.z, v[3] = w.w;
}

// Stager of rowblock_product over the staging arrays, in groups of four stored positions: each
// lane loads four index words and four staged values with one 16-byte load per array, so a
// wavefront reads 1 KB of each contiguously, then stages f(c, slot, v) for each of the four
// entries (c its gather index, slot its CSR-order slot, v its staged value). The stored order
// (group_rank) makes the e-th gathers of a wavefront's lanes consecutive sorted indices.
// Groups start at multiples of four, so a tile's first and last groups also cover positions of
// the neighbouring tiles, which are loaded and ignored; the staging arrays hold 16-byte aligned
// storage for at least kStagePad more elements than nonzeros, so the last group of the last
// tile stays inside them.
// The four gathers of a group are issued together: an entry outside the tile gathers at the
// tile's base (the first gather index of its instance, which exists because the tile has a
// nonzero) and its product is dropped, so no gather waits in a branch of its own for the one
// before it.
// IN_FLIGHT: groups each lane loads before it gathers (0: the measured best for a launch of
// many workgroups, one for a 256-lane workgroup and four for a wavefront); a workgroup that
// runs alone (k_resident) issues a whole tile's groups at once. Only the order of the loads
// changes, not what is staged.
template <int NT, class A, bool PACK, class V, class F, int IN_FLIGHT = 0>
struct GroupStager {
  const int32_t* __restrict__ sidx;
  const uint16_t* __restrict__ pos;
  const V* __restrict__ val;
  F f;
  __device__ __forceinline__ void operator()(const Tile& t, A* stage) const {
    const int p0 = t.p0, p1 = t.p1, qa = p0 & ~3;
    if (p1 <= p0) return;
    const int groups = (p1 - qa + 3) >> 2;
    constexpr int U = IN_FLIGHT > 0 ? IN_FLIGHT : (NT >= kBlock ? 1 : 4);
    for (int g0 = 0; g0 < groups; g0 += U * NT) {
      int w[U][4], sl[U][4];
      V v[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int g = g0 + u * NT + static_cast<int>(threadIdx.x);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          w[u][e] = 0;
          sl[u][e] = 0;
          v[u][e] = V{};
        }
        if (g < groups) {
          load4(sidx + qa + 4 * g, w[u]);
          load4(val + qa + 4 * g, v[u]);
          if (!PACK) load4(pos + qa + 4 * g, sl[u]);
        }
      }
      A prod[U][4];
      bool own[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int g = g0 + u * NT + static_cast<int>(threadIdx.x);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int q = qa + 4 * g + e;
          own[u][e] = g < groups && q >= p0 && q < p1;
          const uint32_t word = static_cast<uint32_t>(w[u][e]);
          const int c = PACK ? t.base + static_cast<int>(word >> kPackShift) : w[u][e];
          if (PACK) sl[u][e] = static_cast<int>(word & 0xFFFu);
          prod[u][e] = f(own[u][e] ? c : t.base, sl[u][e], v[u][e]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (own[u][e]) stage[sl[u][e]] = prod[u][e];
        }
      }
    }
  }
};
// stages val[q] * x[c] (products in A) with NT lanes
template <int NT, class A, bool PACK, class V, class X, int IN_FLIGHT = 0>
__device__ __forceinline__ auto gather_stager(const int32_t* sidx, const uint16_t* pos,
                                              const V* val, const X* x) {
  auto f = [x](int c, int, V v) { return static_cast<A>(v) * static_cast<A>(x[c]); };
  return GroupStager<NT, A, PACK, V, decltype(f), IN_FLIGHT>{sidx, pos, val, f};
}

// CAP: the tile capacity the tables were built for, which sizes the LDS staging buffer (and
// with it how many workgroups a CU holds). PACK: index words are packed (see gather_index).
// kStepLanes: lanes per workgroup of the two sparse step kernels (one tile per workgroup);
// kStepRows: the rows each lane finalises.
constexpr int kStepLanes = kBlock;
constexpr int kStepRows = kMaxTileRows / kStepLanes;

// The sparse product of the two step kernels on one row-block tile: acc_r = sum over row r of
// val[q] * in[gather index of q], with the tile's nonzeros staged in LDS in its staging order
// and each row reduced in its fixed order (rowblock_product; pre and epi as there).
// k_step_dual runs it on CSR(A) with in = x-bar, k_step_aty on CSR(A^T) with in = y', and
// k_probe_product on either table with a caller's vector.
template <class T, bool PACK, int IN_FLIGHT = 0, class Pre, class Epi>
__device__ __forceinline__ void step_product(const Tile& t, const int32_t* __restrict__ ptr,
                                             const int32_t* __restrict__ sidx,
                                             const uint16_t* __restrict__ pos,
                                             const T* __restrict__ val, const T* __restrict__ in,
                                             Pre pre, Epi epi, T* stage) {
  rowblock_product<kStepLanes, T>(
      t, ptr, [&](int q) { return val[q] * in[gather_index<PACK>(t, sidx, q)]; },
      gather_stager<kStepLanes, T, PACK, T, T, IN_FLIGHT>(sidx, pos, val, in), pre, epi, stage);
}

// The dual step of row i from r = (A-hat x-bar)_i (pdhg.cpp:402-408): y' into yn, the running
// sum when the last step was accepted, and (y'_i - y_i)^2 added to d2.
template <class T>
__device__ __forceinline__ void dual_row(const Args<T>& a, const int i, const T yi, const T lo,
                                         const T hi, const T sum, const bool acc, const T sigma,
                                         const T r, T* __restrict__ yn, double& d2) {
  if (acc) a.ysum[i] = sum + yi;
  const T v = yi + sigma * r;
  const T ynew = v - sigma * dev_project<T>(v / sigma, lo, hi);
  yn[i] = ynew;
  const double d = static_cast<double>(ynew - yi);
  d2 += d * d;
}

// A x-bar and the dual step on the rows of tile `tile` of A (pdhg.cpp:402-408), with its
// ||y'-y||^2 partial.
template <class T, bool PACK, int IN_FLIGHT = 0>
__device__ __forceinline__ void dual_tile(const Args<T>& a, const int tile, T* stage, double* red) {
  const Tile t = a.ta[tile];
  const uint32_t f = a.fa[tile];
  if (!a.force && !(f & kFlagStepping)) return;
  const int p = (f & kFlagParity) ? 1 : 0;
  const bool acc = a.force || (f & kFlagAccept);
  // the step size is formed in the epilogue, so waiting for the instance record does not
  // hold back the staging loads
  const InstState& s = a.state[t.inst];
  const double eta = s.eta, omega = s.omega;
  const T* __restrict__ y = p ? a.y1 : a.y0;
  T* __restrict__ yn = p ? a.y0 : a.y1;
  double d2 = 0.0;
  T yi[kStepRows], lo[kStepRows], hi[kStepRows], sum[kStepRows];
  step_product<T, PACK, IN_FLIGHT>(
      t, a.a_ptr, a.a_sidx, a.a_pos, a.a_val, a.xbar,
      [&](int i, int k) {
        yi[k] = y[i];
        lo[k] = a.rlhat[i];
        hi[k] = a.ruhat[i];
        sum[k] = acc ? a.ysum[i] : T(0);
      },
      [&](int i, int k, T r) {
        const T sigma = static_cast<T>(eta * omega);
        dual_row(a, i, yi[k], lo[k], hi[k], sum[k], acc, sigma, r, yn, d2);
      },
      stage);
  d2 = block_sum(d2, red);
  if (threadIdx.x == 0) a.ps_a[tile] = d2;
}

template <class T, int CAP, bool PACK>
__global__ __launch_bounds__(kStepLanes) void k_step_dual(Args<T> a) {
  __shared__ T stage[CAP];
  __shared__ double red[kStepLanes / kWave];
  dual_tile<T, PACK>(a, xcd_tile(blockIdx.x, gridDim.x), stage, red);
}

// Probe of the step kernels' product (tests): out[r] = step_product's sum for every row of one
// row-block table, with the staging arrays the step kernels read (CSR(A) with in = x, or
// CSR(A^T) with in = y) and nothing else computed.
template <class T, int CAP, bool PACK>
__global__ __launch_bounds__(kStepLanes) void k_probe_product(const Tile* __restrict__ tiles,
                                                              const int32_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ sidx,
                                                              const uint16_t* __restrict__ pos,
                                                              const T* __restrict__ val,
                                                              const T* __restrict__ in,
                                                              T* __restrict__ out) {
  __shared__ T stage[CAP];
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  step_product<T, PACK>(
      t, ptr, sidx, pos, val, in, [](int, int) {}, [&](int r, int, T acc) { out[r] = acc; },
      stage);
}

// A lane's share of a table of per-tile partials: acc + part[k * stride] for k = threadIdx.x,
// threadIdx.x + kBlock, ... below count, added in that order. The loads are issued kSumBatch
// at a time before their adds (a plain loop waits for each load before issuing the next), so a
// table of count partials costs about count / (kBlock * kSumBatch) dependent memory round trips
// instead of count / kBlock; the sum is bit for bit the plain loop's. A position past the table
// reloads its last entry and is not added.
constexpr int kSumBatch = 8;
__device__ __forceinline__ double lane_sum(const double* __restrict__ part, int count,
                                           int stride, double acc) {
  for (int k0 = static_cast<int>(threadIdx.x); k0 < count; k0 += kSumBatch * kBlock) {
    double w[kSumBatch];
#pragma unroll
    for (int u = 0; u < kSumBatch; ++u) {
      const int k = k0 + u * kBlock;
      w[u] = part[static_cast<size_t>(k < count ? k : count - 1) * stride];
    }
#pragma unroll
    for (int u = 0; u < kSumBatch; ++u) acc = k0 + u * kBlock < count ? acc + w[u] : acc;
  }
  return acc;
}

// A x-bar over column slices: the row sums of one tile of S (the stacked slices) into xpart.
// XCD s runs the tiles of slice s (xcd_tile over a table of equal slice parts), so the entries
// of x-bar its gathers touch, one slice's columns, stay in that XCD's L2.
template <class T, int CAP, bool PACK>
__global__ __launch_bounds__(kStepLanes) void k_slice_product(Args<T> a) {
  __shared__ T stage[CAP];
  // the batch holds one instance, and every tile of an instance carries the same flags
  if (!a.force && !(a.fa[0] & kFlagStepping)) return;
  const Tile t = a.ts[xcd_tile(blockIdx.x, gridDim.x)];
  step_product<T, PACK>(
      t, a.s_ptr, a.s_sidx, a.s_pos, a.s_val, a.xbar, [](int, int) {},
      [&](int r, int, T acc) { a.xpart[r] = acc; }, stage);
}

// Row i of A x-bar from the slices' partials, added in slice order.
template <class T>
__device__ __forceinline__ T slice_sum(const T* __restrict__ part, const int i, const int rows,
                                       const int slices) {
  T r = part[i];
  for (int k = 1; k < slices; ++k) r = r + part[static_cast<size_t>(k) * rows + i];
  return r;
}

// The dual step of row chunk `tile` from the slices' partials (pdhg.cpp:402-408), with the
// chunk's ||y'-y||^2 partial in ps_r.
template <class T>
__global__ __launch_bounds__(kBlock) void k_slice_dual(Args<T> a) {
  __shared__ double red[kBlock / kWave];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t f = a.fr[tile];
  if (!a.force && !(f & kFlagStepping)) return;
  const Tile t = a.tr[tile];
  const int p = (f & kFlagParity) ? 1 : 0;
  const bool acc = a.force || (f & kFlagAccept);
  const T* __restrict__ y = p ? a.y1 : a.y0;
  T* __restrict__ yn = p ? a.y0 : a.y1;
  constexpr int U = kChunk / kBlock;
  T yi[U], lo[U], hi[U], sum[U], r[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = t.begin + u * kBlock + static_cast<int>(threadIdx.x);
    if (i < t.end) {
      yi[u] = y[i];
      lo[u] = a.rlhat[i];
      hi[u] = a.ruhat[i];
      sum[u] = acc ? a.ysum[i] : T(0);
      r[u] = slice_sum(a.xpart, i, a.slice_rows, a.slices);
    }
  }
  const InstState& s = a.state[t.inst];
  const T sigma = static_cast<T>(s.eta * s.omega);
  double d2 = 0.0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = t.begin + u * kBlock + static_cast<int>(threadIdx.x);
    if (i < t.end) dual_row(a, i, yi[u], lo[u], hi[u], sum[u], acc, sigma, r[u], yn, d2);
  }
  d2 = block_sum(d2, red);
  if (threadIdx.x == 0) a.ps_r[tile] = d2;
}

// Probe of the sliced product (tests): out[i] = row i's sum of the slices' partials.
template <class T>
__global__ __launch_bounds__(kBlock) void k_slice_sum(const T* __restrict__ part,
                                                      T* __restrict__ out, int32_t rows,
                                                      int32_t slices) {
  const int i = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
  if (i < rows) out[i] = slice_sum(part, i, rows, slices);
}

// Three lane_sums whose loads are issued together, B per table per round, so three short
// tables cost one memory round trip instead of three; each sum bit for bit its lane_sum. Every
// count is at least 1 (every instance has a tile in each table).
template <int B>
__device__ __forceinline__ void lane_sums3(const double* __restrict__ p0, int n0,
                                           const double* __restrict__ p1, int n1,
                                           const double* __restrict__ p2, int n2, double* v) {
  const int most = max(n0, max(n1, n2));
  for (int k0 = static_cast<int>(threadIdx.x); k0 < most; k0 += B * kBlock) {
    double w0[B], w1[B], w2[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int k = k0 + u * kBlock;
      w0[u] = p0[k < n0 ? k : n0 - 1];
      w1[u] = p1[k < n1 ? k : n1 - 1];
      w2[u] = p2[k < n2 ? k : n2 - 1];
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int k = k0 + u * kBlock;
      v[0] = k < n0 ? v[0] + w0[u] : v[0];
      v[1] = k < n1 ? v[1] + w1[u] : v[1];
      v[2] = k < n2 ? v[2] + w2[u] : v[2];
    }
  }
}

// Reduce `count` per-tile partials starting at tile `base` with the whole workgroup in a
// fixed order (result valid in every thread).
__device__ __forceinline__ double tile_sum(const double* part, int base, int count, int stride,
                                           int slot, double* red) {
  const double v = lane_sum(part + static_cast<size_t>(base) * stride + slot, count, stride, 0.0);
  return block_sum(v, red);
}

// `done_count` is null in a workgroup that decides only for its own use (a deciding primal
// chunk other than the instance's first): the instance is counted once, by the writer.
__device__ __forceinline__ void finish(InstState& s, int status, uint32_t* done_count) {
  s.done = 1;
  s.status = status;
  s.report_pending = 1;
  if (done_count) atomicAdd(done_count, 1u);
}

// The shrink and grow factors of the step-size rule at iteration k (pdhg.cpp:472-474): from
// the table for k below rule_len (which covers the default limit of 10^6 iterations), else
// computed here with the device's pow.
template <class T>
__device__ __forceinline__ void rule_factors(const Args<T>& a, const int64_t k, double* f) {
  if (k < a.rule_len) {
    f[0] = a.rule[2 * k];
    f[1] = a.rule[2 * k + 1];
  } else {
    const double exponent = static_cast<double>(k + 1 > 2 ? k + 1 : 2);
    f[0] = 1.0 - pow(exponent, -0.3);
    f[1] = 1.0 + pow(exponent, -0.6);
  }
}

// The adaptive step-size rule and acceptance (pdhg.cpp:410-503); shrink and grow are the
// rule's factors at the current iteration (rule_factors).
__device__ void decide_step(InstState& s, double dx2, double dy2, double inter_sum, int has_rows,
                            double shrink, double grow, uint32_t* done_count) {
  const double omega = s.omega;
  const double movement = 0.5 * omega * dx2 + 0.5 * dy2 / omega;
  const double interaction = has_rows ? fabs(inter_sum) : 0.0;
  s.movement = movement;
  s.interaction = interaction;
  s.attempts += 1;
  s.copy_best = 0;
  s.do_restart = 0;
  if (!isfinite(movement) || !isfinite(interaction)) {
    s.nonfinite = 1;
    finish(s, kStNumericalError, done_count);
    return;
  }
  const bool no_information = interaction <= 0.0;
  const double limit = no_information ? INFINITY : movement / interaction;
  const double proposed = dev_min(shrink * limit, grow * s.eta);
  if (s.eta <= limit) {
    s.parity ^= 1;
    s.averaged += 1;
    s.iteration += 1;
    s.pending_accept = 1;
  } else {
    s.pending_accept = 0;
  }
  if (!no_information) {
    s.eta = proposed < 1e-12 ? 1e-12 : (s.eta_ceiling < proposed ? s.eta_ceiling : proposed);
  }
  if (s.iteration != 0 && (s.iteration % 40 == 0 || no_information)) {
    s.eval_pending = 1;
  } else if (s.iteration >= s.iteration_limit) {
    finish(s, kStIterationLimit, done_count);
  }
}

// A^T y' on the rows of tile `tile` of A^T and its share of the interaction (pdhg.cpp:410-438).
template <class T, bool PACK, int IN_FLIGHT = 0>
__device__ __forceinline__ void aty_tile(const Args<T>& a, const int tile, T* stage, double* red) {
  const Tile t = a.tt[tile];
  const uint32_t f = a.ft[tile];
  if (!a.force && !(f & kFlagStepping)) return;
  const int p = (f & kFlagParity) ? 1 : 0;
  const T* __restrict__ x = p ? a.x1 : a.x0;
  const T* __restrict__ xn = p ? a.x0 : a.x1;
  const T* __restrict__ yn = p ? a.y0 : a.y1;
  const T* __restrict__ aty = p ? a.aty1 : a.aty0;
  T* __restrict__ atyn = p ? a.aty0 : a.aty1;
  double inter = 0.0;
  T xj[kStepRows], xnj[kStepRows], atyj[kStepRows];
  step_product<T, PACK, IN_FLIGHT>(
      t, a.t_ptr, a.t_sidx, a.t_pos, a.t_val, yn,
      [&](int j, int k) {
        xj[k] = x[j];
        xnj[k] = xn[j];
        atyj[k] = aty[j];
      },
      [&](int j, int k, T r) {
        atyn[j] = r;
        inter += static_cast<double>(xnj[k] - xj[k]) * static_cast<double>(r - atyj[k]);
      },
      stage);
  inter = block_sum(inter, red);
  if (threadIdx.x == 0) a.ps_t[tile] = inter;
}

template <class T, int CAP, bool PACK>
__global__ __launch_bounds__(kStepLanes) void k_step_aty(Args<T> a) {
  __shared__ T stage[CAP];
  __shared__ double red[kStepLanes / kWave];
  aty_tile<T, PACK>(a, xcd_tile(blockIdx.x, gridDim.x), stage, red);
}

// A^T y' over row slices: the row sums of one tile of Q (the stacked row slices of CSR(A^T))
// into ypart. XCDs 2k and 2k + 1 run the tiles of slice k, so the entries of y' their gathers
// touch, one slice's rows, stay in their L2.
template <class T, int CAP, bool PACK>
__global__ __launch_bounds__(kStepLanes) void k_rslice_product(Args<T> a) {
  __shared__ T stage[CAP];
  // the batch holds one instance, and every tile of an instance carries the same flags
  const uint32_t f = a.ft[0];
  if (!a.force && !(f & kFlagStepping)) return;
  const T* __restrict__ yn = (f & kFlagParity) ? a.y0 : a.y1;
  const Tile t = a.tq[xcd_tile(blockIdx.x, gridDim.x)];
  step_product<T, PACK>(
      t, a.q_ptr, a.q_sidx, a.q_pos, a.q_val, yn, [](int, int) {},
      [&](int r, int, T acc) { a.ypart[r] = acc; }, stage);
}

// A^T y' of column chunk `tile` from the row slices' partials, added in slice order, and the
// chunk's share of the interaction (x'-x)^T (A^T y' - A^T y) in ps_q (pdhg.cpp:410-438).
template <class T>
__global__ __launch_bounds__(kBlock) void k_rslice_aty(Args<T> a) {
  __shared__ double red[kBlock / kWave];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t f = a.fc[tile];
  if (!a.force && !(f & kFlagStepping)) return;
  const Tile t = a.tc[tile];
  const int p = (f & kFlagParity) ? 1 : 0;
  const T* __restrict__ x = p ? a.x1 : a.x0;
  const T* __restrict__ xn = p ? a.x0 : a.x1;
  const T* __restrict__ aty = p ? a.aty1 : a.aty0;
  T* __restrict__ atyn = p ? a.aty0 : a.aty1;
  double inter = 0.0;
#pragma unroll
  for (int u = 0; u < kChunkCols; ++u) {
    const int j = t.begin + u * kBlock + static_cast<int>(threadIdx.x);
    if (j < t.end) {
      const T r = slice_sum(a.ypart, j, a.slice_cols, a.row_slices);
      atyn[j] = r;
      inter += static_cast<double>(xn[j] - x[j]) * static_cast<double>(r - aty[j]);
    }
  }
  inter = block_sum(inter, red);
  if (threadIdx.x == 0) a.ps_q[tile] = inter;
}

// Copies an instance record between global memory and a workgroup's LDS copy, in 8-byte
// words over the lanes of the workgroup.
__device__ __forceinline__ void copy_state(InstState* dst, const InstState* src) {
  static_assert(sizeof(InstState) % sizeof(uint64_t) == 0, "InstState is copied in 8-byte words");
  constexpr int W = static_cast<int>(sizeof(InstState) / sizeof(uint64_t));
  const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
  uint64_t* d = reinterpret_cast<uint64_t*>(dst);
  for (int k = threadIdx.x; k < W; k += kBlock) d[k] = s[k];
}

// The step decision of instance `inst` (pdhg.cpp:410-503), made on the workgroup's copy `sm`
// of its record in[inst]: the step's partials (ps_c of the primal chunks, a.ps_a of the tiles
// of A, a.ps_t of the tiles of A^T) are reduced in tile order and thread 0 applies decide_step.
// With `out`, the copy is then written to out[inst] (out may be in) and, if the instance
// stepped, its tiles' flags are published and a stop is counted in done_count. Without `out`
// nothing is written: every workgroup that decides from the same inputs gets the same bits, so
// the chunks of a deciding primal launch agree without exchanging anything.
// The partials are visible because they were written by earlier launches (or, in
// k_resident, by this workgroup before a barrier); no device-scope fence is needed.
// `red` holds 3 * (kBlock / kWave) values; sm is shared by the workgroup.
template <class T>
__device__ __forceinline__ void decide_instance(const Args<T>& a, const int inst,
                                                const InstDesc& d, const InstState* in,
                                                InstState* out, const double* ps_c,
                                                InstState* sm, double* red) {
  // the tiles' partials and the record are loaded together (the partials of an instance that
  // did not step are read and dropped): ||x'-x||^2, ||y'-y||^2 and the interaction as lane
  // sums over the tiles, then one reduction. The record's words are loaded first and stored to
  // LDS after the partials' loads are issued, so one memory round trip covers both.
  constexpr int W = static_cast<int>(sizeof(InstState) / sizeof(uint64_t));
  static_assert(sizeof(InstState) % sizeof(uint64_t) == 0 && W <= kBlock, "one record word per lane");
  uint64_t word = 0;
  if (static_cast<int>(threadIdx.x) < W) word = reinterpret_cast<const uint64_t*>(in + inst)[threadIdx.x];
  double v[3] = {0.0, 0.0, 0.0};
  lane_sums3<4>(ps_c + d.tc0, d.tc1 - d.tc0,
                a.slices ? a.ps_r + d.tr0 : a.ps_a + d.ta0, a.slices ? d.tr1 - d.tr0 : d.ta1 - d.ta0,
                a.row_slices ? a.ps_q + d.tc0 : a.ps_t + d.tt0,
                a.row_slices ? d.tc1 - d.tc0 : d.tt1 - d.tt0, v);
  if (static_cast<int>(threadIdx.x) < W) reinterpret_cast<uint64_t*>(sm)[threadIdx.x] = word;
  __syncthreads();
  const bool go = stepping(*sm);
  if (go) {
    // the rule's factors at this iteration come with the record; the workgroup that writes
    // the record loads those of the next iteration (after an accepted step) beside the
    // reduction
    double next_rule[2] = {0.0, 0.0};
    if (out && threadIdx.x == 0) rule_factors(a, sm->iteration + 1, next_rule);
    block_reduce_n<3, 0u>(v, red);
    if (threadIdx.x == 0) {
      const int64_t before = sm->iteration;
      decide_step(*sm, v[0], v[1], v[2], d.rows > 0, sm->rule_shrink, sm->rule_grow,
                  out ? a.done_count : nullptr);
      if (out && sm->iteration != before) {
        sm->rule_shrink = next_rule[0];
        sm->rule_grow = next_rule[1];
      }
    }
    __syncthreads();
  }
  if (out && (go || out != in)) {
    copy_state(out + inst, sm);
    if (go) publish_flags(a, d, step_flags(*sm));
  }
}

// One workgroup per instance: the decision of the step that just ran, from the records in
// a.state into a.state_next. A separate launch (rather than a last-arriving tile) makes the
// partials visible across XCDs through the kernel boundary, with no per-tile device-scope
// fence. The host runs it before every evaluation stage and at the end of a captured block;
// the other steps are decided by the next primal launch (k_step_primal with decide = 1).
template <class T>
__global__ __launch_bounds__(kBlock) void k_step_finalize(Args<T> a) {
  __shared__ double red[3 * (kBlock / kWave)];
  __shared__ InstState sm;
  const int inst = static_cast<int>(blockIdx.x);
  decide_instance(a, inst, a.desc[inst], a.state, a.state_next, a.ps_c, &sm, red);
}

// The primal step of column chunk `tile` after deciding the previous step itself: every chunk
// of an instance makes the decision from the same records (a.state) and partials (ps_c_prev,
// ps_a, ps_t), and the instance's first chunk writes the decided records to a.state_next and
// publishes the flags, which the dual and A^T launches of this step read. This saves the
// decision's own launch and its wait on most steps; the arithmetic, and so every answer, is
// bit for bit that of the decision launch. The operands that do not depend on the decision,
// and x and A^T y of the buffer an accepted step makes current (almost every step is
// accepted), are loaded before it; a rejected step reloads them from the other buffer.
template <class T>
__device__ __forceinline__ void primal_chunk_decided(const Args<T>& a, const int tile,
                                                     double* red, InstState* sm) {
  const Tile t = a.tc[tile];
  const InstDesc d = a.cdesc[tile];
  const int p_old = a.state[t.inst].parity;
  const T* __restrict__ xs = p_old ? a.x0 : a.x1;
  const T* __restrict__ atys = p_old ? a.aty0 : a.aty1;
  constexpr int U = kChunkCols;
  T xj[U], c[U], at[U], lo[U], hi[U], sum[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = t.begin + u * kBlock + threadIdx.x;
    if (j < t.end) {
      xj[u] = xs[j];
      c[u] = a.chat[j];
      at[u] = atys[j];
      lo[u] = a.lhat[j];
      hi[u] = a.uhat[j];
      sum[u] = a.xsum[j];
    }
  }
  decide_instance(a, t.inst, d, a.state, tile == d.tc0 ? a.state_next : nullptr, a.ps_c_prev,
                  sm, red);
  const InstState& s = *sm;
  if (!stepping(s)) return;
  const int p = s.parity;
  if (p == p_old) {
    const T* __restrict__ x = p ? a.x1 : a.x0;
    const T* __restrict__ aty = p ? a.aty1 : a.aty0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = t.begin + u * kBlock + threadIdx.x;
      if (j < t.end) {
        xj[u] = x[j];
        at[u] = aty[j];
      }
    }
  }
  primal_update(a, t, tile, p, s.pending_accept != 0, static_cast<T>(s.eta / s.omega), xj, c, at,
                lo, hi, sum, red);
}

// The primal step launch: decide = 0 steps with the published decision (primal_chunk),
// decide = 1 first makes the previous step's decision (primal_chunk_decided).
template <class T>
__global__ __launch_bounds__(kBlock) void k_step_primal(Args<T> a) {
  __shared__ double red[3 * (kBlock / kWave)];
  __shared__ InstState sm;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  if (a.decide) primal_chunk_decided(a, tile, red, &sm);
  else primal_chunk(a, tile, red);
}

// ---- evaluation kernels (FP64) ---------------------------------------------------------------

// k_eval_prepare: for instances with a pending evaluation, fold the last accepted iterate into
// the running sums and write the unscaled candidates (pdhg.cpp:357-366, 505-531), plus the
// squared distances to the restart point used by the primal-weight update (pdhg.cpp:613-624).
// eval_prepare_chunk does it for column chunk `tile` (is_col) or row chunk `tile`.
template <class T>
__device__ __forceinline__ void eval_prepare_chunk(const Args<T>& a, const bool is_col,
                                                   const int tile, double* red) {
  if (!((is_col ? a.fc[tile] : a.fr[tile]) & kFlagEval)) return;
  const Tile t = is_col ? a.tc[tile] : a.tr[tile];
  const InstState& s = a.state[t.inst];
  const int p = s.parity;
  const bool acc = s.pending_accept;
  const double count = static_cast<double>(s.averaged);
  const bool have_avg = s.averaged > 0;
  const T* __restrict__ v = is_col ? (p ? a.x1 : a.x0) : (p ? a.y1 : a.y0);
  const T* __restrict__ rst = is_col ? a.xrst : a.yrst;
  T* __restrict__ sum = is_col ? a.xsum : a.ysum;
  double* __restrict__ tot = is_col ? a.xtot : a.ytot;
  const double* __restrict__ scale = is_col ? a.dc : a.dr;
  double* __restrict__ u0 = is_col ? a.xu0 : a.yu0;
  double* __restrict__ u1 = is_col ? a.xu1 : a.yu1;
  constexpr bool kSplitSum = !std::is_same<T, double>::value;
  double d2 = 0.0;
  for (int k = t.begin + threadIdx.x; k < t.end; k += kBlock) {
    const T vk = v[k];
    double total;
    if constexpr (kSplitSum) {
      total = tot[k] + static_cast<double>(sum[k]);
      if (acc) total += static_cast<double>(vk);
      tot[k] = total;
      sum[k] = static_cast<T>(0);
    } else {
      total = sum[k];
      if (acc) {
        total += vk;
        sum[k] = total;
      }
    }
    u0[k] = static_cast<double>(vk) * scale[k];
    u1[k] = have_avg ? (total / count) * scale[k] : 0.0;
    const double d = static_cast<double>(vk) - static_cast<double>(rst[k]);
    d2 += d * d;
  }
  d2 = block_sum(d2, red);
  if (threadIdx.x == 0) (is_col ? a.pe_c : a.pe_r)[tile * 2] = d2;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_eval_prepare(Args<T> a) {
  __shared__ double red[kBlock / kWave];
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const bool is_col = b < a.n_tc;
  eval_prepare_chunk(a, is_col, is_col ? b : b - a.n_tc, red);
}

// eval_rows_tile: activity = A x for both candidates on the original A, and every row term of
// evaluate() (pdhg.cpp:154-170, 203-222, 227-238). Partials per tile, per candidate c:
// [4c+0] sum violation^2, [4c+1] support, [4c+2] dual violation from y, [4c+3] complementarity.
template <class T, bool PACK>
__device__ __forceinline__ void eval_rows_tile(const Args<T>& a, const int tile, D2* stage,
                                               double* red) {
  if (!(a.fa[tile] & (a.report ? kFlagReport : kFlagEval))) return;
  const Tile t = a.ta[tile];
  const double* __restrict__ xa = a.report ? a.bestx : a.xu0;
  const double* __restrict__ xb = a.report ? a.bestx : a.xu1;
  const double* __restrict__ ya = a.report ? a.besty : a.yu0;
  const double* __restrict__ yb = a.report ? a.besty : a.yu1;
  const double* __restrict__ val = a.a_orig_st;
  const int32_t* __restrict__ sidx = a.a_sidx;
  // the step kernels' staging order, with the FP64 values stored in it too
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const auto prod = [&](int q) {
    const int c = gather_index<PACK>(t, sidx, q);
    const double v = val[q];
    return D2{v * xa[c], v * xb[c]};
  };
  auto f = [xa, xb](int c, int, double v) { return D2{v * xa[c], v * xb[c]}; };
  constexpr int R = kMaxTileRows / kBlock;
  double lo_[R], hi_[R], ya_[R], yb_[R];
  rowblock_product<kBlock, D2>(
      t, a.a_ptr, prod,
      GroupStager<kBlock, D2, PACK, double, decltype(f)>{sidx, a.a_pos, val, f},
      [&](int i, int k) {
        lo_[k] = a.rl[i];
        hi_[k] = a.ru[i];
        ya_[k] = ya[i];
        yb_[k] = yb[i];
      },
      [&](int i, int k, D2 act) {
        const double lo = lo_[k], hi = hi_[k];
        const bool lo_f = dev_finite_bound(lo), hi_f = dev_finite_bound(hi);
        if (a.report) a.activity[i] = act.a;
        for (int c = 0; c < 2; ++c) {
          const double ai = c == 0 ? act.a : act.b;
          const double yi = c == 0 ? ya_[k] : yb_[k];
          double viol = 0.0;
          if (lo_f) viol = dev_max(viol, lo - ai);
          if (hi_f) viol = dev_max(viol, ai - hi);
          acc[4 * c + 0] += viol * viol;
          if (yi > 0.0) {
            if (hi_f) acc[4 * c + 1] += yi * hi;
            else acc[4 * c + 2] += yi * yi;
          } else if (yi < 0.0) {
            if (lo_f) acc[4 * c + 1] += yi * lo;
            else acc[4 * c + 2] += yi * yi;
          }
          if (lo != hi) {
            const double ls = lo_f ? ai - lo : INFINITY;
            const double us = hi_f ? hi - ai : INFINITY;
            acc[4 * c + 3] = dev_max(acc[4 * c + 3], fabs(yi) * dev_min(ls, us));
          }
        }
      },
      stage);
  block_reduce_n<8, (1u << 3) | (1u << 7)>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a.pe_a[(size_t)tile * 8 + k] = acc[k];
  }
}

__device__ __forceinline__ double worst_of(const Residuals& r) {
  return dev_max(dev_max(r.primal, r.dual), r.gap);
}

__device__ __forceinline__ bool meets_request(const Residuals& r, double tol, int absolute_clause) {
  return r.primal <= tol && r.dual <= tol && r.gap <= tol &&
         (!absolute_clause || r.absolute_primal <= 1e-7);
}

__device__ __forceinline__ bool meets_project_standard(const Residuals& r) {
  return r.absolute_primal <= 1e-7 && r.absolute_dual <= 1e-7 && r.gap_as_verified <= 1e-9 &&
         r.complementarity <= 1e-6;
}

// pdhg.cpp:169-170, 223-224, 252-261
__device__ Residuals make_residuals(const InstState& s, double pv2, double dviol, double support,
                                    double bound, double compl_max, double pobj) {
  Residuals r;
  r.absolute_primal = sqrt(pv2);
  r.primal = r.absolute_primal / (1.0 + s.bound_norm);
  r.absolute_dual = sqrt(dviol);
  r.dual = r.absolute_dual / (1.0 + s.cost_norm);
  r.complementarity = compl_max;
  r.primal_objective = pobj;
  r.dual_objective = bound - support;
  const double absolute_gap = fabs(r.primal_objective - r.dual_objective);
  r.gap = absolute_gap / (1.0 + fabs(r.primal_objective) + fabs(r.dual_objective));
  r.gap_as_verified = absolute_gap / dev_max(1.0, fabs(r.primal_objective));
  return r;
}

// Candidate choice, best tracking, termination and restart (pdhg.cpp:512-643).
__device__ void decide_evaluation(InstState& s, const Residuals& cur, const Residuals& avg,
                                  double dx2, double dy2, uint32_t* done_count) {
  s.evaluations += 1;
  s.current = cur;
  s.average = avg;
  const bool have_avg = s.averaged > 0;
  const bool pick_avg = have_avg && worst_of(avg) < worst_of(cur);
  const Residuals chosen = pick_avg ? avg : cur;
  s.chosen_avg = pick_avg;
  s.copy_best = 0;
  s.do_restart = 0;
  if (worst_of(chosen) < s.best_worst) {
    s.best = chosen;
    s.best_worst = worst_of(chosen);
    s.copy_best = 1;
  }
  s.eval_pending = 0;
  s.pending_accept = 0;  // folded into the sums by k_eval_prepare
  if (meets_request(chosen, s.tolerance, s.absolute_clause) &&
      (s.stop_at_request || meets_project_standard(chosen))) {
    s.best = chosen;
    s.best_worst = worst_of(chosen);
    s.copy_best = 1;
    s.converged = 1;
    finish(s, meets_project_standard(chosen) ? kStOptimal : kStFeasible, done_count);
    return;
  }
  if (s.use_restarts) {
    const double kkt = worst_of(chosen);
    const int64_t since = s.iteration - s.last_restart;
    const bool sufficient = kkt <= 0.2 * s.restart_kkt;
    const int64_t floor36 = static_cast<int64_t>(0.36 * static_cast<double>(s.iteration));
    const bool artificial = since >= (floor36 > 40 ? floor36 : 40);
    if (sufficient || artificial) {
      const double dxn = sqrt(dx2), dyn = sqrt(dy2);
      s.dx_norm = dxn;
      s.dy_norm = dyn;
      if (dxn > 1e-12 && dyn > 1e-12) {
        double w = exp(0.5 * log(dyn / dxn) + 0.5 * log(s.omega));
        s.omega = w < 1e-6 ? 1e-6 : (1e6 < w ? 1e6 : w);
      }
      s.do_restart = 1;
      s.averaged = 0;
      s.restart_kkt = kkt;
      s.last_restart = s.iteration;
      s.restarts += 1;
    }
  }
  if (s.iteration >= s.iteration_limit) finish(s, kStIterationLimit, done_count);
}

// eval_cols_tile: reduced cost d = c + A^T y per candidate on the original A and every column
// term of evaluate() (pdhg.cpp:175-199, 239-257). Partials per tile, per candidate c:
// [4c+0] dual violation from d, [4c+1] bound contribution, [4c+2] complementarity,
// [4c+3] primal objective.
template <class T, bool PACK>
__device__ __forceinline__ void eval_cols_tile(const Args<T>& a, const int tile, D2* stage,
                                               double* red) {
  if (!(a.ft[tile] & (a.report ? kFlagReport : kFlagEval))) return;
  const Tile t = a.tt[tile];
  const double* __restrict__ xa = a.report ? a.bestx : a.xu0;
  const double* __restrict__ xb = a.report ? a.bestx : a.xu1;
  const double* __restrict__ ya = a.report ? a.besty : a.yu0;
  const double* __restrict__ yb = a.report ? a.besty : a.yu1;
  const double* __restrict__ val = a.t_orig_st;
  const int32_t* __restrict__ sidx = a.t_sidx;
  // the step kernels' staging order, with the FP64 values stored in it too
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const auto prod = [&](int q) {
    const int r = gather_index<PACK>(t, sidx, q);
    const double v = val[q];
    return D2{v * ya[r], v * yb[r]};
  };
  auto f = [ya, yb](int r, int, double v) { return D2{v * ya[r], v * yb[r]}; };
  constexpr int R = kMaxTileRows / kBlock;
  double lo_[R], hi_[R], c_[R], xa_[R], xb_[R];
  rowblock_product<kBlock, D2>(
      t, a.t_ptr, prod,
      GroupStager<kBlock, D2, PACK, double, decltype(f)>{sidx, a.t_pos, val, f},
      [&](int j, int k) {
        lo_[k] = a.cl[j];
        hi_[k] = a.cu[j];
        c_[k] = a.cmin[j];
        xa_[k] = xa[j];
        xb_[k] = xb[j];
      },
      [&](int j, int k, D2 dot) {
        const double lo = lo_[k], hi = hi_[k];
        const bool lo_f = dev_finite_bound(lo), hi_f = dev_finite_bound(hi);
        const double cj = c_[k];
        for (int c = 0; c < 2; ++c) {
          const double d = cj + (c == 0 ? dot.a : dot.b);
          const double xj = c == 0 ? xa_[k] : xb_[k];
          if (c == 0 && a.report) a.reduced[j] = d;
          if (d > 0.0) {
            if (lo_f) acc[4 * c + 1] += d * lo;
            else acc[4 * c + 0] += d * d;
          } else if (d < 0.0) {
            if (hi_f) acc[4 * c + 1] += d * hi;
            else acc[4 * c + 0] += d * d;
          }
          if (lo != hi) {
            const double ls = lo_f ? xj - lo : INFINITY;
            const double us = hi_f ? hi - xj : INFINITY;
            acc[4 * c + 2] = dev_max(acc[4 * c + 2], fabs(d) * dev_min(ls, us));
          }
          acc[4 * c + 3] += cj * xj;
        }
      },
      stage);
  block_reduce_n<8, (1u << 2) | (1u << 6)>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a.pe_t[(size_t)tile * 8 + k] = acc[k];
  }
}

// The row and the column products of an evaluation in one launch: workgroups [0, rows) take
// the tiles of A, the others those of A^T. Both read only the candidates k_eval_prepare wrote
// and write their own partials, so neither waits for the other.
template <class T, int CAP, bool PACK>
__global__ __launch_bounds__(kBlock) void k_eval_products(Args<T> a, int rows) {
  __shared__ D2 stage[CAP];
  __shared__ double red[8 * (kBlock / kWave)];
  const int b = static_cast<int>(blockIdx.x);
  if (b < rows) eval_rows_tile<T, PACK>(a, xcd_tile(b, rows), stage, red);
  else eval_cols_tile<T, PACK>(a, xcd_tile(b - rows, static_cast<int>(gridDim.x) - rows), stage, red);
}

// For an evaluated instance (one workgroup): reduce every evaluation partial in tile order,
// form both Residuals and decide (or store the report). `red` holds 18 * (kBlock / kWave)
// values; `flags` is shared by the workgroup.
template <class T>
__device__ __forceinline__ void eval_finalize_instance(const Args<T>& a, const int inst,
                                                       double* red, uint32_t* flags) {
  InstState& s = a.state[inst];
  if (!(a.report ? s.report_pending : s.eval_pending)) return;
  const InstDesc d = a.desc[inst];
  // per candidate c, v[8c + ...]: 0 viol^2, 1 support, 2 dual violation (rows), 3 compl (rows),
  // 4 dual violation (cols), 5 bound, 6 compl (cols), 7 primal objective; 16, 17 the squared
  // distances to the restart point. Lane sums (maxima) over the tiles, then one reduction.
  double v[18];
#pragma unroll
  for (int i = 0; i < 18; ++i) v[i] = 0.0;
  // each lane takes its tiles in order, kEvalBatch tiles' records loaded before they are added
  // (as lane_sum; a position past the table reloads its last record and is not added)
  constexpr int kEvalBatch = 4;
  const auto fold = [&](const double* recs, const int count, auto add) {
    for (int k0 = static_cast<int>(threadIdx.x); k0 < count; k0 += kEvalBatch * kBlock) {
      double2 w[kEvalBatch][4];
#pragma unroll
      for (int u = 0; u < kEvalBatch; ++u) {
        const int k = k0 + u * kBlock;
        const double2* q = reinterpret_cast<const double2*>(recs + (size_t)(k < count ? k : count - 1) * 8);
#pragma unroll
        for (int h = 0; h < 4; ++h) w[u][h] = q[h];
      }
#pragma unroll
      for (int u = 0; u < kEvalBatch; ++u) {
        if (k0 + u * kBlock < count) {
          const double q[8] = {w[u][0].x, w[u][0].y, w[u][1].x, w[u][1].y,
                               w[u][2].x, w[u][2].y, w[u][3].x, w[u][3].y};
          add(q);
        }
      }
    }
  };
  fold(a.pe_a + (size_t)d.ta0 * 8, d.ta1 - d.ta0, [&](const double* q) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      v[8 * c + 0] += q[4 * c + 0];
      v[8 * c + 1] += q[4 * c + 1];
      v[8 * c + 2] += q[4 * c + 2];
      v[8 * c + 3] = dev_max(v[8 * c + 3], q[4 * c + 3]);
    }
  });
  fold(a.pe_t + (size_t)d.tt0 * 8, d.tt1 - d.tt0, [&](const double* q) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      v[8 * c + 4] += q[4 * c + 0];
      v[8 * c + 5] += q[4 * c + 1];
      v[8 * c + 6] = dev_max(v[8 * c + 6], q[4 * c + 2]);
      v[8 * c + 7] += q[4 * c + 3];
    }
  });
  if (!a.report) {
    v[16] = lane_sum(a.pe_c + (size_t)d.tc0 * 2, d.tc1 - d.tc0, 2, v[16]);
    v[17] = lane_sum(a.pe_r + (size_t)d.tr0 * 2, d.tr1 - d.tr0, 2, v[17]);
  }
  block_reduce_n<18, (1u << 3) | (1u << 6) | (1u << 11) | (1u << 14)>(v, red);
  if (threadIdx.x == 0) {
    Residuals r[2];
    for (int c = 0; c < 2; ++c) {
      const double* q = v + 8 * c;
      r[c] = make_residuals(s, q[0], q[4] + q[2], q[1], q[5], dev_max(q[3], q[6]), q[7]);
    }
    if (a.report) {
      s.report = r[0];
      s.report_pending = 0;
      // nothing may rewrite a reported instance's answer: later commit stages skip it, so the
      // host can read it back while the rest of the batch keeps running
      s.copy_best = 0;
      s.do_restart = 0;
    } else {
      decide_evaluation(s, r[0], r[1], v[16], v[17], a.done_count);
    }
    *flags = step_flags(s);
  }
  __syncthreads();
  publish_flags(a, d, *flags);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_eval_finalize(Args<T> a) {
  __shared__ double red[18 * (kBlock / kWave)];
  __shared__ uint32_t flags;
  eval_finalize_instance(a, static_cast<int>(blockIdx.x), red, &flags);
}

// eval_commit_chunk: apply the decisions of the evaluation finalize (pdhg.cpp:541-545, 592-594,
// 632-638) to column chunk `tile` (is_col) or row chunk `tile`. Idempotent, so an instance
// that stopped may see it again without effect.
template <class T>
__device__ __forceinline__ void eval_commit_chunk(const Args<T>& a, const bool is_col,
                                                  const int tile) {
  if (!((is_col ? a.fc[tile] : a.fr[tile]) & kFlagCommit)) return;
  const Tile t = is_col ? a.tc[tile] : a.tr[tile];
  const InstState& s = a.state[t.inst];
  const int p = s.parity;
  const bool copy = s.copy_best, restart = s.do_restart, from_avg = s.chosen_avg;
  const T* __restrict__ v = is_col ? (p ? a.x1 : a.x0) : (p ? a.y1 : a.y0);
  T* __restrict__ rst = is_col ? a.xrst : a.yrst;
  T* __restrict__ sum = is_col ? a.xsum : a.ysum;
  double* __restrict__ tot = is_col ? a.xtot : a.ytot;
  const double* __restrict__ src = is_col ? (from_avg ? a.xu1 : a.xu0) : (from_avg ? a.yu1 : a.yu0);
  double* __restrict__ best = is_col ? a.bestx : a.besty;
  constexpr bool kSplitSum = !std::is_same<T, double>::value;
  for (int k = t.begin + threadIdx.x; k < t.end; k += kBlock) {
    if (copy) best[k] = src[k];
    if (restart) {
      sum[k] = static_cast<T>(0);
      if constexpr (kSplitSum) tot[k] = 0.0;
      rst[k] = v[k];
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_eval_commit(Args<T> a) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const bool is_col = b < a.n_tc;
  eval_commit_chunk(a, is_col, is_col ? b : b - a.n_tc);
}

// ---- resident solve of small instances ------------------------------------------------------

// One workgroup per instance runs up to `steps` rounds of its solve in one launch, each round a
// PDHG step or, when one is pending, an evaluation stage; an instance that stops runs its final
// (report) evaluation before the launch ends. Every phase is the per-tile code of the separate
// launches above (primal chunks, tiles of A, tiles of A^T, the decision; evaluation: chunk
// preparation, row and column tiles, finalize, commit), tile after tile in index order with a
// barrier between phases, so the partials and every answer are bit for bit those of the
// separate launches. It is used when every instance of a batch has only a few tiles: such a
// step is a short chain of dependent memory round trips, and running all of them in one
// launch keeps the instance's data in the caches instead of crossing a kernel boundary (and
// its launch gap) four times per step. The step product stages its values in the first
// CAP * sizeof(T) bytes of the evaluation's staging buffer.
// The phases of one round of the resident solve for instance `inst` (whole workgroup).
template <class T, int CAP, bool PACK>
__device__ __forceinline__ void resident_eval(const Args<T>& a, const int inst, const InstDesc& d,
                                              D2* stage, double* red, uint32_t* flags) {
  for (int k = d.tc0; k < d.tc1; ++k) eval_prepare_chunk(a, true, k, red);
  for (int k = d.tr0; k < d.tr1; ++k) eval_prepare_chunk(a, false, k, red);
  __syncthreads();
  for (int k = d.ta0; k < d.ta1; ++k) eval_rows_tile<T, PACK>(a, k, stage, red);
  for (int k = d.tt0; k < d.tt1; ++k) eval_cols_tile<T, PACK>(a, k, stage, red);
  __syncthreads();
  eval_finalize_instance(a, inst, red, flags);
  __syncthreads();
  for (int k = d.tc0; k < d.tc1; ++k) eval_commit_chunk(a, true, k);
  for (int k = d.tr0; k < d.tr1; ++k) eval_commit_chunk(a, false, k);
}

template <class T, int CAP, bool PACK>
__device__ __forceinline__ void resident_step(const Args<T>& a, const int inst, const InstDesc& d,
                                              T* stage, double* red, InstState* sm) {
  // a tile's groups of four staged nonzeros (CAP / 4, one more when the tile does not start
  // on a multiple of four) all in flight at once
  constexpr int kInFlight = (CAP / 4 + 1 + kBlock - 1) / kBlock;
  for (int k = d.tc0; k < d.tc1; ++k) primal_chunk(a, k, red);
  __syncthreads();
  for (int k = d.ta0; k < d.ta1; ++k) dual_tile<T, PACK, kInFlight>(a, k, stage, red);
  __syncthreads();
  for (int k = d.tt0; k < d.tt1; ++k) aty_tile<T, PACK, kInFlight>(a, k, stage, red);
  __syncthreads();
  decide_instance(a, inst, d, a.state, a.state, a.ps_c, sm, red);
}

// The final (report) evaluation of a stopped instance.
template <class T, int CAP, bool PACK>
__device__ __forceinline__ void resident_report(const Args<T>& a, const int inst,
                                                const InstDesc& d, D2* stage, double* red,
                                                uint32_t* flags) {
  Args<T> r = a;
  r.report = 1;
  for (int k = d.ta0; k < d.ta1; ++k) eval_rows_tile<T, PACK>(r, k, stage, red);
  for (int k = d.tt0; k < d.tt1; ++k) eval_cols_tile<T, PACK>(r, k, stage, red);
  __syncthreads();
  eval_finalize_instance(r, inst, red, flags);
}

template <class T, int CAP, bool PACK>
__global__ __launch_bounds__(kBlock) void k_resident(Args<T> a, int32_t steps) {
  __shared__ D2 stage[CAP];
  __shared__ double red[18 * (kBlock / kWave)];
  __shared__ InstState sm;
  __shared__ uint32_t flags;
  const int inst = static_cast<int>(blockIdx.x);
  const InstDesc d = a.desc[inst];
  const InstState* rec = a.state + inst;
  for (int round = 0; round < steps; ++round) {
    // the record was last written by this workgroup's previous phase
    __syncthreads();
    const int32_t done = rec->done, evaluate = rec->eval_pending;
    if (done) break;
    if (evaluate) resident_eval<T, CAP, PACK>(a, inst, d, stage, red, &flags);
    else resident_step<T, CAP, PACK>(a, inst, d, reinterpret_cast<T*>(stage), red, &sm);
  }
  __syncthreads();
  if (rec->report_pending) resident_report<T, CAP, PACK>(a, inst, d, stage, red, &flags);
}

// ---- set-up kernels (FP64, no FMA contraction: bit-for-bit the reference's arithmetic) -------

// Ruiz (max) or Pock-Chambolle (sum) statistic of every row of a CSR matrix, as the pass
// multiplier 1/sqrt(stat) (1 for an empty row). Sums run in stored order, which is the
// reference's accumulation order (scaling.cpp:44-100): one lane per row, over the tile's
// |values| staged in LDS with coalesced loads (a row longer than the staging capacity is read
// from memory by its lane).
__global__ __launch_bounds__(kBlock) void k_row_pass(const Tile* __restrict__ tiles,
                                                      const int32_t* __restrict__ ptr,
                                                      const double* __restrict__ val,
                                                      int32_t use_sum, double* __restrict__ pass) {
  __shared__ double mag[kMaxTileNnz];
  const Tile t = tiles[blockIdx.x];
  const int nrows = t.end - t.begin, nnz = t.p1 - t.p0;
  const bool staged = t.kind != kTileVector;
  if (staged) {
    for (int k = threadIdx.x; k < nnz; k += kBlock) mag[k] = fabs(val[t.p0 + k]);
    __syncthreads();
  }
  for (int row = threadIdx.x; row < nrows; row += kBlock) {
    const int r = t.begin + row;
    const int b = ptr[r] - t.p0, e = ptr[r + 1] - t.p0;
    double acc = 0.0;
    for (int k = b; k < e; ++k) {
      const double m = staged ? mag[k] : fabs(val[t.p0 + k]);
      acc = use_sum ? acc + m : dev_max(acc, m);
    }
    pass[r] = acc > 0.0 ? 1.0 / sqrt(acc) : 1.0;
  }
}

// value * row_pass[i] * column_pass[j] in both layouts (scaling.cpp:14-40), one workgroup per
// row-block tile, nonzero-parallel: a tile of several rows first writes each nonzero's row
// (within the tile) into LDS, then consecutive lanes scale consecutive nonzeros, so every load
// and store is coalesced. `row_major`: rows of this CSR are rows of A (the index array holds
// columns); otherwise its rows are columns of A.
__global__ __launch_bounds__(kBlock) void k_apply_pass(const Tile* __restrict__ tiles,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx,
                                                        double* __restrict__ val, int32_t row_major,
                                                        const double* __restrict__ row_pass,
                                                        const double* __restrict__ col_pass) {
  __shared__ uint16_t row_of[kMaxTileNnz];
  const Tile t = tiles[blockIdx.x];
  const int nrows = t.end - t.begin, nnz = t.p1 - t.p0;
  if (nrows <= 0 || nnz <= 0) return;
  const bool one_row = nrows == 1;  // also every vector tile (more than kMaxTileNnz nonzeros)
  if (!one_row) {
    // at most kMaxTileRows rows: lane l expands rows l, l + kBlock
    for (int row = threadIdx.x; row < nrows; row += kBlock) {
      const int r = t.begin + row;
      for (int p = ptr[r]; p < ptr[r + 1]; ++p) row_of[p - t.p0] = static_cast<uint16_t>(row);
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < nnz; k += kBlock) {
    const int p = t.p0 + k;
    const int r = t.begin + (one_row ? 0 : static_cast<int>(row_of[k]));
    const double ri = row_major ? row_pass[r] : row_pass[idx[p]];
    const double cj = row_major ? col_pass[idx[p]] : col_pass[r];
    val[p] = val[p] * ri * cj;
  }
}

__global__ __launch_bounds__(kBlock) void k_scale_mul(double* __restrict__ acc,
                                                       const double* __restrict__ pass, int32_t n) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < n) acc[k] *= pass[k];
}

// Scaled vectors (scaling.cpp:127-145) in the iterate precision, absent bounds as +-inf.
template <class T>
__global__ __launch_bounds__(kBlock) void k_scaled_vectors(Args<T> a, T* chat, T* lhat, T* uhat,
                                                            T* rlhat, T* ruhat, int32_t ncols,
                                                            int32_t nrows) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < ncols) {
    const double dc = a.dc[k];
    chat[k] = static_cast<T>(a.cmin[k] * dc);
    const double lo = a.cl[k], hi = a.cu[k];
    lhat[k] = dev_finite_bound(lo) ? static_cast<T>(lo * (1.0 / dc)) : static_cast<T>(-INFINITY);
    uhat[k] = dev_finite_bound(hi) ? static_cast<T>(hi * (1.0 / dc)) : static_cast<T>(INFINITY);
  }
  if (k < nrows) {
    const double dr = a.dr[k];
    const double lo = a.rl[k], hi = a.ru[k];
    rlhat[k] = dev_finite_bound(lo) ? static_cast<T>(lo * dr) : static_cast<T>(-INFINITY);
    ruhat[k] = dev_finite_bound(hi) ? static_cast<T>(hi * dr) : static_cast<T>(INFINITY);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_convert(const double* __restrict__ src, T* __restrict__ dst,
                                                     int64_t n) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) dst[k] = static_cast<T>(src[k]);
}

// Rank, in gather-index order, of the entry stored at position j (from the tile's first
// nonzero) of a stream tile of nnz entries that starts o = p0 % 4 positions past a multiple of
// four. GroupStager has lane l of a wavefront load the four stored entries 4l .. 4l+3 of a
// 256-entry block (blocks aligned to p0 - o) and gather entry 4l+e with its e-th gather
// instruction; storing the rank order down the lanes (entry 4l+e of a block holds the block's
// (64e + l)-th entry, counting only the tile's own positions) makes each gather instruction's
// 64 lanes read 64 consecutive sorted indices. A bijection of [0, nnz) for every o and nnz.
__device__ __forceinline__ int group_rank(int j, int o, int nnz) {
  const int jj = j + o, blk = jj >> 8, l = (jj & 255) >> 2, e = jj & 3;
  const int lo = blk == 0 ? o : 0;                 // the block's first own position
  const int hi = min(256, o + nnz - 256 * blk);    // one past its last
  // lanes l' >= 0 with 4 l' + e' < x
  auto below = [](int x, int ep) { return max(0, (x - ep + 3) >> 2); };
  int r = blk == 0 ? 0 : 256 * blk - o;
  for (int ep = 0; ep < e; ++ep) r += below(hi, ep) - below(lo, ep);
  return r + l - below(lo, e);
}

// Staging order of every row-block tile (structure only; run once per upload). A stream tile's
// nonzeros are sorted by gather index (column of A for CSR(A), row of A for CSR(A^T)), ties
// broken by CSR position, with a bitonic sort of (index << 16 | slot) keys in LDS, and stored
// in GroupStager's lane order (group_rank); sidx gets the indices and pos the CSR-order slot of
// each. The keys are distinct, so the order is unique. With `sorted` = 0, or for a vector tile,
// the staging order is the CSR order.
__global__ __launch_bounds__(kBlock) void k_sort_tiles(const Tile* __restrict__ tiles,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx,
                                                        int32_t sorted, int32_t pack,
                                                        int32_t* __restrict__ sidx,
                                                        uint16_t* __restrict__ pos) {
  __shared__ uint64_t keys[kMaxTileNnz];
  const Tile t = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int p0 = ptr[t.begin], nnz = ptr[t.end] - p0;
  const bool vector = t.kind == kTileVector;
  // one staging entry: gather index g, CSR slot k (a vector tile never uses its slots)
  auto put = [&](int q, int g, int k) {
    if (pack) {
      sidx[q] = static_cast<int32_t>((static_cast<uint32_t>(g - t.base) << kPackShift) |
                                     static_cast<uint32_t>(vector ? 0 : k));
    } else {
      sidx[q] = g;
      pos[q] = static_cast<uint16_t>(k);
    }
  };
  if (!sorted || vector) {
    for (int k = tid; k < nnz; k += kBlock) put(p0 + k, idx[p0 + k], k);
    return;
  }
  // a stream tile holds at most cap <= kMaxTileNnz nonzeros, so slots fit 16 bits
  int n2 = 1;
  while (n2 < nnz) n2 <<= 1;
  for (int k = tid; k < n2; k += kBlock)
    keys[k] = k < nnz ? (static_cast<uint64_t>(static_cast<uint32_t>(idx[p0 + k])) << 16) |
                            static_cast<uint64_t>(k)
                      : ~0ull;
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < n2 / 2; i += kBlock) {
        const int lo = 2 * stride * (i / stride) + (i % stride), hi = lo + stride;
        const bool ascending = (lo & size) == 0;
        const uint64_t u = keys[lo], v = keys[hi];
        if ((u > v) == ascending) {
          keys[lo] = v;
          keys[hi] = u;
        }
      }
      __syncthreads();
    }
  }
  const int o = p0 & 3;
  for (int j = tid; j < nnz; j += kBlock) {
    const uint64_t key = keys[group_rank(j, o, nnz)];
    put(p0 + j, static_cast<int32_t>(key >> 16), static_cast<int>(key & 0xFFFFu));
  }
}

// The scaled values in each tile's staging order, in the iterate precision (a vector tile keeps
// the CSR order; every other tile reads through its slot map).
template <class T>
__global__ __launch_bounds__(kBlock) void k_stage_values(const Tile* __restrict__ tiles,
                                                          const int32_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ sidx,
                                                          const uint16_t* __restrict__ pos,
                                                          int32_t pack, const double* __restrict__ src,
                                                          T* __restrict__ dst) {
  const Tile t = tiles[blockIdx.x];
  const int p0 = ptr[t.begin], nnz = ptr[t.end] - p0;
  const bool vector = t.kind == kTileVector;
  for (int k = threadIdx.x; k < nnz; k += kBlock) {
    const int slot = vector ? k
                     : pack ? static_cast<int>(static_cast<uint32_t>(sidx[p0 + k]) & 0xFFFu)
                            : static_cast<int>(pos[p0 + k]);
    dst[p0 + k] = static_cast<T>(src[p0 + slot]);
  }
}

// The same for the stacked column slices: S's CSR position q holds A's CSR entry map[q].
template <class T>
__global__ __launch_bounds__(kBlock) void k_stage_slice_values(const Tile* __restrict__ tiles,
                                                                const int32_t* __restrict__ ptr,
                                                                const int32_t* __restrict__ sidx,
                                                                const uint16_t* __restrict__ pos,
                                                                int32_t pack,
                                                                const int32_t* __restrict__ map,
                                                                const double* __restrict__ src,
                                                                T* __restrict__ dst) {
  const Tile t = tiles[blockIdx.x];
  const int p0 = ptr[t.begin], nnz = ptr[t.end] - p0;
  const bool vector = t.kind == kTileVector;
  for (int k = threadIdx.x; k < nnz; k += kBlock) {
    const int slot = vector ? k
                     : pack ? static_cast<int>(static_cast<uint32_t>(sidx[p0 + k]) & 0xFFFu)
                            : static_cast<int>(pos[p0 + k]);
    dst[p0 + k] = static_cast<T>(src[map[p0 + slot]]);
  }
}

// ---- power iteration (scaling.cpp:161-194) ---------------------------------------------------

// y = A x over the row-block tiles of a CSR matrix with the step kernels' staged schedule:
// the scaled values in the staging order and the iterate precision V (gathers coalesced,
// rows reduced in CSR order), products and sums in FP64. With V = double the products are
// those of the FP64 scaled matrix; with V = float, of its FP32 copy that the steps use. The
// power iteration only sets the starting step size and its ceiling.
template <class V, bool PACK>
__global__ __launch_bounds__(kBlock) void k_spmv_tiles(const Tile* __restrict__ tiles,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ sidx,
                                                        const uint16_t* __restrict__ pos,
                                                        const V* __restrict__ val,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y) {
  __shared__ double stage[kMaxTileNnz];
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  rowblock_product<kBlock, double>(
      t, ptr,
      [&](int q) { return static_cast<double>(val[q]) * x[gather_index<PACK>(t, sidx, q)]; },
      gather_stager<kBlock, double, PACK>(sidx, pos, val, x), [](int, int) {},
      [&](int r, int, double acc) { y[r] = acc; }, stage);
}

// Per column chunk: sum of v^2 (slot 0 of pe_c).
__global__ __launch_bounds__(kBlock) void k_chunk_sumsq(const Tile* __restrict__ tc,
                                                         const double* __restrict__ v,
                                                         double* __restrict__ part) {
  __shared__ double red[kBlock / kWave];
  const Tile t = tc[blockIdx.x];
  double acc = 0.0;
  for (int k = t.begin + threadIdx.x; k < t.end; k += kBlock) acc += v[k] * v[k];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// One workgroup per instance: length = sqrt(sum of its chunk partials); a zero length is
// remembered (the reference then returns 1, scaling.cpp:176-179).
__global__ __launch_bounds__(kBlock) void k_inst_norm(const InstDesc* __restrict__ desc,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ out,
                                                       int32_t* __restrict__ zero_seen) {
  __shared__ double red[kBlock / kWave];
  const InstDesc d = desc[blockIdx.x];
  const double s = tile_sum(part, d.tc0, d.tc1 - d.tc0, 1, 0, red);
  if (threadIdx.x == 0) {
    out[blockIdx.x] = sqrt(s);
    if (zero_seen && s == 0.0) zero_seen[blockIdx.x] = 1;
  }
}

// Spectral norm estimate and the initial step sizes (scaling.cpp:187-193, pdhg.cpp:368,
// 495). `len` holds ||A^T A v|| after the last step, so the estimate is sqrt(len) * 1.01.
__global__ __launch_bounds__(kBlock) void k_power_finish(const double* __restrict__ len,
                                                          const int32_t* __restrict__ zero_seen,
                                                          InstState* __restrict__ state,
                                                          int32_t count) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= count) return;
  InstState& s = state[k];
  double norm = 1.0;
  if (!s.norm_trivial && !zero_seen[k]) {
    const double r = sqrt(len[k]);
    norm = r > 0.0 ? r * 1.01 : 1.0;
  }
  s.spectral_norm = norm;
  s.eta = norm > 0.0 ? 1.0 / norm : 1.0;
  s.eta_ceiling = 1.0e3 / dev_max(norm, 1e-12);
  // a warm start keeps the step size of the run it continues, within the rule's clamp
  if (s.eta_start > 0.0) s.eta = dev_min(dev_max(s.eta_start, 1e-12), s.eta_ceiling);
}

__global__ __launch_bounds__(kBlock) void k_fill(double* __restrict__ dst, double value, int64_t n) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) dst[k] = value;
}

__global__ __launch_bounds__(kBlock) void k_fill_i32(int32_t* __restrict__ dst, int32_t value, int64_t n) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) dst[k] = value;
}

__global__ __launch_bounds__(kBlock) void k_copy(const double* __restrict__ src, double* __restrict__ dst,
                                                  int64_t n) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

// v /= length of the instance (a zero length leaves v as is; the estimate is then 1).
__global__ __launch_bounds__(kBlock) void k_chunk_div(const Tile* __restrict__ tc,
                                                       const double* __restrict__ length,
                                                       double* __restrict__ v) {
  const Tile t = tc[blockIdx.x];
  const double len = length[t.inst];
  if (len == 0.0) return;
  for (int k = t.begin + threadIdx.x; k < t.end; k += kBlock) v[k] /= len;
}

// ---- initial state (pdhg.cpp:327-379) ----------------------------------------------------------

template <class T>
__global__ __launch_bounds__(kBlock) void k_init_iterates(Args<T> a, const double* __restrict__ lhat64,
                                                           const double* __restrict__ uhat64,
                                                           int32_t ncols, int32_t nrows) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < ncols) {
    const T x = dev_project<T>(static_cast<T>(0), a.lhat[k], a.uhat[k]);
    a.x0[k] = x;
    a.x1[k] = x;
    a.xrst[k] = x;
    a.xbar[k] = x;
    a.aty0[k] = static_cast<T>(0);
    a.aty1[k] = static_cast<T>(0);
    a.xsum[k] = static_cast<T>(0);
    if constexpr (!std::is_same<T, double>::value) a.xtot[k] = 0.0;
    // pdhg.cpp:377: best starts as the scaled start point, in FP64
    double b = 0.0;
    if (dev_finite_bound(lhat64[k]) && b < lhat64[k]) b = lhat64[k];
    else if (dev_finite_bound(uhat64[k]) && b > uhat64[k]) b = uhat64[k];
    a.bestx[k] = b;
  }
  if (k < nrows) {
    a.y0[k] = static_cast<T>(0);
    a.y1[k] = static_cast<T>(0);
    a.yrst[k] = static_cast<T>(0);
    a.ysum[k] = static_cast<T>(0);
    if constexpr (!std::is_same<T, double>::value) a.ytot[k] = 0.0;
    a.besty[k] = 0.0;
  }
}

// A warm start (sk_batch_warm_start): the columns and rows that have a start value (not NaN)
// begin there instead of at proj(0) and 0, in the scaled space of the device set-up:
// x-hat = proj(x / D_c), y-hat = y / D_r (pdhg.cpp:357-366 maps the other way). The running
// sums stay zero and the restart point and best point start at the warm point, as the cold
// start does at its own (pdhg.cpp:327-379).
template <class T>
__global__ __launch_bounds__(kBlock) void k_warm_iterates(Args<T> a, const double* __restrict__ wx,
                                                          const double* __restrict__ wy,
                                                          const double* __restrict__ lhat64,
                                                          const double* __restrict__ uhat64,
                                                          int32_t ncols, int32_t nrows) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < ncols && !isnan(wx[k])) {
    double v = wx[k] * (1.0 / a.dc[k]);
    if (v < lhat64[k]) v = lhat64[k];
    if (v > uhat64[k]) v = uhat64[k];
    const T x = dev_project<T>(static_cast<T>(v), a.lhat[k], a.uhat[k]);
    a.x0[k] = x;
    a.x1[k] = x;
    a.xrst[k] = x;
    a.xbar[k] = x;
    a.bestx[k] = v;
  }
  if (k < nrows && !isnan(wy[k])) {
    const double v = wy[k] * (1.0 / a.dr[k]);
    a.y0[k] = static_cast<T>(v);
    a.y1[k] = static_cast<T>(v);
    a.yrst[k] = static_cast<T>(v);
    a.besty[k] = v;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_copy_t(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

}  // namespace sk
