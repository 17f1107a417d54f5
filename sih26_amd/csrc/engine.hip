// Host driver of the batched restarted-PDHG engine (see engine.hpp and kernels.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "threads.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "status.hpp"

namespace sk {

namespace {

#define SK_HIP(call, code)                                                              \
  do {                                                                                  \
    const hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                             \
      set_last_error(std::string(#call) + ": " + hipGetErrorString(e_));                \
      return (code);                                                                    \
    }                                                                                   \
  } while (0)

// The step-size rule's factors (pdhg.cpp:472-474) for iterations [0, kRuleIterations), which
// covers the default limit of 10^6 iterations: at iteration k, with e = max(k + 1, 2),
// shrink = 1 - e^-0.3 and grow = 1 + e^-0.6, computed with the host's libm pow as the
// reference computes them. One table per process, copied once to each device and kept until
// the process ends; a decision then loads the factors instead of evaluating two FP64 powers.
constexpr int64_t kRuleIterations = (int64_t{1} << 20) + 1;

const std::vector<double>& rule_table_host() {
  static std::once_flag once;
  static std::vector<double> table;
  std::call_once(once, [] {
    table.resize(2 * static_cast<size_t>(kRuleIterations));
    for (int64_t k = 0; k < kRuleIterations; ++k) {
      const double exponent = static_cast<double>(k + 1 > 2 ? k + 1 : 2);
      table[2 * static_cast<size_t>(k)] = 1.0 - std::pow(exponent, -0.3);
      table[2 * static_cast<size_t>(k) + 1] = 1.0 + std::pow(exponent, -0.6);
    }
  });
  return table;
}

// The table on the current device (allocated and filled on first use).
int rule_table_device(const double** out) {
  static std::mutex mu;
  static std::vector<double*> per_device;
  int dev = 0;
  SK_HIP(hipGetDevice(&dev), kStatusInitFailed);
  const std::vector<double>& host = rule_table_host();
  std::lock_guard<std::mutex> lk(mu);
  if (per_device.size() <= static_cast<size_t>(dev)) per_device.resize(static_cast<size_t>(dev) + 1, nullptr);
  double*& d = per_device[static_cast<size_t>(dev)];
  if (!d) {
    double* p = nullptr;
    SK_HIP(hipMalloc(&p, host.size() * sizeof(double)), kStatusOutOfMemory);
    if (hipMemcpy(p, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      set_last_error("hipMemcpy of the step-size rule table failed");
      (void)hipFree(p);
      return kStatusCopyFailed;
    }
    d = p;
  }
  *out = d;
  return kStatusSuccess;
}

#define SK_LAUNCH(...)                                                                  \
  do {                                                                                  \
    if (fault_launch_fails()) {                                                         \
      set_last_error("kernel launch: injected failure (sk_fault_inject)");              \
      return kStatusLaunchFailed;                                                       \
    }                                                                                   \
    __VA_ARGS__;                                                                        \
    const hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess) {                                                             \
      set_last_error(std::string("kernel launch: ") + hipGetErrorString(e_));           \
      return kStatusLaunchFailed;                                                       \
    }                                                                                   \
  } while (0)

inline unsigned blocks_for(int64_t n) {
  return static_cast<unsigned>(std::max<int64_t>(1, (n + kBlock - 1) / kBlock));
}

// Waits for a stream when it goes out of scope: keeps host sources of queued asynchronous
// copies alive until the copies are done, on every return path.
struct StreamDrain {
  hipStream_t stream;
  ~StreamDrain() {
    if (stream) (void)hipStreamSynchronize(stream);
  }
};

// Device buffer owned by the engine.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (ptr) (void)hipFree(ptr);
  }
};

// Row-block tiles over rows [row0, row0 + rows) of a CSR matrix. Short rows (at most kLongRow
// nonzeros) share kTileShort tiles of at most kMaxTileRows rows; long rows share kTileLong
// tiles of at most kMaxTileRows / kWave rows; either kind holds at most `cap` nonzeros, and a row with more
// than `cap` nonzeros is a kTileVector tile of its own. A row's sum does not depend on the
// tiling (see kLongRow), so `cap` is free to tune. Every instance gets at least one tile.
void build_row_tiles(const std::vector<int32_t>& ptr, int32_t row0, int32_t rows, int32_t inst,
                     int32_t cap, int32_t base, std::vector<Tile>* out) {
  const size_t before = out->size();
  auto emit = [&](int32_t b, int32_t e, int32_t kind) {
    out->push_back(Tile{b, e, inst, cap, ptr[static_cast<size_t>(b)], ptr[static_cast<size_t>(e)], base, kind});
  };
  int32_t start = row0, kind = kTileShort;
  int64_t acc = 0;
  const int32_t end = row0 + rows;
  for (int32_t r = row0; r < end; ++r) {
    const int64_t len = ptr[static_cast<size_t>(r) + 1] - ptr[static_cast<size_t>(r)];
    if (len > cap) {
      if (start < r) emit(start, r, kind);
      emit(r, r + 1, kTileVector);
      start = r + 1;
      acc = 0;
      continue;
    }
    const int32_t k = len > kLongRow ? kTileLong : kTileShort;
    const int32_t max_rows = k == kTileLong ? kMaxTileRows / kWave : kMaxTileRows;
    if (start < r && (k != kind || acc + len > cap || r - start == max_rows)) {
      emit(start, r, kind);
      start = r;
      acc = 0;
    }
    if (start == r) kind = k;
    acc += len;
  }
  if (start < end) emit(start, end, kind);
  if (out->size() == before) emit(row0, row0, kTileShort);
}

// The resident solve (k_resident) applies when every instance has at most this many SpMV
// tiles (A and A^T together) and column chunks; a launch runs up to kResidentRounds
// steps_per_block rounds of each instance (steps or evaluation stages).
constexpr int32_t kFusedMaxTiles = 4;
constexpr int32_t kFusedMaxChunks = 2;
constexpr int32_t kResidentRounds = 4;

// A step decided inside the next primal launch is decided by every column chunk of the
// instance, each reading all of the instance's per-tile partials. That costs chunks x partials
// reads per launch, which grows with the square of an instance's size: a batch whose sum of it
// over the instances exceeds this many reads decides every step in a launch of its own
// (k_step_finalize, one workgroup per instance) instead. Measured on the 2M x 5M LP (4,883
// chunks, ~24,400 partials, 119M reads): a deciding primal launch 99 us against 28 us for the
// plain one plus 28 us for the decision's launch; on the 200k x 500k LP (1.5M reads) folding
// still saves ~2 us per step.
constexpr int64_t kFoldMaxReads = int64_t{1} << 24;


// Column order on the device. For an instance of at least kRelabelColumns columns (mode 0) or
// any instance (mode 1), its kHotBytes / w columns with the most nonzeros (the x-bar entries
// that the gathers of A x-bar touch most often) come first, largest first with ties in index
// order, and the other columns follow in their own order: that head of x-bar stays resident
// in every L2 while the rest streams, and the long rows of A^T share tiles. Measured: a 5M-
// and a 500k-column power-law LP solve 16% and 14% faster; instances of up to 15k columns
// neither gain nor lose. Equal degrees give the identity. Returns new -> old, or nothing for
// the identity. The order depends on the instance alone, so an instance gets the same order
// alone or in a batch.
constexpr size_t kRelabelColumns = size_t{1} << 16;
constexpr size_t kHotBytes = size_t{2} << 20;
std::vector<int32_t> relabel_columns(const HostLp& lp, size_t word, int32_t mode) {
  const auto n = static_cast<size_t>(lp.num_cols);
  if (mode == 2 || (mode == 0 && n < kRelabelColumns)) return {};
  const size_t hot = std::min(n, kHotBytes / word);
  // the head: columns by degree, largest first, ties in index order (a counting sort over the
  // degrees, which are at most the row count); the tail: the other columns in index order
  auto degree = [&](size_t j) {
    return static_cast<size_t>(lp.col_starts[j + 1] - lp.col_starts[j]);
  };
  size_t top = 0;
  for (size_t j = 0; j < n; ++j) top = std::max(top, degree(j));
  std::vector<size_t> first(top + 2, 0);  // first[top - d]: where degree d starts
  for (size_t j = 0; j < n; ++j) ++first[top - degree(j) + 1];
  for (size_t d = 0; d <= top; ++d) first[d + 1] += first[d];
  std::vector<int32_t> by_degree(n);
  for (size_t j = 0; j < n; ++j) by_degree[first[top - degree(j)]++] = static_cast<int32_t>(j);
  std::vector<int32_t> order(by_degree.begin(), by_degree.begin() + static_cast<std::ptrdiff_t>(hot));
  order.resize(n);
  std::vector<uint8_t> taken(n, 0);
  for (size_t k = 0; k < hot; ++k) taken[static_cast<size_t>(order[k])] = 1;
  size_t next = hot;
  for (size_t j = 0; j < n; ++j)
    if (!taken[j]) order[next++] = static_cast<int32_t>(j);
  bool identity = true;
  for (size_t j = 0; j < n && identity; ++j) identity = order[j] == static_cast<int32_t>(j);
  if (identity) order.clear();
  return order;
}

// f(k, scratch) for every instance k of a batch, on up to 16 host threads that take instances
// in turn (one thread below ~100k nonzeros); `scratch` is the calling thread's vector, each
// other thread gets its own. An exception in a worker is rethrown here after all have joined.
template <class V, class F>
void for_each_instance(int32_t count, size_t nnz, std::vector<V>* scratch, F&& f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t threads = std::min<int64_t>({16, hw ? static_cast<int64_t>(hw) : 1, count,
                                            static_cast<int64_t>(nnz / 100000) + 1});
  if (threads <= 1) {
    for (int32_t k = 0; k < count; ++k) f(k, *scratch);
    return;
  }
  std::atomic<int32_t> next{0};
  std::exception_ptr failed;
  std::mutex failed_mu;
  auto work = [&](std::vector<V>& mine) {
    try {
      for (int32_t k; (k = next.fetch_add(1)) < count;) f(k, mine);
    } catch (...) {
      std::lock_guard<std::mutex> lk(failed_mu);
      if (!failed) failed = std::current_exception();
      next.store(count);
    }
  };
  std::vector<std::thread> pool;
  {
    JoinAll joiner{pool};
    for (int64_t t = 1; t < threads; ++t)
      pool.emplace_back([&]() {
        std::vector<V> mine;
        work(mine);
      });
    work(*scratch);
  }
  if (failed) std::rethrow_exception(failed);
}

void build_chunks(int32_t begin, int32_t count, int32_t inst, std::vector<Tile>* out) {
  const size_t before = out->size();
  for (int32_t k = 0; k < count; k += kChunk)
    out->push_back(Tile{begin + k, begin + std::min(count, k + kChunk), inst, 0});
  if (out->size() == before) out->push_back(Tile{begin, begin, inst, 0});
}

template <class T>
class EngineImpl final : public Engine {
 public:
  explicit EngineImpl(const EngineOptions& o) : opt_(o) {}
  ~EngineImpl() override {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (setup_exec_) (void)hipGraphExecDestroy(setup_exec_);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
    if (host_done_) (void)hipHostFree(host_done_);
    if (host_state_) (void)hipHostFree(host_state_);
    for (hipEvent_t e : block_ev_)
      if (e) (void)hipEventDestroy(e);
  }

  int32_t add(const HostLp* lp) override {
    lps_.push_back(lp);
    warm_.emplace_back();
    return static_cast<int32_t>(lps_.size()) - 1;
  }
  int warm_start(int32_t index, const double* x, const double* y, double omega, double eta) override {
    if (index < 0 || static_cast<size_t>(index) >= lps_.size() || !x || !y) {
      set_last_error("warm_start(): no such instance, or a null start");
      return kStatusInvalidArgument;
    }
    const HostLp& lp = *lps_[static_cast<size_t>(index)];
    WarmStart& w = warm_[static_cast<size_t>(index)];
    w.set = true;
    w.x.assign(x, x + lp.num_cols);
    w.y.assign(y, y + lp.num_rows);
    w.omega = omega;
    w.eta = eta;
    return kStatusSuccess;
  }
  void clear() override {
    lps_.clear();
    warm_.clear();
  }
  int32_t size() const override { return static_cast<int32_t>(lps_.size()); }
  const EngineTimings& timings() const override { return timings_; }
  int64_t total_nnz() const override { return nnz_; }
  int64_t total_rows() const override { return m_; }
  int64_t total_cols() const override { return n_; }
  int stage_choice() const override { return stage_; }
  int32_t streams() const override { return 1; }
  int32_t fused() const override { return fused_ ? 1 : 0; }
  int32_t column_slices() const override { return slices_; }
  int32_t tile_capacity() const override { return cap_; }

  int upload() override;
  int solve() override;
  int start() override;
  int poll(bool wait, bool* finished, std::vector<int32_t>* newly_done) override;
  int64_t device_bytes() const override;
  void release() override;
  int result(int32_t index, InstanceResult* out) const override;
  int time_kernel(int kernel, int reps, double* mean_ms) override;
  double kernel_bytes(int kernel) const override;
  int probe_spmv(int32_t index, int which, const double* in, double* out) override;
  int probe_scaling(int32_t index, double* row_scale, double* col_scale, double* values) override;
  int column_order(int32_t index, int32_t* device_to_lp) const override;

 private:
  // Grow-only device buffers: a reloaded engine (a serving slot taking its next batch) keeps
  // every allocation that is already large enough.
  template <class U>
  int alloc(DevBuf& b, size_t count, U** typed) {
    const size_t need = std::max<size_t>(1, count) * sizeof(U);
    if (b.ptr && b.bytes < need) {
      (void)hipFree(b.ptr);
      b.ptr = nullptr;
      b.bytes = 0;
    }
    if (!b.ptr) {
      if (fault_allocation_fails()) {
        set_last_error("hipMalloc: injected failure (sk_fault_inject)");
        return kStatusOutOfMemory;
      }
      SK_HIP(hipMalloc(&b.ptr, need), kStatusOutOfMemory);
      b.bytes = need;
    }
    *typed = static_cast<U*>(b.ptr);
    return kStatusSuccess;
  }
  template <class U>
  int put(DevBuf& b, const std::vector<U>& host, U** typed) {
    int st = alloc(b, host.size(), typed);
    if (st) return st;
    if (!host.empty())
      SK_HIP(hipMemcpyAsync(b.ptr, host.data(), host.size() * sizeof(U), hipMemcpyHostToDevice,
                            stream_),
             kStatusCopyFailed);
    return kStatusSuccess;
  }

  // Stacked slices of one CSR table of the batch's one instance: row k * rows + r of the
  // result holds the entries of the table's row row0 + r whose gather index lies in slice k of
  // [g0, g0 + count) (slice k: [g0 + k count / S, g0 + (k + 1) count / S)), in the table's order,
  // with src the table position of each. A row's entries of one slice are one run of it, since
  // the gather indices ascend in every row (CSR(A) of a sliced instance keeps the LP's column
  // order; CSR(A^T) holds rows ascending). The row-block tiles go slice by slice, each slice's
  // part padded with empty tiles to one length that its 8 / S XCDs share evenly, so that
  // xcd_tile gives slice k's tiles to XCDs k 8 / S, ..., (k + 1) 8 / S - 1.
  void build_stacked(const std::vector<int32_t>& ptr, const std::vector<int32_t>& idx,
                     int32_t row0, int32_t rows, int32_t g0, int32_t count, int32_t S,
                     int32_t cap, std::vector<int32_t>& sp, std::vector<int32_t>& sidx,
                     std::vector<int32_t>& src, std::vector<Tile>& tiles) {
    const auto R = static_cast<size_t>(rows);
    std::vector<int32_t> bound(static_cast<size_t>(S) + 1);
    for (int32_t k = 0; k <= S; ++k)
      bound[static_cast<size_t>(k)] = g0 + static_cast<int32_t>(static_cast<int64_t>(count) * k / S);
    sp.assign(static_cast<size_t>(S) * R + 1, 0);
    auto walk = [&](int32_t r, auto&& f) {
      int32_t k = 0;
      for (int32_t p = ptr[static_cast<size_t>(row0 + r)]; p < ptr[static_cast<size_t>(row0 + r) + 1]; ++p) {
        while (idx[static_cast<size_t>(p)] >= bound[static_cast<size_t>(k) + 1]) ++k;
        f(k, p);
      }
    };
    for (int32_t r = 0; r < rows; ++r)
      walk(r, [&](int32_t k, int32_t) { ++sp[static_cast<size_t>(k) * R + static_cast<size_t>(r) + 1]; });
    for (size_t q = 0; q + 1 < sp.size(); ++q) sp[q + 1] += sp[q];
    const auto Z = static_cast<size_t>(sp.back());
    sidx.resize(Z);
    src.resize(Z);
    std::vector<int32_t>& next = stage_host_.next;
    next.assign(sp.begin(), sp.end() - 1);
    for (int32_t r = 0; r < rows; ++r)
      walk(r, [&](int32_t k, int32_t p) {
        const auto q = static_cast<size_t>(next[static_cast<size_t>(k) * R + static_cast<size_t>(r)]++);
        sidx[q] = idx[static_cast<size_t>(p)];
        src[q] = p;
      });
    std::vector<std::vector<Tile>> part(static_cast<size_t>(S));
    const size_t share = static_cast<size_t>(kXcds / S);  // XCDs per slice
    size_t longest = 0;
    for (int32_t k = 0; k < S; ++k) {
      build_row_tiles(sp, k * rows, rows, 0, cap, g0, &part[static_cast<size_t>(k)]);
      longest = std::max(longest, part[static_cast<size_t>(k)].size());
    }
    longest = (longest + share - 1) / share * share;
    tiles.clear();
    for (int32_t k = 0; k < S; ++k) {
      const std::vector<Tile>& tk = part[static_cast<size_t>(k)];
      tiles.insert(tiles.end(), tk.begin(), tk.end());
      const int32_t e = (k + 1) * rows;
      const int32_t pe = sp[static_cast<size_t>(e)];
      for (size_t u = tk.size(); u < longest; ++u) tiles.push_back(Tile{e, e, 0, cap, pe, pe, g0, kTileShort});
    }
  }

  int setup_device();
  int enqueue_setup();
  int capture(int (EngineImpl<T>::*body)(), hipGraphExec_t* exec);
  int run_loop();
  int report();
  int enqueue_block();
  int enqueue_eval(int report_mode);
  int launch_block();
  // One step SpMV launch (1: A xbar + dual step, 2: A^T y' + interaction), instantiated for
  // the staging capacity the tile tables were built with.
  // With slices, A x-bar is the column slices' product followed by their sum and the dual step,
  // and A^T y' the row slices' product followed by their sum and the interaction.
  int launch_spmv(int kernel, const Args<T>& a) {
    hipStream_t s = stream_;
    const unsigned grid = static_cast<unsigned>(kernel == 1 ? n_ta_ : n_tt_);
    return with_layout([&](auto cap, auto pack) -> int {
      constexpr int C = decltype(cap)::value;
      constexpr bool P = decltype(pack)::value;
      if (kernel == 1 && slices_) {
        SK_LAUNCH(k_slice_product<T, C, P><<<static_cast<unsigned>(n_ts_), kStepLanes, 0, s>>>(a));
        SK_LAUNCH(k_slice_dual<T><<<static_cast<unsigned>(n_tr_), kBlock, 0, s>>>(a));
      } else if (kernel == 1) {
        SK_LAUNCH(k_step_dual<T, C, P><<<grid, kStepLanes, 0, s>>>(a));
      } else if (slices_) {
        SK_LAUNCH(k_rslice_product<T, C, P><<<static_cast<unsigned>(n_tq_), kStepLanes, 0, s>>>(a));
        SK_LAUNCH(k_rslice_aty<T><<<static_cast<unsigned>(n_tc_), kBlock, 0, s>>>(a));
      } else {
        SK_LAUNCH(k_step_aty<T, C, P><<<grid, kStepLanes, 0, s>>>(a));
      }
      return kStatusSuccess;
    });
  }
  // Calls f(capacity, packing) with the staging capacity (sizing LDS) and the index packing
  // of the tile tables as compile-time constants.
  template <class F>
  int with_layout(F f) {
    auto packed = [&](auto cap) -> int {
      return pack_ ? f(cap, std::true_type{}) : f(cap, std::false_type{});
    };
    if (cap_ <= 1024) return packed(std::integral_constant<int, 1024>{});
    if (cap_ <= kDefaultTileNnz) return packed(std::integral_constant<int, kDefaultTileNnz>{});
    return packed(std::integral_constant<int, kMaxTileNnz>{});
  }
  int extract_finished(const InstState* snapshot, std::vector<int32_t>* newly_done);
  int stop_on_time_limit(std::vector<int32_t>* newly_done);

  EngineOptions opt_;
  std::vector<const HostLp*> lps_;
  std::vector<WarmStart> warm_;  // per instance, applied at upload
  bool warm_any_ = false;        // the uploaded batch has a warm start
  int64_t m_ = 0, n_ = 0, nnz_ = 0;
  int32_t K_ = 0;
  int stage_ = 1;
  int32_t cap_ = kDefaultTileNnz;  // row-block tile capacity (nonzeros)
  bool fused_ = false;  // the resident solve (k_resident): every instance has few tiles
  bool fold_ = true;    // steps decided in the next primal launch (kFoldMaxReads)
  bool pack_ = true;    // packed index words in the staging order
  int32_t slices_ = 0;  // column slices of A x-bar (kColumnSlices), or 0
  int32_t cus_ = 256;   // compute units of the device

  // host staging of an upload (the sources of its copies), reused by the next one
  struct HostStage {
    std::vector<int32_t> a_ptr, a_idx, t_ptr, t_idx, next, dev_of, s_ptr, s_idx, s_src, q_ptr,
        q_idx, q_src;
    std::vector<double> a_orig, t_orig, cmin, cl, cu, rl, ru, draw;
    std::vector<Tile> ta, tt, tc, tr, ts, tq;
    std::vector<InstDesc> cdesc;
  };
  HostStage stage_host_;
  // host copies needed after upload
  std::vector<InstDesc> desc_;
  std::vector<InstState> init_state_;
  std::vector<int32_t> sense_;
  std::vector<double> v0_;  // power-iteration start, all instances
  std::vector<double> wx_host_, wy_host_;  // warm start in the device order (copy sources)
  std::vector<uint8_t> trivial_norm_;
  int32_t n_ta_ = 0, n_tt_ = 0, n_tc_ = 0, n_tr_ = 0, n_ts_ = 0, n_tq_ = 0;
  uint32_t initially_done_ = 0;
  bool setup_done_ = false;  // the device set-up of the uploaded batch has been enqueued

  hipStream_t stream_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  hipGraphExec_t setup_exec_ = nullptr;
  uint32_t* host_done_ = nullptr;

  // device buffers
  DevBuf b_a_ptr_, b_a_idx_, b_a_orig_, b_t_ptr_, b_t_idx_, b_t_orig_;
  DevBuf b_sa_, b_st_, b_a_val_, b_t_val_, b_a_sidx_, b_t_sidx_, b_a_pos_, b_t_pos_;
  DevBuf b_cmin_, b_cl_, b_cu_, b_rl_, b_ru_, b_dc_, b_dr_, b_rpass_, b_cpass_;
  DevBuf b_chat_, b_lhat_, b_uhat_, b_rlhat_, b_ruhat_, b_lhat64_, b_uhat64_;
  DevBuf b_x0_, b_x1_, b_y0_, b_y1_, b_aty0_, b_aty1_, b_xbar_, b_xrst_, b_yrst_;
  DevBuf b_xsum_, b_ysum_, b_xtot_, b_ytot_, b_xu0_, b_xu1_, b_yu0_, b_yu1_, b_bestx_, b_besty_, b_act_, b_red_;
  DevBuf b_tc_, b_tr_, b_ta_, b_tt_, b_desc_, b_cdesc_, b_state_;
  DevBuf b_ps_c_, b_ps_a_, b_ps_t_, b_pe_c_, b_pe_r_, b_pe_a_, b_pe_t_;
  DevBuf b_done_, b_v_, b_v0_, b_av_, b_len_, b_zero_, b_fa_, b_ft_, b_fc_, b_fr_;
  DevBuf b_a_orig_st_, b_t_orig_st_;
  DevBuf b_wx_, b_wy_;  // warm start, device order, NaN where cold
  DevBuf b_s_ptr_, b_s_idx_, b_s_src_, b_s_sidx_, b_s_pos_, b_s_val_, b_ts_, b_xpart_, b_ps_r_;
  DevBuf b_q_ptr_, b_q_idx_, b_q_src_, b_q_sidx_, b_q_pos_, b_q_val_, b_tq_, b_ypart_, b_ps_q_;
  std::vector<DevBuf*> buffers() {
    return {&b_a_ptr_, &b_a_idx_, &b_a_orig_, &b_t_ptr_, &b_t_idx_, &b_t_orig_, &b_sa_, &b_st_,
            &b_a_val_, &b_t_val_, &b_a_sidx_, &b_t_sidx_, &b_a_pos_, &b_t_pos_, &b_cmin_, &b_cl_, &b_cu_, &b_rl_, &b_ru_, &b_dc_, &b_dr_,
            &b_rpass_, &b_cpass_, &b_chat_, &b_lhat_, &b_uhat_, &b_rlhat_, &b_ruhat_,
            &b_lhat64_, &b_uhat64_, &b_x0_, &b_x1_, &b_y0_, &b_y1_, &b_aty0_, &b_aty1_,
            &b_xbar_, &b_xrst_, &b_yrst_, &b_xsum_, &b_ysum_, &b_xtot_, &b_ytot_, &b_xu0_, &b_xu1_, &b_yu0_,
            &b_yu1_, &b_bestx_, &b_besty_, &b_act_, &b_red_, &b_tc_, &b_tr_, &b_ta_, &b_tt_,
            &b_desc_, &b_cdesc_, &b_state_, &b_ps_c_, &b_ps_a_, &b_ps_t_, &b_pe_c_, &b_pe_r_, &b_pe_a_,
            &b_pe_t_, &b_done_, &b_v_, &b_v0_, &b_av_, &b_len_, &b_zero_, &b_fa_, &b_ft_,
            &b_fc_, &b_fr_, &b_a_orig_st_, &b_t_orig_st_, &b_wx_, &b_wy_, &b_s_ptr_, &b_s_idx_,
            &b_s_src_, &b_s_sidx_, &b_s_pos_, &b_s_val_, &b_ts_, &b_xpart_, &b_ps_r_, &b_q_ptr_,
            &b_q_idx_, &b_q_src_, &b_q_sidx_, &b_q_pos_, &b_q_val_, &b_tq_, &b_ypart_, &b_ps_q_};
  }
  std::vector<const DevBuf*> buffers() const {
    auto all = const_cast<EngineImpl<T>*>(this)->buffers();
    return std::vector<const DevBuf*>(all.begin(), all.end());
  }

  Args<T> args_{};
  double* d_sa_ = nullptr;
  double* d_st_ = nullptr;
  const int32_t* s_src_dev_ = nullptr;  // the column slices' CSR(A) positions (build_stacked)
  const int32_t* q_src_dev_ = nullptr;  // the row slices' CSR(A^T) positions
  double* d_rpass_ = nullptr;
  double* d_cpass_ = nullptr;
  double* d_lhat64_ = nullptr;
  double* d_uhat64_ = nullptr;
  double* d_v_ = nullptr;
  double* d_v0_ = nullptr;
  double* d_av_ = nullptr;
  double* d_len_ = nullptr;
  int32_t* d_zero_seen_ = nullptr;
  double* d_wx_ = nullptr;
  double* d_wy_ = nullptr;

  // serving: blocks in flight (a ring of 2: event, finished count and state snapshot per
  // block), instances already read back, and a second stream for those read-backs
  hipStream_t copy_stream_ = nullptr;
  InstState* host_state_ = nullptr;  // pinned, 2 * capacity records
  size_t host_state_cap_ = 0;
  hipEvent_t block_ev_[2] = {nullptr, nullptr};
  int64_t launched_ = 0;
  int inflight_ = 0;
  uint32_t extracted_count_ = 0;
  std::chrono::steady_clock::time_point started_;
  std::vector<uint8_t> extracted_;

  // results
  std::vector<InstState> final_state_;
  std::vector<double> bestx_, besty_, act_, red_;
  // per instance, device column -> the LP's column (relabel_columns); empty for the identity
  std::vector<std::vector<int32_t>> col_perm_;
  EngineTimings timings_;
};

template <class T>
int EngineImpl<T>::upload() {
  // a reload (new instance set) invalidates the captured graphs: grids and tables change
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (setup_exec_) (void)hipGraphExecDestroy(setup_exec_);
  graph_exec_ = setup_exec_ = nullptr;
  inflight_ = 0;
  setup_done_ = false;
  K_ = static_cast<int32_t>(lps_.size());
  m_ = n_ = nnz_ = 0;
  initially_done_ = 0;
  for (size_t k = 0; k < lps_.size(); ++k) {
    const HostLp* lp = lps_[k];
    const std::string problem = lp->validate();
    if (!problem.empty()) {
      set_last_error("instance " + std::to_string(k) + ": " + problem);
      return kStatusInvalidArgument;
    }
    m_ += lp->num_rows;
    n_ += lp->num_cols;
    nnz_ += lp->nnz();
  }
  if (m_ >= INT32_MAX || n_ >= INT32_MAX || nnz_ >= INT32_MAX) {
    set_last_error("batch exceeds 32-bit row, column or nonzero offsets");
    return kStatusInvalidArgument;
  }
  slices_ = K_ == 1 && opt_.column_slices != 2 && lps_[0]->num_cols >= kColumnSlices &&
                    (opt_.column_slices == 1 || lps_[0]->num_cols >= kSliceMinCols)
                ? kColumnSlices
                : 0;
  if (static_cast<int64_t>(slices_) * m_ >= INT32_MAX || static_cast<int64_t>(kRowSlices) * n_ >= INT32_MAX ||
      m_ < kRowSlices)
    slices_ = 0;
  const auto M = static_cast<size_t>(m_), N = static_cast<size_t>(n_), Z = static_cast<size_t>(nnz_);

  // Block-diagonal CSR(A) (columns ascending in each row, the order of the reference's CSC
  // scatter) and CSR(A^T) (= the CSC of A, rows ascending in each column).
  // host staging of the device arrays, kept by the engine between uploads so a serving slot
  // does not fault in fresh pages for every batch (every element is written below)
  HostStage& h = stage_host_;
  std::vector<int32_t>& a_ptr = h.a_ptr;
  std::vector<int32_t>& a_idx = h.a_idx;
  std::vector<int32_t>& t_ptr = h.t_ptr;
  std::vector<int32_t>& t_idx = h.t_idx;
  std::vector<double>& a_orig = h.a_orig;
  std::vector<double>& t_orig = h.t_orig;
  std::vector<double>& cmin = h.cmin;
  std::vector<double>& cl = h.cl;
  std::vector<double>& cu = h.cu;
  std::vector<double>& rl = h.rl;
  std::vector<double>& ru = h.ru;
  a_ptr.assign(M + 1, 0);
  t_ptr.assign(N + 1, 0);
  a_idx.resize(Z);
  t_idx.resize(Z);
  a_orig.resize(Z);
  t_orig.resize(Z);
  cmin.resize(N);
  cl.resize(N);
  cu.resize(N);
  rl.resize(M);
  ru.resize(M);
  desc_.assign(static_cast<size_t>(K_), InstDesc{});
  init_state_.assign(static_cast<size_t>(K_), InstState{});
  sense_.assign(static_cast<size_t>(K_), 1);
  trivial_norm_.assign(static_cast<size_t>(K_), 0);
  v0_.assign(N, 0.0);
  std::vector<Tile>& ta = h.ta;
  std::vector<Tile>& tt = h.tt;
  std::vector<Tile>& tc = h.tc;
  std::vector<Tile>& tr = h.tr;
  ta.clear();
  tt.clear();
  tc.clear();
  tr.clear();
  col_perm_.assign(static_cast<size_t>(K_), {});
  const int64_t limit = opt_.iteration_limit < 0 ? 1000000 : opt_.iteration_limit;
  // every instance's offsets first, then the instances prepared independently (their rows,
  // columns and nonzeros are disjoint ranges of the batch arrays), on host threads
  std::vector<int32_t> row_off(static_cast<size_t>(K_) + 1, 0), col_off(row_off), nz_off(row_off);
  for (int32_t k = 0; k < K_; ++k) {
    const HostLp& lp = *lps_[static_cast<size_t>(k)];
    row_off[static_cast<size_t>(k) + 1] = row_off[static_cast<size_t>(k)] + lp.num_rows;
    col_off[static_cast<size_t>(k) + 1] = col_off[static_cast<size_t>(k)] + lp.num_cols;
    nz_off[static_cast<size_t>(k) + 1] = nz_off[static_cast<size_t>(k)] + static_cast<int32_t>(lp.nnz());
  }
  for_each_instance(K_, Z, &h.draw, [&](int32_t k, std::vector<double>& draw) {
    const HostLp& lp = *lps_[static_cast<size_t>(k)];
    const int32_t m = lp.num_rows, n = lp.num_cols;
    const int32_t row0 = row_off[static_cast<size_t>(k)], col0 = col_off[static_cast<size_t>(k)];
    const int32_t nz0 = nz_off[static_cast<size_t>(k)];
    // counts per row
    for (int32_t p = 0; p < lp.nnz(); ++p) ++a_ptr[static_cast<size_t>(row0 + lp.row_indices[p]) + 1];
    std::vector<int32_t>& perm = col_perm_[static_cast<size_t>(k)];
    perm = relabel_columns(lp, sizeof(T), slices_ ? 2 : opt_.column_order);
    auto lp_col = [&](int32_t j) { return perm.empty() ? j : perm[static_cast<size_t>(j)]; };
    int32_t nz = nz0;
    if (k == 0) t_ptr[0] = 0;  // every later start is the end its predecessor writes
    for (int32_t j = 0; j < n; ++j) {
      const int32_t o = lp_col(j);
      for (int32_t p = lp.col_starts[o]; p < lp.col_starts[o + 1]; ++p, ++nz) {
        t_idx[static_cast<size_t>(nz)] = row0 + lp.row_indices[p];
        t_orig[static_cast<size_t>(nz)] = lp.values[p];
      }
      t_ptr[static_cast<size_t>(col0 + j) + 1] = nz;
    }
    const double sense = static_cast<double>(lp.sense);
    sense_[static_cast<size_t>(k)] = lp.sense;
    double cost_square = 0.0, bound_square = 0.0;
    for (int32_t j = 0; j < n; ++j) cost_square += (sense * lp.cost[j]) * (sense * lp.cost[j]);
    for (int32_t j = 0; j < n; ++j) {
      const int32_t o = lp_col(j);
      cmin[static_cast<size_t>(col0 + j)] = sense * lp.cost[o];
      cl[static_cast<size_t>(col0 + j)] = lp.col_lower[o];
      cu[static_cast<size_t>(col0 + j)] = lp.col_upper[o];
    }
    for (int32_t i = 0; i < m; ++i) {
      rl[static_cast<size_t>(row0 + i)] = lp.row_lower[i];
      ru[static_cast<size_t>(row0 + i)] = lp.row_upper[i];
      const double b = is_finite_bound(lp.row_lower[i])
                           ? lp.row_lower[i]
                           : (is_finite_bound(lp.row_upper[i]) ? lp.row_upper[i] : 0.0);
      bound_square += b * b;
    }
    // power-iteration start (scaling.cpp:168-171), the same libstdc++ engine and distribution
    if (m > 0 && n > 0 && lp.nnz() > 0) {
      // drawn in the LP's column order, placed in the device order
      std::mt19937 rng(static_cast<unsigned>(opt_.random_seed) + 1u);
      std::uniform_real_distribution<double> spread(-1.0, 1.0);
      draw.resize(static_cast<size_t>(n));
      for (int32_t j = 0; j < n; ++j) draw[static_cast<size_t>(j)] = spread(rng);
      for (int32_t j = 0; j < n; ++j) v0_[static_cast<size_t>(col0 + j)] = draw[static_cast<size_t>(lp_col(j))];
    } else {
      trivial_norm_[static_cast<size_t>(k)] = 1;
      init_state_[static_cast<size_t>(k)].norm_trivial = 1;
    }
    InstDesc& d = desc_[static_cast<size_t>(k)];
    d.row0 = row0;
    d.rows = m;
    d.col0 = col0;
    d.cols = n;
    InstState& s = init_state_[static_cast<size_t>(k)];
    s.tolerance = opt_.tolerance;
    s.cost_norm = std::sqrt(cost_square);
    s.bound_norm = std::sqrt(bound_square);
    s.omega = 1.0;
    s.rule_shrink = rule_table_host()[0];
    s.rule_grow = rule_table_host()[1];
    if (k < static_cast<int32_t>(warm_.size()) && warm_[static_cast<size_t>(k)].set) {
      const WarmStart& w = warm_[static_cast<size_t>(k)];
      if (w.omega > 0.0) s.omega = std::min(1e6, std::max(1e-6, w.omega));  // pdhg.cpp:627-629
      if (w.eta > 0.0) s.eta_start = w.eta;
    }
    s.restart_kkt = INFINITY;
    s.best_worst = INFINITY;
    s.iteration_limit = limit;
    s.use_restarts = opt_.restart != 0;
    s.stop_at_request = opt_.stop_at_request != 0;
    s.absolute_clause = opt_.absolute_clause != 0;
    s.best.primal = s.best.dual = s.best.gap = INFINITY;
    if (limit <= 0) {  // pdhg.cpp:384: the loop stops before its first step
      s.done = 1;
      s.status = kStIterationLimit;
      s.report_pending = 1;
    }
  });
  if (limit <= 0) initially_done_ = static_cast<uint32_t>(K_);
  {
    // an injected fault: a NaN cost in the first column of one instance (its first step is
    // non-finite, the others run on)
    const int64_t bad = fault_take_nan_instance();
    if (bad >= 0 && bad < K_ && desc_[static_cast<size_t>(bad)].cols > 0)
      cmin[static_cast<size_t>(desc_[static_cast<size_t>(bad)].col0)] = std::nan("");
  }
  for (size_t i = 0; i < M; ++i) a_ptr[i + 1] += a_ptr[i];
  {
    // Each row of CSR(A) holds its entries in the LP's column order, also where the device
    // holds the columns in another order: that is the order in which the reference's column
    // scatter (sparse.cpp:202-212) and its row statistics (scaling.cpp:44-100) add a row's
    // terms, so row sums (scaling, A x-bar, the evaluation's activities) do not depend on the
    // device order. The staging order sorts by device column anyway, for the gathers.
    // The scatter reads the LP's own column-major arrays in order (its columns ascending, each
    // column's rows as stored) and writes each entry to the next free slot of its row. A large
    // instance alone in the batch is scattered by several host threads, each owning a range of
    // rows and skipping the entries of the others: every row still receives its entries in the
    // same order, so the arrays are the same as one thread's.
    std::vector<int32_t>& next = h.next;
    next.assign(a_ptr.begin(), a_ptr.end() - 1);
    const unsigned hw = std::thread::hardware_concurrency();
    const int32_t row_threads = K_ == 1 ? static_cast<int32_t>(std::min<int64_t>(
                                              {8, hw ? static_cast<int64_t>(hw) : 1,
                                               static_cast<int64_t>(Z >> 17) + 1}))
                                        : 1;
    for_each_instance(K_, Z, &h.dev_of, [&](int32_t k, std::vector<int32_t>& dev_of) {
      const HostLp& lp = *lps_[static_cast<size_t>(k)];
      const InstDesc& d = desc_[static_cast<size_t>(k)];
      const std::vector<int32_t>& perm = col_perm_[static_cast<size_t>(k)];  // device -> LP
      dev_of.assign(perm.size(), 0);
      for (size_t j = 0; j < perm.size(); ++j) dev_of[static_cast<size_t>(perm[j])] = static_cast<int32_t>(j);
      auto scatter = [&](int32_t r_lo, int32_t r_hi) {
        for (int32_t o = 0; o < d.cols; ++o) {
          const int32_t j = d.col0 + (perm.empty() ? o : dev_of[static_cast<size_t>(o)]);
          for (int32_t p = lp.col_starts[o]; p < lp.col_starts[o + 1]; ++p) {
            const int32_t r = lp.row_indices[p];
            if (r < r_lo || r >= r_hi) continue;
            const auto slot = static_cast<size_t>(next[static_cast<size_t>(d.row0 + r)]++);
            a_idx[slot] = j;
            a_orig[slot] = lp.values[p];
          }
        }
      };
      if (row_threads <= 1) {
        scatter(0, d.rows);
        return;
      }
      std::vector<std::thread> pool;
      JoinAll joiner{pool};
      for (int32_t t = 1; t < row_threads; ++t)
        pool.emplace_back(scatter, static_cast<int32_t>(static_cast<int64_t>(d.rows) * t / row_threads),
                          static_cast<int32_t>(static_cast<int64_t>(d.rows) * (t + 1) / row_threads));
      scatter(0, static_cast<int32_t>(d.rows / row_threads));
    });
  }
  // Staging order of the SpMV tiles: sorted by gather index unless told otherwise.
  stage_ = opt_.stage < 0 ? 1 : opt_.stage;
  auto build_tables = [&](const int32_t cap) {
    cap_ = cap;
    ta.clear();
    tt.clear();
    tc.clear();
    tr.clear();
    for (int32_t k = 0; k < K_; ++k) {
      InstDesc& d = desc_[static_cast<size_t>(k)];
      d.ta0 = static_cast<int32_t>(ta.size());
      build_row_tiles(a_ptr, d.row0, d.rows, k, cap, d.col0, &ta);
      d.ta1 = static_cast<int32_t>(ta.size());
      d.tt0 = static_cast<int32_t>(tt.size());
      build_row_tiles(t_ptr, d.col0, d.cols, k, cap, d.row0, &tt);
      d.tt1 = static_cast<int32_t>(tt.size());
      d.tc0 = static_cast<int32_t>(tc.size());
      build_chunks(d.col0, d.cols, k, &tc);
      d.tc1 = static_cast<int32_t>(tc.size());
      d.tr0 = static_cast<int32_t>(tr.size());
      build_chunks(d.row0, d.rows, k, &tr);
      d.tr1 = static_cast<int32_t>(tr.size());
    }
  };
  build_tables(opt_.tile_nnz > 0 ? std::min(opt_.tile_nnz, kMaxTileNnz) : kDefaultTileNnz);
  // A batch whose A^T launch would not fill one round of the workgroups the device holds at
  // once (8 per CU) is bound by its slowest tile, and smaller tiles shorten that. Measured at
  // 1,024 / 1,280 / 1,536 / 2,048 / 2,560 / 3,072 nonzeros on the 200k x 500k power-law LP
  // (1,634 tiles of A^T at 2,560): 3.28 / 3.29 / 3.22 / 3.25 / 3.40 / 3.53 s at the same
  // iterations; on the serving stream (3,000 requests, batches of ~5) 1,536 for every batch gives
  // latency p50 35.0-35.5 / p99 79-80 ms against 36.9-37.7 / 87 ms at 2,560, and 35.8 / 92-94 ms
  // when batches that can run as the resident solve at 2,560 kept it. A batch that fills a round
  // (the 256-LP batch: 3,931 tiles) keeps the default, and so does a sub-batch of a split batch,
  // whose launches share the round with its siblings' (three sub-batches of the 256-LP batch at
  // 1,536: 61.0 ms per solve against 56.3 ms at 2,560).
  // Below a quarter of a round, 1,024: the serving stream's p50 34.0-34.5 ms against 35.1-35.5
  // at 1,536, and the 500 x 1,000 LP 11.7 against 12.5 ms.
  const int64_t round = 8 * static_cast<int64_t>(cus_);
  if (opt_.tile_nnz == 0 && opt_.split_parts <= 1 && static_cast<int64_t>(tt.size()) < round)
    build_tables(4 * static_cast<int64_t>(tt.size()) < round ? kQuarterRoundTileNnz : kOneRoundTileNnz);
  n_ta_ = static_cast<int32_t>(ta.size());
  n_tt_ = static_cast<int32_t>(tt.size());
  std::vector<Tile>& ts = h.ts;
  ts.clear();
  std::vector<Tile>& tq = h.tq;
  tq.clear();
  if (slices_) {
    const InstDesc& d = desc_[0];
    build_stacked(h.a_ptr, h.a_idx, d.row0, d.rows, d.col0, d.cols, slices_, cap_, h.s_ptr, h.s_idx,
                  h.s_src, ts);
    build_stacked(h.t_ptr, h.t_idx, d.col0, d.cols, d.row0, d.rows, kRowSlices, cap_, h.q_ptr,
                  h.q_idx, h.q_src, tq);
  }
  n_tq_ = static_cast<int32_t>(tq.size());
  n_ts_ = static_cast<int32_t>(ts.size());
  // A step of an instance with a few tiles is a chain of short dependent launches; one
  // workgroup then runs its tiles one after another in a single launch
  fused_ = (opt_.step_fusion & 1) == 0 && K_ > 0 && !slices_;
  for (const InstDesc& d : desc_)
    fused_ = fused_ && (d.ta1 - d.ta0) + (d.tt1 - d.tt0) <= kFusedMaxTiles && d.tc1 - d.tc0 <= kFusedMaxChunks;
  int64_t fold_reads = 0;
  for (const InstDesc& d : desc_) {
    const int64_t chunks = d.tc1 - d.tc0;
    fold_reads += chunks * (chunks + (d.ta1 - d.ta0) + (d.tt1 - d.tt0));
  }
  fold_ = (opt_.step_fusion & 2) == 0 && fold_reads <= kFoldMaxReads;
  n_tc_ = static_cast<int32_t>(tc.size());
  n_tr_ = static_cast<int32_t>(tr.size());

  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) {
    set_last_error("no HIP device is available");
    return kStatusNoDevice;
  }
  if (!stream_) {
    SK_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), kStatusInitFailed);
    int dev = 0, cus = 0;
    SK_HIP(hipGetDevice(&dev), kStatusInitFailed);
    SK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), kStatusInitFailed);
    cus_ = std::max(1, cus);
  }
  if (!host_done_) SK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_done_), 4 * sizeof(uint32_t)), kStatusOutOfMemory);
  if (host_state_cap_ < static_cast<size_t>(K_)) {
    if (host_state_) (void)hipHostFree(host_state_);
    host_state_ = nullptr;
    host_state_cap_ = 0;
    SK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_state_), 2 * static_cast<size_t>(K_) * sizeof(InstState)),
           kStatusOutOfMemory);
    host_state_cap_ = static_cast<size_t>(K_);
  }
  for (hipEvent_t& e : block_ev_)
    if (!e) SK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), kStatusInitFailed);
  // The host vectors above are the sources of the asynchronous copies queued below and are
  // destroyed when upload() returns; every exit from here on, an error return included, first
  // waits for the stream (declared after the vectors, so destroyed before them).
  const StreamDrain drain{stream_};

  Args<T>& a = args_;
  int st = 0;
  int32_t *ap, *ai, *tp, *ti;
  double *ao, *to;
#define SK_TRY(expr) \
  if ((st = (expr)) != 0) return st
  SK_TRY(put(b_a_ptr_, a_ptr, &ap));
  SK_TRY(put(b_a_idx_, a_idx, &ai));
  SK_TRY(put(b_a_orig_, a_orig, &ao));
  SK_TRY(put(b_t_ptr_, t_ptr, &tp));
  SK_TRY(put(b_t_idx_, t_idx, &ti));
  SK_TRY(put(b_t_orig_, t_orig, &to));
  a.a_ptr = ap;
  a.a_idx = ai;
  a.a_orig = ao;
  a.t_ptr = tp;
  a.t_idx = ti;
  a.t_orig = to;
  SK_TRY(alloc(b_sa_, Z, &d_sa_));
  SK_TRY(alloc(b_st_, Z, &d_st_));
  T *av, *tv;
  // the staging-order arrays carry kStagePad elements of slack (read in aligned groups of 4)
  SK_TRY(alloc(b_a_val_, Z + kStagePad, &av));
  SK_TRY(alloc(b_t_val_, Z + kStagePad, &tv));
  a.a_val = av;
  a.t_val = tv;
  double *pcm, *pcl, *pcu, *prl, *pru, *pdc, *pdr;
  SK_TRY(put(b_cmin_, cmin, &pcm));
  SK_TRY(put(b_cl_, cl, &pcl));
  SK_TRY(put(b_cu_, cu, &pcu));
  SK_TRY(put(b_rl_, rl, &prl));
  SK_TRY(put(b_ru_, ru, &pru));
  SK_TRY(alloc(b_dc_, N, &pdc));
  SK_TRY(alloc(b_dr_, M, &pdr));
  a.cmin = pcm;
  a.cl = pcl;
  a.cu = pcu;
  a.rl = prl;
  a.ru = pru;
  a.dc = pdc;
  a.dr = pdr;
  SK_TRY(alloc(b_rpass_, M, &d_rpass_));
  SK_TRY(alloc(b_cpass_, N, &d_cpass_));
  T *chat, *lhat, *uhat, *rlhat, *ruhat;
  SK_TRY(alloc(b_chat_, N, &chat));
  SK_TRY(alloc(b_lhat_, N, &lhat));
  SK_TRY(alloc(b_uhat_, N, &uhat));
  SK_TRY(alloc(b_rlhat_, M, &rlhat));
  SK_TRY(alloc(b_ruhat_, M, &ruhat));
  a.chat = chat;
  a.lhat = lhat;
  a.uhat = uhat;
  a.rlhat = rlhat;
  a.ruhat = ruhat;
  SK_TRY(alloc(b_lhat64_, N, &d_lhat64_));
  SK_TRY(alloc(b_uhat64_, N, &d_uhat64_));
  SK_TRY(alloc(b_x0_, N, &a.x0));
  SK_TRY(alloc(b_x1_, N, &a.x1));
  SK_TRY(alloc(b_y0_, M, &a.y0));
  SK_TRY(alloc(b_y1_, M, &a.y1));
  SK_TRY(alloc(b_aty0_, N, &a.aty0));
  SK_TRY(alloc(b_aty1_, N, &a.aty1));
  SK_TRY(alloc(b_xbar_, N, &a.xbar));
  SK_TRY(alloc(b_xrst_, N, &a.xrst));
  SK_TRY(alloc(b_yrst_, M, &a.yrst));
  SK_TRY(alloc(b_xsum_, N, &a.xsum));
  SK_TRY(alloc(b_ysum_, M, &a.ysum));
  {
    // FP64 totals of the FP32 running sums (unused with FP64 iterates)
    const bool split = !std::is_same<T, double>::value;
    SK_TRY(alloc(b_xtot_, split ? N : 0, &a.xtot));
    SK_TRY(alloc(b_ytot_, split ? M : 0, &a.ytot));
  }
  SK_TRY(alloc(b_xu0_, N, &a.xu0));
  SK_TRY(alloc(b_xu1_, N, &a.xu1));
  SK_TRY(alloc(b_yu0_, M, &a.yu0));
  SK_TRY(alloc(b_yu1_, M, &a.yu1));
  SK_TRY(alloc(b_bestx_, N, &a.bestx));
  SK_TRY(alloc(b_besty_, M, &a.besty));
  SK_TRY(alloc(b_act_, M, &a.activity));
  SK_TRY(alloc(b_red_, N, &a.reduced));
  Tile *ptc, *ptr_, *pta, *ptt;
  SK_TRY(put(b_tc_, tc, &ptc));
  SK_TRY(put(b_tr_, tr, &ptr_));
  SK_TRY(put(b_ta_, ta, &pta));
  SK_TRY(put(b_tt_, tt, &ptt));
  a.tc = ptc;
  a.tr = ptr_;
  a.ta = pta;
  a.tt = ptt;
  {
    // staging order of the step kernels' tiles (structure only, so once per upload)
    // Packed index words (4 B per nonzero instead of 6) when every instance's local indices
    // fit 20 bits; the slot arrays are then not needed.
    pack_ = cap_ <= kMaxTileNnz;
    for (const InstDesc& d : desc_) pack_ = pack_ && d.rows < kPackMaxLocal && d.cols < kPackMaxLocal;
    int32_t *asx, *tsx;
    uint16_t *apos, *tpos;
    SK_TRY(alloc(b_a_sidx_, Z + kStagePad, &asx));
    SK_TRY(alloc(b_t_sidx_, Z + kStagePad, &tsx));
    SK_TRY(alloc(b_a_pos_, pack_ ? 0 : Z + kStagePad, &apos));
    SK_TRY(alloc(b_t_pos_, pack_ ? 0 : Z + kStagePad, &tpos));
    const int32_t sorted = stage_ != 0, pack = pack_ ? 1 : 0;
    SK_LAUNCH(k_sort_tiles<<<n_ta_, kBlock, 0, stream_>>>(pta, ap, ai, sorted, pack, asx, apos));
    SK_LAUNCH(k_sort_tiles<<<n_tt_, kBlock, 0, stream_>>>(ptt, tp, ti, sorted, pack, tsx, tpos));
    a.a_sidx = asx;
    a.t_sidx = tsx;
    a.a_pos = apos;
    a.t_pos = tpos;
    // the original FP64 values in the same order, for the evaluation kernels
    double *aos, *tos;
    SK_TRY(alloc(b_a_orig_st_, Z + kStagePad, &aos));
    SK_TRY(alloc(b_t_orig_st_, Z + kStagePad, &tos));
    SK_LAUNCH(k_stage_values<double><<<n_ta_, kBlock, 0, stream_>>>(pta, ap, asx, apos, pack, a.a_orig, aos));
    SK_LAUNCH(k_stage_values<double><<<n_tt_, kBlock, 0, stream_>>>(ptt, tp, tsx, tpos, pack, a.t_orig, tos));
    a.a_orig_st = aos;
    a.t_orig_st = tos;
    if (slices_) {
      // the stacked column slices: structure, staging order, and the partial row sums
      const size_t zs = h.s_idx.size();
      int32_t *sp, *si, *ss, *ssx;
      uint16_t* spos;
      Tile* pts;
      T *sv, *xp;
      SK_TRY(put(b_s_ptr_, h.s_ptr, &sp));
      SK_TRY(put(b_s_idx_, h.s_idx, &si));
      SK_TRY(put(b_s_src_, h.s_src, &ss));
      SK_TRY(put(b_ts_, h.ts, &pts));
      SK_TRY(alloc(b_s_sidx_, zs + kStagePad, &ssx));
      SK_TRY(alloc(b_s_pos_, pack_ ? 0 : zs + kStagePad, &spos));
      SK_TRY(alloc(b_s_val_, zs + kStagePad, &sv));
      SK_TRY(alloc(b_xpart_, static_cast<size_t>(slices_) * M, &xp));
      SK_LAUNCH(k_sort_tiles<<<n_ts_, kBlock, 0, stream_>>>(pts, sp, si, sorted, pack, ssx, spos));
      a.ts = pts;
      a.s_ptr = sp;
      a.s_sidx = ssx;
      a.s_pos = spos;
      a.s_val = sv;
      a.xpart = xp;
      s_src_dev_ = ss;
      // the stacked row slices of CSR(A^T), the same way
      const size_t zq = h.q_idx.size();
      int32_t *qp, *qi, *qs, *qsx;
      uint16_t* qpos;
      Tile* ptq;
      T *qv, *yp;
      SK_TRY(put(b_q_ptr_, h.q_ptr, &qp));
      SK_TRY(put(b_q_idx_, h.q_idx, &qi));
      SK_TRY(put(b_q_src_, h.q_src, &qs));
      SK_TRY(put(b_tq_, h.tq, &ptq));
      SK_TRY(alloc(b_q_sidx_, zq + kStagePad, &qsx));
      SK_TRY(alloc(b_q_pos_, pack_ ? 0 : zq + kStagePad, &qpos));
      SK_TRY(alloc(b_q_val_, zq + kStagePad, &qv));
      SK_TRY(alloc(b_ypart_, static_cast<size_t>(kRowSlices) * N, &yp));
      SK_LAUNCH(k_sort_tiles<<<n_tq_, kBlock, 0, stream_>>>(ptq, qp, qi, sorted, pack, qsx, qpos));
      a.tq = ptq;
      a.q_ptr = qp;
      a.q_sidx = qsx;
      a.q_pos = qpos;
      a.q_val = qv;
      a.ypart = yp;
      q_src_dev_ = qs;
    }
    a.n_ts = n_ts_;
    a.slices = slices_;
    a.slice_rows = slices_ ? static_cast<int32_t>(M) : 0;
    a.n_tq = n_tq_;
    a.row_slices = slices_ ? kRowSlices : 0;
    a.slice_cols = slices_ ? static_cast<int32_t>(N) : 0;
  }
  InstDesc* pdesc;
  SK_TRY(put(b_desc_, desc_, &pdesc));
  a.desc = pdesc;
  {
    std::vector<InstDesc>& cd = stage_host_.cdesc;
    cd.resize(static_cast<size_t>(n_tc_));
    for (int32_t k = 0; k < K_; ++k) {
      const InstDesc& d = desc_[static_cast<size_t>(k)];
      for (int32_t c = d.tc0; c < d.tc1; ++c) cd[static_cast<size_t>(c)] = d;
    }
    InstDesc* pcd;
    SK_TRY(put(b_cdesc_, cd, &pcd));
    a.cdesc = pcd;
  }
  // two buffers of instance records and of primal partials: a primal launch that decides the
  // previous step reads one and writes the other (enqueue_block)
  SK_TRY(alloc(b_state_, 2 * static_cast<size_t>(K_), &a.state));
  SK_TRY(alloc(b_ps_c_, 2 * static_cast<size_t>(n_tc_), &a.ps_c));
  a.state_next = a.state;
  a.ps_c_prev = a.ps_c + n_tc_;
  SK_TRY(alloc(b_ps_a_, static_cast<size_t>(n_ta_), &a.ps_a));
  SK_TRY(alloc(b_ps_t_, static_cast<size_t>(n_tt_), &a.ps_t));
  SK_TRY(alloc(b_ps_r_, slices_ ? static_cast<size_t>(n_tr_) : 0, &a.ps_r));
  SK_TRY(alloc(b_ps_q_, slices_ ? static_cast<size_t>(n_tc_) : 0, &a.ps_q));
  SK_TRY(alloc(b_pe_c_, static_cast<size_t>(n_tc_) * 2, &a.pe_c));
  SK_TRY(alloc(b_pe_r_, static_cast<size_t>(n_tr_) * 2, &a.pe_r));
  SK_TRY(alloc(b_pe_a_, static_cast<size_t>(n_ta_) * 8, &a.pe_a));
  SK_TRY(alloc(b_pe_t_, static_cast<size_t>(n_tt_) * 8, &a.pe_t));
  SK_TRY(alloc(b_done_, 1, &a.done_count));
  SK_TRY(put(b_v0_, v0_, &d_v0_));
  warm_any_ = false;
  for (int32_t k = 0; k < K_ && k < static_cast<int32_t>(warm_.size()); ++k) warm_any_ = warm_any_ || warm_[static_cast<size_t>(k)].set;
  if (warm_any_) {
    // in the device order; NaN marks the cold elements
    wx_host_.assign(N, std::nan(""));
    wy_host_.assign(M, std::nan(""));
    for (int32_t k = 0; k < K_; ++k) {
      const WarmStart& w = warm_[static_cast<size_t>(k)];
      if (!w.set) continue;
      const InstDesc& d = desc_[static_cast<size_t>(k)];
      const std::vector<int32_t>& perm = col_perm_[static_cast<size_t>(k)];
      for (int32_t j = 0; j < d.cols; ++j)
        wx_host_[static_cast<size_t>(d.col0 + j)] = w.x[static_cast<size_t>(perm.empty() ? j : perm[static_cast<size_t>(j)])];
      for (int32_t i = 0; i < d.rows; ++i) wy_host_[static_cast<size_t>(d.row0 + i)] = w.y[static_cast<size_t>(i)];
    }
    SK_TRY(put(b_wx_, wx_host_, &d_wx_));
    SK_TRY(put(b_wy_, wy_host_, &d_wy_));
  }
  SK_TRY(alloc(b_v_, N, &d_v_));
  SK_TRY(alloc(b_av_, M, &d_av_));
  SK_TRY(alloc(b_len_, static_cast<size_t>(K_), &d_len_));
  SK_TRY(alloc(b_zero_, static_cast<size_t>(K_), &d_zero_seen_));
  SK_TRY(alloc(b_fa_, static_cast<size_t>(n_ta_), &a.fa));
  SK_TRY(alloc(b_ft_, static_cast<size_t>(n_tt_), &a.ft));
  SK_TRY(alloc(b_fc_, static_cast<size_t>(n_tc_), &a.fc));
  SK_TRY(alloc(b_fr_, static_cast<size_t>(n_tr_), &a.fr));
  {
    const double* rule = nullptr;
    const int st_rule = rule_table_device(&rule);
    if (st_rule) return st_rule;
    a.rule = rule;
    a.rule_len = static_cast<int32_t>(kRuleIterations);
  }
  a.n_tc = n_tc_;
  a.n_ta = n_ta_;
  a.n_tt = n_tt_;
  a.force = 0;
  a.report = 0;
  a.decide = 0;
#undef SK_TRY
  // the host vectors above are sources of asynchronous copies (the drain also covers errors)
  SK_HIP(hipStreamSynchronize(stream_), kStatusSyncFailed);
  final_state_.assign(static_cast<size_t>(K_), InstState{});
  extracted_.assign(static_cast<size_t>(K_), 0);
  bestx_.assign(static_cast<size_t>(n_), 0.0);
  besty_.assign(static_cast<size_t>(m_), 0.0);
  act_.assign(static_cast<size_t>(m_), 0.0);
  red_.assign(static_cast<size_t>(n_), 0.0);
  return kStatusSuccess;
}

// Device set-up of one solve, enqueued on the engine's stream (captured once as a graph):
// scaling (scaling.cpp:111-159), scaled vectors, the power iteration (scaling.cpp:161-194),
// the initial step sizes and iterates (pdhg.cpp:327-379).
template <class T>
int EngineImpl<T>::enqueue_setup() {
  Args<T>& a = args_;
  hipStream_t s = stream_;
  const auto M = static_cast<int32_t>(m_), N = static_cast<int32_t>(n_);
  const auto Z = static_cast<int64_t>(nnz_);
  SK_LAUNCH(k_copy<<<blocks_for(Z), kBlock, 0, s>>>(a.a_orig, d_sa_, Z));
  SK_LAUNCH(k_copy<<<blocks_for(Z), kBlock, 0, s>>>(a.t_orig, d_st_, Z));
  SK_LAUNCH(k_fill<<<blocks_for(M), kBlock, 0, s>>>(const_cast<double*>(a.dr), 1.0, M));
  SK_LAUNCH(k_fill<<<blocks_for(N), kBlock, 0, s>>>(const_cast<double*>(a.dc), 1.0, N));
  // 10 Ruiz rounds, then one Pock-Chambolle round (scaling.cpp:121-122)
  for (int pass = 0; pass < 11; ++pass) {
    const int use_sum = pass == 10;
    SK_LAUNCH(k_row_pass<<<n_ta_, kBlock, 0, s>>>(a.ta, a.a_ptr, d_sa_, use_sum, d_rpass_));
    SK_LAUNCH(k_row_pass<<<n_tt_, kBlock, 0, s>>>(a.tt, a.t_ptr, d_st_, use_sum, d_cpass_));
    SK_LAUNCH(k_apply_pass<<<n_ta_, kBlock, 0, s>>>(a.ta, a.a_ptr, a.a_idx, d_sa_, 1, d_rpass_, d_cpass_));
    SK_LAUNCH(k_apply_pass<<<n_tt_, kBlock, 0, s>>>(a.tt, a.t_ptr, a.t_idx, d_st_, 0, d_rpass_, d_cpass_));
    SK_LAUNCH(k_scale_mul<<<blocks_for(M), kBlock, 0, s>>>(const_cast<double*>(a.dr), d_rpass_, M));
    SK_LAUNCH(k_scale_mul<<<blocks_for(N), kBlock, 0, s>>>(const_cast<double*>(a.dc), d_cpass_, N));
  }
  const int32_t MN = std::max(M, N);
  SK_LAUNCH(k_scaled_vectors<T><<<blocks_for(MN), kBlock, 0, s>>>(
      a, const_cast<T*>(a.chat), const_cast<T*>(a.lhat), const_cast<T*>(a.uhat),
      const_cast<T*>(a.rlhat), const_cast<T*>(a.ruhat), N, M));
  {
    // FP64 scaled column bounds for the reference's initial `best` point
    Args<double> a64{};
    a64.dc = a.dc;
    a64.dr = a.dr;
    a64.cmin = a.cmin;
    a64.cl = a.cl;
    a64.cu = a.cu;
    a64.rl = a.rl;
    a64.ru = a.ru;
    SK_LAUNCH(k_scaled_vectors<double><<<blocks_for(N), kBlock, 0, s>>>(
        a64, a.xu0, d_lhat64_, d_uhat64_, a.yu0, a.yu1, N, 0));
  }
  const int32_t pack = pack_ ? 1 : 0;
  SK_LAUNCH(k_stage_values<T><<<n_ta_, kBlock, 0, s>>>(a.ta, a.a_ptr, a.a_sidx, a.a_pos, pack, d_sa_,
                                                       const_cast<T*>(a.a_val)));
  SK_LAUNCH(k_stage_values<T><<<n_tt_, kBlock, 0, s>>>(a.tt, a.t_ptr, a.t_sidx, a.t_pos, pack, d_st_,
                                                       const_cast<T*>(a.t_val)));
  if (slices_) {
    SK_LAUNCH(k_stage_slice_values<T><<<n_ts_, kBlock, 0, s>>>(a.ts, a.s_ptr, a.s_sidx, a.s_pos, pack,
                                                               s_src_dev_, d_sa_, const_cast<T*>(a.s_val)));
    SK_LAUNCH(k_stage_slice_values<T><<<n_tq_, kBlock, 0, s>>>(a.tq, a.q_ptr, a.q_sidx, a.q_pos, pack,
                                                               q_src_dev_, d_st_, const_cast<T*>(a.q_val)));
  }
  // power iteration on A^T A, 30 steps (scaling.cpp:161-194)
  SK_LAUNCH(k_copy<<<blocks_for(N), kBlock, 0, s>>>(d_v0_, d_v_, N));
  SK_LAUNCH(k_fill_i32<<<blocks_for(K_), kBlock, 0, s>>>(d_zero_seen_, 0, K_));
  for (int step = 0; step < 30; ++step) {
    SK_LAUNCH(k_chunk_sumsq<<<n_tc_, kBlock, 0, s>>>(a.tc, d_v_, a.ps_c));
    SK_LAUNCH(k_inst_norm<<<K_, kBlock, 0, s>>>(a.desc, a.ps_c, d_len_, d_zero_seen_));
    SK_LAUNCH(k_chunk_div<<<n_tc_, kBlock, 0, s>>>(a.tc, d_len_, d_v_));
    if (pack_) {
      SK_LAUNCH(k_spmv_tiles<T, true><<<n_ta_, kBlock, 0, s>>>(a.ta, a.a_ptr, a.a_sidx, a.a_pos, a.a_val, d_v_, d_av_));
      SK_LAUNCH(k_spmv_tiles<T, true><<<n_tt_, kBlock, 0, s>>>(a.tt, a.t_ptr, a.t_sidx, a.t_pos, a.t_val, d_av_, d_v_));
    } else {
      SK_LAUNCH(k_spmv_tiles<T, false><<<n_ta_, kBlock, 0, s>>>(a.ta, a.a_ptr, a.a_sidx, a.a_pos, a.a_val, d_v_, d_av_));
      SK_LAUNCH(k_spmv_tiles<T, false><<<n_tt_, kBlock, 0, s>>>(a.tt, a.t_ptr, a.t_sidx, a.t_pos, a.t_val, d_av_, d_v_));
    }
  }
  SK_LAUNCH(k_chunk_sumsq<<<n_tc_, kBlock, 0, s>>>(a.tc, d_v_, a.ps_c));
  SK_LAUNCH(k_inst_norm<<<K_, kBlock, 0, s>>>(a.desc, a.ps_c, d_len_, nullptr));
  SK_LAUNCH(k_power_finish<<<blocks_for(K_), kBlock, 0, s>>>(d_len_, d_zero_seen_, a.state, K_));
  SK_LAUNCH(k_init_iterates<T><<<blocks_for(MN), kBlock, 0, s>>>(a, d_lhat64_, d_uhat64_, N, M));
  if (warm_any_) {
    // warm iterates, then the carried A^T y of the warm y with the step kernels' product
    SK_LAUNCH(k_warm_iterates<T><<<blocks_for(MN), kBlock, 0, s>>>(a, d_wx_, d_wy_, d_lhat64_, d_uhat64_, N, M));
    const int st = with_layout([&](auto cap, auto pack) -> int {
      constexpr int C = decltype(cap)::value;
      constexpr bool P = decltype(pack)::value;
      if (slices_) {
        // the product the steps form: the row slices' partials, added in slice order
        SK_LAUNCH(k_probe_product<T, C, P><<<static_cast<unsigned>(n_tq_), kStepLanes, 0, s>>>(
            a.tq, a.q_ptr, a.q_sidx, a.q_pos, a.q_val, a.y0, a.ypart));
        SK_LAUNCH(k_slice_sum<T><<<blocks_for(N), kBlock, 0, s>>>(a.ypart, a.aty0, a.slice_cols, a.row_slices));
      } else {
        SK_LAUNCH(k_probe_product<T, C, P><<<static_cast<unsigned>(n_tt_), kStepLanes, 0, s>>>(
            a.tt, a.t_ptr, a.t_sidx, a.t_pos, a.t_val, a.y0, a.aty0));
      }
      return kStatusSuccess;
    });
    if (st) return st;
    SK_LAUNCH(k_copy_t<T><<<blocks_for(N), kBlock, 0, s>>>(a.aty0, a.aty1, N));
  }
  if (K_ > 0) SK_LAUNCH(k_publish_flags<T><<<K_, kBlock, 0, s>>>(a));
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::capture(int (EngineImpl<T>::*body)(), hipGraphExec_t* exec) {
  hipStream_t s = stream_;
  SK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), kStatusLaunchFailed);
  const int st = (this->*body)();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &graph);
  if (st) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) {
    set_last_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return kStatusLaunchFailed;
  }
  const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    set_last_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    *exec = nullptr;
    return kStatusLaunchFailed;
  }
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::setup_device() {
  hipStream_t s = stream_;
  // per-solve state: host-built records (norms, options, instances stopped before step 1)
  host_done_[2] = initially_done_;
  SK_HIP(hipMemcpyAsync(args_.state, init_state_.data(), init_state_.size() * sizeof(InstState),
                        hipMemcpyHostToDevice, s), kStatusCopyFailed);
  SK_HIP(hipMemcpyAsync(args_.done_count, &host_done_[2], sizeof(uint32_t), hipMemcpyHostToDevice, s),
         kStatusCopyFailed);
  if (opt_.use_graph) {
    if (!setup_exec_) {
      const int st = capture(&EngineImpl<T>::enqueue_setup, &setup_exec_);
      if (st) return st;
    }
    SK_HIP(hipGraphLaunch(setup_exec_, s), kStatusLaunchFailed);
  } else {
    const int st = enqueue_setup();
    if (st) return st;
  }
  setup_done_ = true;
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::enqueue_eval(int report_mode) {
  Args<T> a = args_;
  a.report = report_mode;
  hipStream_t s = stream_;
  if (!report_mode) SK_LAUNCH(k_eval_prepare<T><<<n_tc_ + n_tr_, kBlock, 0, s>>>(a));
  const int st = with_layout([&](auto cap, auto pack) -> int {
    constexpr int C = decltype(cap)::value;
    constexpr bool P = decltype(pack)::value;
    SK_LAUNCH(k_eval_products<T, C, P><<<n_ta_ + n_tt_, kBlock, 0, s>>>(a, n_ta_));
    return kStatusSuccess;
  });
  if (st) return st;
  SK_LAUNCH(k_eval_finalize<T><<<K_, kBlock, 0, s>>>(a));
  if (!report_mode) SK_LAUNCH(k_eval_commit<T><<<n_tc_ + n_tr_, kBlock, 0, s>>>(a));
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::enqueue_block() {
  hipStream_t s = stream_;
  const int B = opt_.steps_per_block;
  const auto eval_after = [&](int k) { return (k + 1) % opt_.eval_stride == 0 || k + 1 == B; };
  if (fused_) {
    // one launch: each instance's steps, evaluations and final evaluation (k_resident)
    const int32_t rounds = kResidentRounds * B;
    return with_layout([&](auto cap, auto pack) -> int {
      constexpr int C = decltype(cap)::value;
      constexpr bool P = decltype(pack)::value;
      SK_LAUNCH(k_resident<T, C, P><<<K_, kBlock, 0, s>>>(args_, rounds));
      return kStatusSuccess;
    });
  }
  // alone[k]: step k is decided by its own launch (k_step_finalize), which writes the records
  // into buffer 0; that happens before every evaluation stage and at the end of the block.
  // Every other step is decided by the next primal launch, which reads the records from one
  // buffer and writes them to the other. The number of such switches is made even, so the
  // block ends, and every replay of it starts, with the records in buffer 0. Step k's primal
  // partials are in buffer k % 2, so the launch deciding step k never overwrites them.
  std::vector<uint8_t> alone(static_cast<size_t>(B), 1);
  int switches = 0;
  if (fold_) {
    for (int k = 0; k < B; ++k) alone[static_cast<size_t>(k)] = eval_after(k) ? 1 : 0;
    for (int k = 0; k + 1 < B; ++k) switches += alone[static_cast<size_t>(k)] ? 0 : 1;
    for (int k = B - 2; k >= 0 && switches % 2; --k) {
      if (!alone[static_cast<size_t>(k)]) {
        alone[static_cast<size_t>(k)] = 1;
        --switches;
      }
    }
  }
  InstState* const state[2] = {args_.state, args_.state + K_};
  double* const ps_c[2] = {args_.ps_c, args_.ps_c + n_tc_};
  int cur = 0;  // the buffer that holds the records
  for (int k = 0; k < B; ++k) {
    Args<T> a = args_;
    a.state = state[cur];
    a.ps_c = ps_c[k & 1];
    if (k > 0 && !alone[static_cast<size_t>(k - 1)]) {
      a.decide = 1;
      a.state_next = state[1 - cur];
      a.ps_c_prev = ps_c[(k - 1) & 1];
      cur = 1 - cur;
    }
    SK_LAUNCH(k_step_primal<T><<<n_tc_, kBlock, 0, s>>>(a));
    a.decide = 0;
    a.state = state[cur];
    if (const int st = launch_spmv(1, a)) return st;
    if (const int st = launch_spmv(2, a)) return st;
    if (alone[static_cast<size_t>(k)]) {
      a.state_next = state[0];
      SK_LAUNCH(k_step_finalize<T><<<K_, kBlock, 0, s>>>(a));
      cur = 0;
      // an instance that became due for evaluation waits at most eval_stride steps
      if (eval_after(k)) {
        if (const int e = enqueue_eval(0)) return e;
      }
    }
  }
  // the final evaluation of every instance that stopped in this block, so its answer can be
  // read back as soon as the block completes
  return enqueue_eval(1);
}

template <class T>
int EngineImpl<T>::run_loop() {
  hipStream_t s = stream_;
  int st = 0;
  if (opt_.use_graph && !graph_exec_) {
    if ((st = capture(&EngineImpl<T>::enqueue_block, &graph_exec_))) return st;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t target = static_cast<uint32_t>(K_);
  int64_t launched = 0;
  hipEvent_t ev[2];
  SK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming), kStatusInitFailed);
  SK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming), kStatusInitFailed);
  bool finished = host_done_[2] >= target;
  bool timed_out = false;
  // Block k copies the finished-instance counter into ring slot k % 2 and records event k % 2.
  // Block k + 1 is launched before block k is waited on, so the device never idles on the
  // host's read-back.
  auto launch = [&]() -> int {
    if (graph_exec_) {
      SK_HIP(hipGraphLaunch(graph_exec_, s), kStatusLaunchFailed);
    } else {
      const int r = enqueue_block();
      if (r) return r;
    }
    SK_HIP(hipMemcpyAsync(&host_done_[launched & 1], args_.done_count, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s), kStatusCopyFailed);
    SK_HIP(hipEventRecord(ev[launched & 1], s), kStatusLaunchFailed);
    ++launched;
    return kStatusSuccess;
  };
  if (!finished) st = launch();
  for (int64_t k = 0; !finished && !st; ++k) {
    if ((st = launch())) break;  // block k + 1
    const hipError_t e = hipEventSynchronize(ev[k & 1]);
    if (e != hipSuccess) {
      set_last_error(std::string("hipEventSynchronize: ") + hipGetErrorString(e));
      st = kStatusSyncFailed;
      break;
    }
    finished = host_done_[k & 1] >= target;
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!finished && elapsed > opt_.time_limit) {
      timed_out = true;
      break;
    }
  }
  if (!st) {
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_last_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
      st = kStatusSyncFailed;
    }
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  timings_.graph_launches = launched;
  if (st) return st;
  if (timed_out) {
    // Mark every unfinished instance as stopped by the time limit (host side, then upload).
    std::vector<InstState> cur(static_cast<size_t>(K_));
    SK_HIP(hipMemcpy(cur.data(), args_.state, cur.size() * sizeof(InstState), hipMemcpyDeviceToHost), kStatusCopyFailed);
    for (auto& x : cur) {
      if (!x.done) {
        x.done = 1;
        x.status = kStTimeLimit;
        x.report_pending = 1;
        x.eval_pending = 0;
      }
    }
    SK_HIP(hipMemcpy(args_.state, cur.data(), cur.size() * sizeof(InstState), hipMemcpyHostToDevice), kStatusCopyFailed);
    SK_LAUNCH(k_publish_flags<T><<<K_, kBlock, 0, s>>>(args_));
  }
  return kStatusSuccess;
}

// ---- serving: non-blocking solve ------------------------------------------------------------

template <class T>
int EngineImpl<T>::launch_block() {
  // ring slot of this block; poll() has consumed the block that used it before
  const size_t slot = static_cast<size_t>(launched_ & 1);
  hipStream_t s = stream_;
  if (graph_exec_) {
    SK_HIP(hipGraphLaunch(graph_exec_, s), kStatusLaunchFailed);
  } else {
    const int st = enqueue_block();
    if (st) return st;
  }
  SK_HIP(hipMemcpyAsync(&host_done_[slot], args_.done_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
         kStatusCopyFailed);
  SK_HIP(hipMemcpyAsync(host_state_ + slot * static_cast<size_t>(K_), args_.state,
                        static_cast<size_t>(K_) * sizeof(InstState), hipMemcpyDeviceToHost, s),
         kStatusCopyFailed);
  SK_HIP(hipEventRecord(block_ev_[slot], s), kStatusLaunchFailed);
  ++launched_;
  ++inflight_;
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::start() {
  if (!stream_) {
    set_last_error("start() before upload()");
    return kStatusInvalidState;
  }
  if (inflight_ > 0) {
    set_last_error("start() while the previous solve is still running");
    return kStatusInvalidState;
  }
  // the read-back stream of the serving path (created on first use: solve() does not need it,
  // and every stream of a process takes one of its few hardware queues)
  if (!copy_stream_) SK_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking), kStatusInitFailed);
  started_ = std::chrono::steady_clock::now();
  int st = setup_device();
  if (st) return st;
  if (opt_.use_graph && !graph_exec_) {
    if ((st = capture(&EngineImpl<T>::enqueue_block, &graph_exec_))) return st;
  }
  launched_ = 0;
  inflight_ = 0;
  extracted_count_ = 0;
  std::fill(extracted_.begin(), extracted_.end(), uint8_t{0});
  timings_ = EngineTimings{};
  // two blocks in flight, so the device never waits for the host's read-back; one is enough
  // when every instance stopped before its first step (it runs their final evaluation)
  if ((st = launch_block())) return st;
  if (initially_done_ < static_cast<uint32_t>(K_)) st = launch_block();
  return st;
}

// Reads back the answers of the instances that a block's snapshot shows as stopped and that
// were not read yet. Their report is complete (the block ends with the final evaluation of
// every instance that stopped in it) and no later kernel writes their slices, so the copies
// run on the second stream while the next block computes.
template <class T>
int EngineImpl<T>::extract_finished(const InstState* snapshot, std::vector<int32_t>* newly_done) {
  hipStream_t c = copy_stream_;
  std::vector<int32_t> fresh;
  for (int32_t k = 0; k < K_; ++k) {
    if (snapshot[k].done && !extracted_[static_cast<size_t>(k)] && !snapshot[k].report_pending)
      fresh.push_back(k);
  }
  for (const int32_t k : fresh) {
    final_state_[static_cast<size_t>(k)] = snapshot[k];
    const InstDesc& d = desc_[static_cast<size_t>(k)];
    const auto c0 = static_cast<size_t>(d.col0), r0 = static_cast<size_t>(d.row0);
    const auto nc = static_cast<size_t>(d.cols), nr = static_cast<size_t>(d.rows);
    if (nc) {
      SK_HIP(hipMemcpyAsync(bestx_.data() + c0, args_.bestx + c0, nc * sizeof(double), hipMemcpyDeviceToHost, c), kStatusCopyFailed);
      SK_HIP(hipMemcpyAsync(red_.data() + c0, args_.reduced + c0, nc * sizeof(double), hipMemcpyDeviceToHost, c), kStatusCopyFailed);
    }
    if (nr) {
      SK_HIP(hipMemcpyAsync(besty_.data() + r0, args_.besty + r0, nr * sizeof(double), hipMemcpyDeviceToHost, c), kStatusCopyFailed);
      SK_HIP(hipMemcpyAsync(act_.data() + r0, args_.activity + r0, nr * sizeof(double), hipMemcpyDeviceToHost, c), kStatusCopyFailed);
    }
  }
  if (!fresh.empty()) SK_HIP(hipStreamSynchronize(c), kStatusSyncFailed);
  for (const int32_t k : fresh) {
    extracted_[static_cast<size_t>(k)] = 1;
    if (newly_done) newly_done->push_back(k);
  }
  extracted_count_ += static_cast<uint32_t>(fresh.size());
  return kStatusSuccess;
}

// Time limit of a non-blocking solve: drain the device, stop every running instance as the
// blocking loop does, run their final evaluation and read everything back.
template <class T>
int EngineImpl<T>::stop_on_time_limit(std::vector<int32_t>* newly_done) {
  hipStream_t s = stream_;
  SK_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  inflight_ = 0;
  std::vector<InstState> cur(static_cast<size_t>(K_));
  SK_HIP(hipMemcpy(cur.data(), args_.state, cur.size() * sizeof(InstState), hipMemcpyDeviceToHost), kStatusCopyFailed);
  for (auto& x : cur) {
    if (!x.done) {
      x.done = 1;
      x.status = kStTimeLimit;
      x.report_pending = 1;
      x.eval_pending = 0;
    }
  }
  SK_HIP(hipMemcpy(args_.state, cur.data(), cur.size() * sizeof(InstState), hipMemcpyHostToDevice), kStatusCopyFailed);
  SK_LAUNCH(k_publish_flags<T><<<K_, kBlock, 0, s>>>(args_));
  int st = enqueue_eval(1);
  if (st) return st;
  SK_HIP(hipMemcpyAsync(host_state_, args_.state, cur.size() * sizeof(InstState), hipMemcpyDeviceToHost, s),
         kStatusCopyFailed);
  SK_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  return extract_finished(host_state_, newly_done);
}

template <class T>
int EngineImpl<T>::poll(bool wait, bool* finished, std::vector<int32_t>* newly_done) {
  *finished = false;
  if (inflight_ == 0) {
    *finished = extracted_count_ >= static_cast<uint32_t>(K_);
    return kStatusSuccess;
  }
  const size_t slot = static_cast<size_t>((launched_ - inflight_) & 1);  // the oldest block
  const hipError_t q = wait ? hipEventSynchronize(block_ev_[slot]) : hipEventQuery(block_ev_[slot]);
  if (q == hipErrorNotReady) return kStatusSuccess;
  if (q != hipSuccess) {
    set_last_error(std::string("hipEventQuery: ") + hipGetErrorString(q));
    return kStatusSyncFailed;
  }
  --inflight_;
  int st = kStatusSuccess;
  if (host_done_[slot] > extracted_count_) {
    st = extract_finished(host_state_ + slot * static_cast<size_t>(K_), newly_done);
    if (st) return st;
  }
  timings_.graph_launches = launched_;
  if (extracted_count_ >= static_cast<uint32_t>(K_)) {
    *finished = inflight_ == 0;  // else the block still in flight is drained by the next poll
    return kStatusSuccess;
  }
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - started_).count();
  if (elapsed > opt_.time_limit) {
    st = stop_on_time_limit(newly_done);
    *finished = !st;
    return st;
  }
  return launch_block();
}

template <class T>
int EngineImpl<T>::report() {
  int st = enqueue_eval(1);
  if (st) return st;
  hipStream_t s = stream_;
  final_state_.resize(static_cast<size_t>(K_));
  bestx_.resize(static_cast<size_t>(n_));
  besty_.resize(static_cast<size_t>(m_));
  act_.resize(static_cast<size_t>(m_));
  red_.resize(static_cast<size_t>(n_));
  SK_HIP(hipMemcpyAsync(final_state_.data(), args_.state, final_state_.size() * sizeof(InstState), hipMemcpyDeviceToHost, s), kStatusCopyFailed);
  SK_HIP(hipMemcpyAsync(bestx_.data(), args_.bestx, bestx_.size() * sizeof(double), hipMemcpyDeviceToHost, s), kStatusCopyFailed);
  SK_HIP(hipMemcpyAsync(besty_.data(), args_.besty, besty_.size() * sizeof(double), hipMemcpyDeviceToHost, s), kStatusCopyFailed);
  SK_HIP(hipMemcpyAsync(act_.data(), args_.activity, act_.size() * sizeof(double), hipMemcpyDeviceToHost, s), kStatusCopyFailed);
  SK_HIP(hipMemcpyAsync(red_.data(), args_.reduced, red_.size() * sizeof(double), hipMemcpyDeviceToHost, s), kStatusCopyFailed);
  SK_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::solve() {
  if (!stream_) {
    set_last_error("solve() before upload()");
    return kStatusInvalidState;
  }
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  int st = setup_device();
  if (st) return st;
  const auto t1 = clock::now();
  st = run_loop();
  if (st) return st;
  const auto t2 = clock::now();
  st = report();
  if (st) return st;
  const auto t3 = clock::now();
  timings_.setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  timings_.loop_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  timings_.report_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::result(int32_t index, InstanceResult* out) const {
  if (index < 0 || index >= K_ || final_state_.empty()) {
    set_last_error("result(): no such solved instance");
    return kStatusInvalidArgument;
  }
  const InstState& s = final_state_[static_cast<size_t>(index)];
  const InstDesc& d = desc_[static_cast<size_t>(index)];
  const double sense = static_cast<double>(sense_[static_cast<size_t>(index)]);
  const HostLp* lp = lps_[static_cast<size_t>(index)];
  out->status = s.status;
  out->converged = s.converged;
  out->iterations = s.iteration;
  out->restarts = s.restarts;
  out->evaluations = s.evaluations;
  out->attempts = s.attempts;
  out->spectral_norm = s.spectral_norm;
  out->omega = s.omega;
  out->eta = s.eta;
  out->final_residuals = s.report;
  // pdhg.cpp:647-716: kOptimal only for a converged point meeting the project standard
  const Residuals& f = s.report;
  const bool standard = f.absolute_primal <= 1e-7 && f.absolute_dual <= 1e-7 &&
                        f.gap_as_verified <= 1e-9 && f.complementarity <= 1e-6;
  const bool verifiable = s.converged && standard;
  if (s.converged) out->status = verifiable ? kStOptimal : kStFeasible;
  out->dual_bound = verifiable ? sense * f.dual_objective + (lp ? lp->objective_offset : 0.0)
                               : (sense < 0 ? INFINITY : -INFINITY);
  const auto c0 = static_cast<size_t>(d.col0), r0 = static_cast<size_t>(d.row0);
  // back to the LP's column order
  const std::vector<int32_t>& perm = col_perm_[static_cast<size_t>(index)];
  out->col_value.resize(static_cast<size_t>(d.cols));
  out->col_dual.resize(static_cast<size_t>(d.cols));
  for (int32_t j = 0; j < d.cols; ++j) {
    const auto o = static_cast<size_t>(perm.empty() ? j : perm[static_cast<size_t>(j)]);
    out->col_value[o] = bestx_[c0 + static_cast<size_t>(j)];
    out->col_dual[o] = sense * red_[c0 + static_cast<size_t>(j)];
  }
  out->row_dual.resize(static_cast<size_t>(d.rows));
  for (int32_t i = 0; i < d.rows; ++i) out->row_dual[static_cast<size_t>(i)] = sense * (-besty_[r0 + static_cast<size_t>(i)]);
  out->row_activity.assign(act_.begin() + static_cast<std::ptrdiff_t>(r0), act_.begin() + static_cast<std::ptrdiff_t>(r0 + static_cast<size_t>(d.rows)));
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::time_kernel(int kernel, int reps, double* mean_ms) {
  if (!stream_ || reps <= 0) {
    set_last_error("time_kernel(): engine not uploaded or reps <= 0");
    return kStatusInvalidArgument;
  }
  Args<T> a = args_;
  // kernels 3..5 are 0..2 in normal mode: instances that stopped skip their tiles
  a.force = kernel < 3 ? 1 : 0;
  kernel %= 3;
  hipStream_t s = stream_;
  hipEvent_t e0, e1;
  SK_HIP(hipEventCreate(&e0), kStatusInitFailed);
  SK_HIP(hipEventCreate(&e1), kStatusInitFailed);
  auto one = [&]() -> int {
    if (kernel == 0) SK_LAUNCH(k_step_primal<T><<<n_tc_, kBlock, 0, s>>>(a));
    else return launch_spmv(kernel, a);
    return kStatusSuccess;
  };
  int st = one();  // warm
  if (st) return st;
  SK_HIP(hipEventRecord(e0, s), kStatusLaunchFailed);
  for (int r = 0; r < reps; ++r)
    if ((st = one())) return st;
  SK_HIP(hipEventRecord(e1, s), kStatusLaunchFailed);
  SK_HIP(hipEventSynchronize(e1), kStatusSyncFailed);
  float ms = 0.0f;
  SK_HIP(hipEventElapsedTime(&ms, e0, e1), kStatusSyncFailed);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *mean_ms = static_cast<double>(ms) / reps;
  return kStatusSuccess;
}

// Algorithmic bytes per launch (each array element counted once):
//   0 primal: per column read x, c, l, u, A^T y (5 words), write x', xbar (2 words),
//             read-modify-write the running sum (2 words)
//   1 dual:   CSR(A): (w + 4) B per nonzero + 4 (m + 1) row pointers + w n for xbar;
//             per row read y, rl, ru (3 words), write y' (1 word), sum RMW (2 words)
//   2 A^T:    CSR(A^T): (w + 4) B per nonzero + 4 (n + 1) + w m for y';
//             per column read x, x', A^T y (3 words), write A^T y' (1 word)
template <class T>
double EngineImpl<T>::kernel_bytes(int kernel) const {
  kernel %= 3;
  const double w = sizeof(T), z = static_cast<double>(nnz_);
  const double m = static_cast<double>(m_), n = static_cast<double>(n_);
  if (kernel == 0) return n * 9.0 * w;
  if (kernel == 1) return (w + 4.0) * z + 4.0 * (m + 1.0) + w * n + m * 6.0 * w;
  return (w + 4.0) * z + 4.0 * (n + 1.0) + w * m + n * 4.0 * w;
}

template <class T>
int64_t EngineImpl<T>::device_bytes() const {
  int64_t total = 0;
  for (const DevBuf* b : buffers()) total += static_cast<int64_t>(b->bytes);
  return total;
}

template <class T>
void EngineImpl<T>::release() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (setup_exec_) (void)hipGraphExecDestroy(setup_exec_);
  graph_exec_ = setup_exec_ = nullptr;
  for (DevBuf* b : buffers()) {
    if (b->ptr) (void)hipFree(b->ptr);
    b->ptr = nullptr;
    b->bytes = 0;
  }
  args_ = Args<T>{};
  stage_host_ = HostStage{};
  inflight_ = 0;
  setup_done_ = false;
  K_ = 0;
}

template <class T>
int EngineImpl<T>::column_order(int32_t index, int32_t* device_to_lp) const {
  if (index < 0 || index >= K_ || !device_to_lp) {
    set_last_error("column_order(): no such instance");
    return kStatusInvalidArgument;
  }
  const std::vector<int32_t>& perm = col_perm_[static_cast<size_t>(index)];
  const int32_t n = desc_[static_cast<size_t>(index)].cols;
  for (int32_t j = 0; j < n; ++j) device_to_lp[j] = perm.empty() ? j : perm[static_cast<size_t>(j)];
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::probe_spmv(int32_t index, int which, const double* in, double* out) {
  if (index < 0 || index >= K_ || (which != 1 && which != 2) || !in || !out) {
    set_last_error("probe_spmv(): bad instance, product or array");
    return kStatusInvalidArgument;
  }
  if (!setup_done_) {
    set_last_error("probe_spmv(): the device set-up has not run (solve first)");
    return kStatusInvalidState;
  }
  SK_HIP(hipStreamSynchronize(stream_), kStatusSyncFailed);
  const InstDesc& d = desc_[static_cast<size_t>(index)];
  const std::vector<int32_t>& perm = col_perm_[static_cast<size_t>(index)];
  auto lp_col = [&](int32_t j) { return static_cast<size_t>(perm.empty() ? j : perm[static_cast<size_t>(j)]); };
  const size_t nin = static_cast<size_t>(which == 1 ? n_ : m_);
  const size_t nout = static_cast<size_t>(which == 1 ? m_ : n_);
  // the other instances' inputs are zero; only this instance's outputs are read back
  std::vector<T> hin(nin, static_cast<T>(0)), hout(nout);
  if (which == 1) {
    for (int32_t j = 0; j < d.cols; ++j) hin[static_cast<size_t>(d.col0 + j)] = static_cast<T>(in[lp_col(j)]);
  } else {
    for (int32_t i = 0; i < d.rows; ++i) hin[static_cast<size_t>(d.row0 + i)] = static_cast<T>(in[i]);
  }
  DevBuf bin, bout;
  T *din = nullptr, *dout = nullptr;
  int st = alloc(bin, nin, &din);
  if (!st) st = alloc(bout, nout, &dout);
  if (st) return st;
  SK_HIP(hipMemcpy(din, hin.data(), nin * sizeof(T), hipMemcpyHostToDevice), kStatusCopyFailed);
  const Args<T>& a = args_;
  hipStream_t s = stream_;
  st = with_layout([&](auto cap, auto pack) -> int {
    constexpr int C = decltype(cap)::value;
    constexpr bool P = decltype(pack)::value;
    if (which == 1 && slices_) {
      // the step kernels' sliced product: each slice's row sums, then their sum per row (the
      // solve's partial buffers serve as scratch: every step rewrites them before reading)
      SK_LAUNCH(k_probe_product<T, C, P><<<static_cast<unsigned>(n_ts_), kStepLanes, 0, s>>>(
          a.ts, a.s_ptr, a.s_sidx, a.s_pos, a.s_val, din, a.xpart));
      SK_LAUNCH(k_slice_sum<T><<<blocks_for(m_), kBlock, 0, s>>>(a.xpart, dout, a.slice_rows, a.slices));
    } else if (which == 1) {
      SK_LAUNCH(k_probe_product<T, C, P><<<static_cast<unsigned>(n_ta_), kStepLanes, 0, s>>>(
          a.ta, a.a_ptr, a.a_sidx, a.a_pos, a.a_val, din, dout));
    } else if (slices_) {
      SK_LAUNCH(k_probe_product<T, C, P><<<static_cast<unsigned>(n_tq_), kStepLanes, 0, s>>>(
          a.tq, a.q_ptr, a.q_sidx, a.q_pos, a.q_val, din, a.ypart));
      SK_LAUNCH(k_slice_sum<T><<<blocks_for(n_), kBlock, 0, s>>>(a.ypart, dout, a.slice_cols, a.row_slices));
    } else
      SK_LAUNCH(k_probe_product<T, C, P><<<static_cast<unsigned>(n_tt_), kStepLanes, 0, s>>>(
          a.tt, a.t_ptr, a.t_sidx, a.t_pos, a.t_val, din, dout));
    return kStatusSuccess;
  });
  if (st) return st;
  SK_HIP(hipStreamSynchronize(s), kStatusSyncFailed);
  SK_HIP(hipMemcpy(hout.data(), dout, nout * sizeof(T), hipMemcpyDeviceToHost), kStatusCopyFailed);
  if (which == 1) {
    for (int32_t i = 0; i < d.rows; ++i) out[i] = static_cast<double>(hout[static_cast<size_t>(d.row0 + i)]);
  } else {
    for (int32_t j = 0; j < d.cols; ++j) out[lp_col(j)] = static_cast<double>(hout[static_cast<size_t>(d.col0 + j)]);
  }
  return kStatusSuccess;
}

template <class T>
int EngineImpl<T>::probe_scaling(int32_t index, double* row_scale, double* col_scale,
                                 double* values) {
  if (index < 0 || index >= K_) {
    set_last_error("probe_scaling(): no such instance");
    return kStatusInvalidArgument;
  }
  if (!setup_done_) {
    set_last_error("probe_scaling(): the device set-up has not run (solve first)");
    return kStatusInvalidState;
  }
  SK_HIP(hipStreamSynchronize(stream_), kStatusSyncFailed);
  const InstDesc& d = desc_[static_cast<size_t>(index)];
  const std::vector<int32_t>& perm = col_perm_[static_cast<size_t>(index)];
  const auto rows = static_cast<size_t>(d.rows), cols = static_cast<size_t>(d.cols);
  if (row_scale && rows)
    SK_HIP(hipMemcpy(row_scale, args_.dr + d.row0, rows * sizeof(double), hipMemcpyDeviceToHost), kStatusCopyFailed);
  if (col_scale && cols) {
    std::vector<double> dc(cols);
    SK_HIP(hipMemcpy(dc.data(), args_.dc + d.col0, cols * sizeof(double), hipMemcpyDeviceToHost), kStatusCopyFailed);
    for (size_t j = 0; j < cols; ++j) col_scale[perm.empty() ? j : static_cast<size_t>(perm[j])] = dc[j];
  }
  if (values && cols) {
    // CSR(A^T) holds the CSC of A in the device column order: each device column's entries
    // (rows ascending) go to the LP column it holds
    std::vector<int32_t> tp(cols + 1);
    SK_HIP(hipMemcpy(tp.data(), args_.t_ptr + d.col0, (cols + 1) * sizeof(int32_t), hipMemcpyDeviceToHost), kStatusCopyFailed);
    const auto z = static_cast<size_t>(tp[cols] - tp[0]);
    std::vector<double> st(z);
    if (z) SK_HIP(hipMemcpy(st.data(), d_st_ + tp[0], z * sizeof(double), hipMemcpyDeviceToHost), kStatusCopyFailed);
    std::vector<size_t> lp_start(cols + 1, 0);
    for (size_t j = 0; j < cols; ++j) {
      const size_t o = perm.empty() ? j : static_cast<size_t>(perm[j]);
      lp_start[o + 1] = static_cast<size_t>(tp[j + 1] - tp[j]);
    }
    for (size_t o = 0; o < cols; ++o) lp_start[o + 1] += lp_start[o];
    for (size_t j = 0; j < cols; ++j) {
      const size_t o = perm.empty() ? j : static_cast<size_t>(perm[j]);
      const auto b = static_cast<size_t>(tp[j] - tp[0]), e = static_cast<size_t>(tp[j + 1] - tp[0]);
      std::copy(st.begin() + static_cast<std::ptrdiff_t>(b), st.begin() + static_cast<std::ptrdiff_t>(e),
                values + lp_start[o]);
    }
  }
  return kStatusSuccess;
}

}  // namespace

int query_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    set_last_error("no HIP device is available");
    return kStatusNoDevice;
  }
  size_t f = 0, t = 0;
  SK_HIP(hipMemGetInfo(&f, &t), kStatusInitFailed);
  *free_bytes = static_cast<int64_t>(f);
  *total_bytes = static_cast<int64_t>(t);
  return kStatusSuccess;
}

std::unique_ptr<Engine> make_single_engine(const EngineOptions& options) {
  if (options.precision == 1) return std::make_unique<EngineImpl<double>>(options);
  return std::make_unique<EngineImpl<float>>(options);
}

int current_device(int* device) {
  SK_HIP(hipGetDevice(device), kStatusNoDevice);
  return kStatusSuccess;
}

int use_device(int device) {
  SK_HIP(hipSetDevice(device), kStatusNoDevice);
  return kStatusSuccess;
}

}  // namespace sk
