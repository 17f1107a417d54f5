// A batch split into sub-batches that run concurrently, each a block-diagonal batch on its own
// stream. The step kernels of one batch run as rounds of workgroups (a CU holds eight of their
// 256-lane workgroups), and each launch ends with a partial round and a drain; with two or
// three batches in flight on separate hardware queues the device fills one launch's tail with
// another's work. Instances do not interact, so each instance's answer is bitwise the same in
// any split (and alone).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "status.hpp"

namespace sk {
namespace {

// Batches below this many nonzeros in total run on one stream: their launches are short
// and mostly latency, so a split only adds host threads.
constexpr int64_t kSplitMinNnz = int64_t{1} << 20;
// Sub-batches of an automatic split. Each takes one of the process's hardware queues (four by
// default), and three measured best on the 256-LP batch.
constexpr int32_t kAutoStreams = 3;

class SplitEngine final : public Engine {
 public:
  explicit SplitEngine(const EngineOptions& options) : opt_(options) {}

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

  int upload() override {
    const int32_t k = static_cast<int32_t>(lps_.size());
    int64_t nnz = 0;
    for (const HostLp* lp : lps_) nnz += lp->nnz();
    int32_t parts = opt_.streams > 0 ? opt_.streams : (nnz >= kSplitMinNnz ? kAutoStreams : 1);
    parts = std::max(1, std::min(parts, k));
    // Longest-processing-time assignment by nonzeros (ties to the lower part), then each part
    // keeps its instances in batch order.
    std::vector<int32_t> order(static_cast<size_t>(k));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return lps_[a]->nnz() > lps_[b]->nnz(); });
    std::vector<int64_t> load(static_cast<size_t>(parts), 0);
    std::vector<int32_t> owner(static_cast<size_t>(k), 0);
    for (int32_t i : order) {
      const auto p = static_cast<int32_t>(std::min_element(load.begin(), load.end()) - load.begin());
      owner[static_cast<size_t>(i)] = p;
      load[static_cast<size_t>(p)] += lps_[static_cast<size_t>(i)]->nnz() + 1;
    }
    if (static_cast<int32_t>(parts_.size()) != parts) {
      parts_.clear();
      EngineOptions one = opt_;
      one.streams = 1;
      one.split_parts = parts;
      for (int32_t p = 0; p < parts; ++p) parts_.push_back(make_single_engine(one));
    } else {
      for (auto& e : parts_) e->clear();
    }
    where_.assign(static_cast<size_t>(k), Place{});
    back_.assign(static_cast<size_t>(parts), {});
    for (int32_t i = 0; i < k; ++i) {
      const int32_t p = owner[static_cast<size_t>(i)];
      where_[static_cast<size_t>(i)] = Place{p, parts_[static_cast<size_t>(p)]->add(lps_[static_cast<size_t>(i)])};
      back_[static_cast<size_t>(p)].push_back(i);
      const WarmStart& w = warm_[static_cast<size_t>(i)];
      if (w.set) {
        const int st = parts_[static_cast<size_t>(p)]->warm_start(where_[static_cast<size_t>(i)].local,
                                                                  w.x.data(), w.y.data(), w.omega, w.eta);
        if (st) return st;
      }
    }
    for (auto& e : parts_) {
      const int st = e->upload();
      if (st) return st;
    }
    return current_device(&device_);
  }

  int solve() override {
    if (parts_.empty()) {
      set_last_error("solve() before upload()");
      return kStatusInvalidState;
    }
    const size_t n = parts_.size();
    std::vector<int> status(n, kStatusSuccess);
    std::vector<std::string> errors(n);
    auto run = [&](size_t p) {
      int st = use_device(device_);
      if (!st) st = parts_[p]->solve();
      status[p] = st;
      if (st) errors[p] = last_error();
    };
    std::vector<std::thread> workers;
    try {
      for (size_t p = 1; p < n; ++p) workers.emplace_back(run, p);
    } catch (const std::exception& e) {
      for (auto& w : workers) w.join();
      set_last_error(std::string("cannot start a sub-batch thread: ") + e.what());
      return kStatusInitFailed;
    }
    run(0);
    for (auto& w : workers) w.join();
    for (size_t p = 0; p < n; ++p) {
      if (status[p]) {
        set_last_error(errors[p]);
        return status[p];
      }
    }
    gather_timings();
    return kStatusSuccess;
  }

  int result(int32_t index, InstanceResult* out) const override {
    if (index < 0 || static_cast<size_t>(index) >= where_.size()) {
      set_last_error("result(): no such instance");
      return kStatusInvalidArgument;
    }
    const Place w = where_[static_cast<size_t>(index)];
    return parts_[static_cast<size_t>(w.part)]->result(w.local, out);
  }

  void clear() override {
    lps_.clear();
    warm_.clear();
    where_.clear();
    back_.clear();
    for (auto& e : parts_) e->clear();
  }

  int start() override {
    if (parts_.empty()) {
      set_last_error("start() before upload()");
      return kStatusInvalidState;
    }
    done_.assign(parts_.size(), false);
    for (auto& e : parts_) {
      const int st = e->start();
      if (st) return st;
    }
    return kStatusSuccess;
  }

  int poll(bool wait, bool* finished, std::vector<int32_t>* newly_done) override {
    bool waited = false;
    std::vector<int32_t> local;
    for (size_t p = 0; p < parts_.size(); ++p) {
      if (done_[p]) continue;
      bool fin = false;
      local.clear();
      // block on the first unfinished sub-batch only; the others keep blocks queued
      const int st = parts_[p]->poll(wait && !waited, &fin, &local);
      waited = true;
      if (st) return st;
      for (int32_t i : local) newly_done->push_back(back_[p][static_cast<size_t>(i)]);
      done_[p] = fin;
    }
    *finished = std::all_of(done_.begin(), done_.end(), [](bool d) { return d; });
    if (*finished) gather_timings();
    return kStatusSuccess;
  }

  int64_t device_bytes() const override {
    int64_t b = 0;
    for (const auto& e : parts_) b += e->device_bytes();
    return b;
  }
  void release() override {
    for (auto& e : parts_) e->release();
  }
  const EngineTimings& timings() const override { return timings_; }
  int32_t size() const override { return static_cast<int32_t>(lps_.size()); }

  // One launch of a step kernel over the batch as the engine runs it: every sub-batch launches
  // it on its own stream, all at once, `reps` times in a row; the result is the wall time per
  // repetition from the common start until the last sub-batch is done. With one sub-batch this
  // is the plain kernel timing (HIP events on its stream).
  int time_kernel(int kernel, int reps, double* mean_ms) override {
    if (parts_.empty()) {
      set_last_error("time_kernel(): engine not uploaded");
      return kStatusInvalidState;
    }
    if (parts_.size() == 1) return parts_[0]->time_kernel(kernel, reps, mean_ms);
    const size_t n = parts_.size();
    std::vector<int> status(n, kStatusSuccess);
    std::vector<std::string> errors(n);
    std::vector<double> part_ms(n, 0.0);
    std::atomic<size_t> ready{0};
    std::atomic<bool> go{false};
    auto run = [&](size_t p) {
      int st = use_device(device_);
      ready.fetch_add(1);
      while (!go.load()) std::this_thread::yield();
      if (!st) st = parts_[p]->time_kernel(kernel, reps, &part_ms[p]);
      status[p] = st;
      if (st) errors[p] = last_error();
    };
    std::vector<std::thread> workers;
    try {
      for (size_t p = 0; p < n; ++p) workers.emplace_back(run, p);
    } catch (const std::exception& e) {
      go.store(true);  // the started workers time their part and end
      for (auto& w : workers) w.join();
      set_last_error(std::string("cannot start a sub-batch thread: ") + e.what());
      return kStatusInitFailed;
    }
    while (ready.load() < n) std::this_thread::yield();
    const auto t0 = std::chrono::steady_clock::now();
    go.store(true);
    for (auto& w : workers) w.join();
    const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (size_t p = 0; p < n; ++p) {
      if (status[p]) {
        set_last_error(errors[p]);
        return status[p];
      }
    }
    *mean_ms = total / reps;
    return kStatusSuccess;
  }
  double kernel_bytes(int kernel) const override {
    double b = 0.0;
    for (const auto& e : parts_) b += e->kernel_bytes(kernel);
    return b;
  }
  int64_t total_nnz() const override { return sum(&Engine::total_nnz); }
  int64_t total_rows() const override { return sum(&Engine::total_rows); }
  int64_t total_cols() const override { return sum(&Engine::total_cols); }
  int probe_spmv(int32_t index, int which, const double* in, double* out) override {
    const Place* w = place(index);
    return w ? parts_[static_cast<size_t>(w->part)]->probe_spmv(w->local, which, in, out) : kStatusInvalidArgument;
  }
  int probe_scaling(int32_t index, double* row_scale, double* col_scale, double* values) override {
    const Place* w = place(index);
    return w ? parts_[static_cast<size_t>(w->part)]->probe_scaling(w->local, row_scale, col_scale, values)
             : kStatusInvalidArgument;
  }
  int column_order(int32_t index, int32_t* device_to_lp) const override {
    const Place* w = place(index);
    return w ? parts_[static_cast<size_t>(w->part)]->column_order(w->local, device_to_lp) : kStatusInvalidArgument;
  }
  int stage_choice() const override { return parts_.empty() ? 1 : parts_[0]->stage_choice(); }
  int32_t streams() const override { return static_cast<int32_t>(parts_.size()); }
  int32_t fused() const override {
    // the least of the parts: a split batch is resident only if every part is
    int32_t least = parts_.empty() ? 0 : 1;
    for (const auto& e : parts_) least = std::min(least, e->fused());
    return least;
  }
  int32_t tile_capacity() const override {
    int32_t most = 0;
    for (const auto& e : parts_) most = std::max(most, e->tile_capacity());
    return most;
  }
  int32_t column_slices() const override {
    // a part slices only when it holds a single instance
    int32_t most = 0;
    for (const auto& e : parts_) most = std::max(most, e->column_slices());
    return most;
  }

 private:
  struct Place {
    int32_t part = 0;
    int32_t local = 0;
  };
  const Place* place(int32_t index) const {
    if (index < 0 || static_cast<size_t>(index) >= where_.size()) {
      set_last_error("no such instance");
      return nullptr;
    }
    return &where_[static_cast<size_t>(index)];
  }
  int64_t sum(int64_t (Engine::*f)() const) const {
    int64_t v = 0;
    for (const auto& e : parts_) v += ((*e).*f)();
    return v;
  }
  // the sub-batches ran side by side: each phase took as long as its slowest sub-batch
  void gather_timings() {
    timings_ = EngineTimings{};
    for (const auto& e : parts_) {
      const EngineTimings& t = e->timings();
      timings_.setup_ms = std::max(timings_.setup_ms, t.setup_ms);
      timings_.loop_ms = std::max(timings_.loop_ms, t.loop_ms);
      timings_.report_ms = std::max(timings_.report_ms, t.report_ms);
      timings_.graph_launches += t.graph_launches;
    }
  }

  EngineOptions opt_;
  std::vector<const HostLp*> lps_;
  std::vector<WarmStart> warm_;
  std::vector<Place> where_;                 // batch index -> sub-batch and its index there
  std::vector<std::vector<int32_t>> back_;   // sub-batch index -> batch index
  std::vector<std::unique_ptr<Engine>> parts_;
  std::vector<bool> done_;
  EngineTimings timings_;
  int device_ = 0;
};

}  // namespace

std::unique_ptr<Engine> make_engine(const EngineOptions& options) {
  if (options.streams == 1) return make_single_engine(options);
  return std::make_unique<SplitEngine>(options);
}

}  // namespace sk
