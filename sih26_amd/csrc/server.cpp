// Continuous-batching request loop (see server.hpp).
#include "server.hpp"

#include <algorithm>
#include <utility>

#include "status.hpp"

namespace sk {

namespace {

// Sums of one candidate batch, for its footprint estimate.
struct Totals {
  int64_t rows = 0, cols = 0, nnz = 0, instances = 0;
  void add(const HostLp& lp) {
    rows += lp.num_rows;
    cols += lp.num_cols;
    nnz += lp.nnz();
    instances += 1;
  }
};

// The engine's 32-bit batch offsets bound a batch's totals.
constexpr int64_t kMaxBatchExtent = INT32_MAX - 1;

}  // namespace

Server::Server(const ServerOptions& options, Hook on_complete)
    : opt_(options), hook_(std::move(on_complete)), t0_(std::chrono::steady_clock::now()) {
  opt_.slots = std::max(1, opt_.slots);
  opt_.max_batch = std::max(1, opt_.max_batch);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stopping_ = true;
  }
  work_cv_.notify_all();
  for (auto& t : threads_) {
    if (t.joinable()) t.join();
  }
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int Server::open() {
  if (!threads_.empty()) {
    set_last_error("server already open");
    return kStatusInvalidState;
  }
  int64_t free_bytes = 0, total_bytes = 0;
  const int st = query_device_memory(&free_bytes, &total_bytes);
  if (st) return st;
  if (opt_.memory_budget <= 0) opt_.memory_budget = static_cast<int64_t>(0.9 * static_cast<double>(free_bytes));
  stats_.memory_budget = opt_.memory_budget;
  slots_.resize(static_cast<size_t>(opt_.slots));
  for (auto& s : slots_) s.engine = make_single_engine(opt_.engine);
  for (int i = 0; i < opt_.slots; ++i) threads_.emplace_back(&Server::run_slot, this, i);
  return kStatusSuccess;
}

int Server::submit(std::unique_ptr<HostLp>& lp, double arrival, int64_t* id) {
  if (!lp) {
    set_last_error("submit: no LP");
    return kStatusInvalidArgument;
  }
  const std::string problem = lp->validate();
  if (!problem.empty()) {
    set_last_error("submit: " + problem);
    return kStatusInvalidArgument;
  }
  auto r = std::make_unique<Request>();
  r->footprint = estimate_device_bytes(lp->num_rows, lp->num_cols, lp->nnz(), 1, opt_.engine.precision);
  r->arrival = arrival < 0.0 ? now() : arrival;
  std::lock_guard<std::mutex> lk(mu_);
  if (threads_.empty() || stopping_) {
    set_last_error("submit: server is not open");
    return kStatusInvalidState;
  }
  if (r->footprint > opt_.memory_budget || lp->num_rows > kMaxBatchExtent ||
      lp->num_cols > kMaxBatchExtent || lp->nnz() > kMaxBatchExtent) {
    ++stats_.rejected;
    set_last_error("submit: the LP needs about " + std::to_string(r->footprint) +
                   " device bytes, more than the server's budget of " +
                   std::to_string(opt_.memory_budget));
    return kStatusOutOfMemory;
  }
  r->lp = std::move(lp);
  r->id = next_id_++;
  if (id) *id = r->id;
  queue_.push_back(std::move(r));
  ++stats_.submitted;
  ++outstanding_;
  stats_.max_queue = std::max<int64_t>(stats_.max_queue, static_cast<int64_t>(queue_.size()));
  work_cv_.notify_all();
  return kStatusSuccess;
}

bool Server::wait_idle(double timeout) {
  std::unique_lock<std::mutex> lk(mu_);
  return idle_cv_.wait_for(lk, std::chrono::duration<double>(std::max(0.0, timeout)),
                           [&] { return outstanding_ == 0; });
}

ServerStats Server::stats() const {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

int64_t Server::held_by_others(int index) const {
  int64_t total = 0;
  for (size_t k = 0; k < slots_.size(); ++k) {
    if (static_cast<int>(k) != index) total += slots_[k].held;
  }
  return total;
}

// Forms the slot's next batch from the head of the queue (FIFO, no overtaking). Returns false
// when the server stops.
bool Server::take_batch(int index, std::vector<std::unique_ptr<Request>>* batch) {
  std::unique_lock<std::mutex> lk(mu_);
  Slot& me = slots_[static_cast<size_t>(index)];
  me.busy = false;
  for (;;) {
    if (stopping_) return false;
    if (release_wanted_ && release_requester_ != index && me.held > 0) {
      // an idle slot gives its buffers back so a waiting request can fit
      me.engine->release();
      me.held = 0;
      me.extent = Extent{};
      ++stats_.releases;
      work_cv_.notify_all();
    }
    if (!queue_.empty()) {
      const int64_t others = held_by_others(index);
      Totals t;
      size_t take = 0;
      for (const auto& r : queue_) {
        if (static_cast<int32_t>(take) >= opt_.max_batch) break;
        Totals next = t;
        next.add(*r->lp);
        if (next.rows > kMaxBatchExtent || next.cols > kMaxBatchExtent || next.nnz > kMaxBatchExtent) break;
        const int64_t est = estimate_device_bytes(
            std::max(next.rows, me.extent.rows), std::max(next.cols, me.extent.cols),
            std::max(next.nnz, me.extent.nnz), std::max(next.instances, me.extent.instances),
            opt_.engine.precision);
        if (others + std::max(me.held, est) > opt_.memory_budget) break;
        t = next;
        ++take;
      }
      if (take > 0) {
        me.extent.rows = std::max(me.extent.rows, t.rows);
        me.extent.cols = std::max(me.extent.cols, t.cols);
        me.extent.nnz = std::max(me.extent.nnz, t.nnz);
        me.extent.instances = std::max(me.extent.instances, t.instances);
        // reserve the projection now, so slots admitting batches meanwhile count it; the
        // slot's real holding (never larger) replaces it after the upload
        me.held = std::max(me.held, estimate_device_bytes(me.extent.rows, me.extent.cols,
                                                          me.extent.nnz, me.extent.instances,
                                                          opt_.engine.precision));
        const double when = now();
        for (size_t k = 0; k < take; ++k) {
          std::unique_ptr<Request> r = std::move(queue_.front());
          queue_.pop_front();
          r->admitted = when;
          r->slot = index;
          r->batch_size = static_cast<int32_t>(take);
          batch->push_back(std::move(r));
        }
        me.busy = true;
        release_wanted_ = false;  // the head moved on
        ++stats_.batches;
        stats_.max_batch_seen = std::max<int64_t>(stats_.max_batch_seen, static_cast<int64_t>(take));
        return true;
      }
      // This slot's buffers grow per dimension, so the projection above takes the maxima of
      // its earlier batches and the head (a row-heavy batch, then an nnz-heavy request):
      // when the head fits on its own next to the other slots, the slot gives its buffers
      // back and forms the batch from nothing.
      if (me.held > 0) {
        const HostLp& head = *queue_.front()->lp;
        const int64_t alone = estimate_device_bytes(head.num_rows, head.num_cols, head.nnz(), 1,
                                                    opt_.engine.precision);
        if (others + alone <= opt_.memory_budget) {
          me.engine->release();
          me.held = 0;
          me.extent = Extent{};
          ++stats_.releases;
          continue;
        }
      }
      // the head does not fit next to what the other slots hold: ask the idle ones to release
      if (!release_wanted_) {
        release_wanted_ = true;
        release_requester_ = index;
        work_cv_.notify_all();
      }
    }
    work_cv_.wait(lk);
  }
}

void Server::complete(std::unique_ptr<Request> r) {
  if (hook_) hook_(*r);
  const bool failed = r->error != 0;
  r.reset();  // frees the LP and the answer
  std::lock_guard<std::mutex> lk(mu_);
  ++stats_.completed;
  if (failed) ++stats_.failed;
  --outstanding_;
  idle_cv_.notify_all();
}

void Server::run_slot(int index) {
  Slot& slot = slots_[static_cast<size_t>(index)];
  Engine& e = *slot.engine;
  std::vector<std::unique_ptr<Request>> batch;
  std::vector<int32_t> fresh;
  while (true) {
    batch.clear();
    if (!take_batch(index, &batch)) return;
    e.clear();
    for (const auto& r : batch) e.add(r->lp.get());
    int st = e.upload();
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot.held = e.device_bytes();
      int64_t total = 0;
      for (const auto& s : slots_) total += s.held;
      stats_.peak_held = std::max(stats_.peak_held, total);
    }
    if (!st) st = e.start();
    while (!st) {
      bool finished = false;
      fresh.clear();
      st = e.poll(true, &finished, &fresh);
      for (const int32_t k : fresh) {
        std::unique_ptr<Request>& r = batch[static_cast<size_t>(k)];
        const int rs = e.result(k, &r->result);
        r->finished = now();
        if (rs) {
          r->error = rs;
          r->error_message = last_error();
        }
        complete(std::move(r));
      }
      if (finished) break;
    }
    if (st) {
      // the batch failed: every request still in it completes with the error, and the
      // engine starts over from nothing
      const std::string message = last_error();
      for (auto& r : batch) {
        if (!r) continue;
        r->finished = now();
        r->error = st;
        r->error_message = message;
        complete(std::move(r));
      }
      e.release();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot.held = e.device_bytes();
      if (st) slot.extent = Extent{};
    }
    work_cv_.notify_all();
  }
}

}  // namespace sk
