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
constexpr int kMaxLanes = 8;

}  // namespace

Server::Server(const ServerOptions& options, Hook on_complete)
    : opt_(options), hook_(std::move(on_complete)), t0_(std::chrono::steady_clock::now()) {
  opt_.slots = std::max(1, opt_.slots);
  opt_.max_batch = std::max(1, opt_.max_batch);
  opt_.lanes = std::min(kMaxLanes, std::max(1, opt_.lanes));
  opt_.tail_fraction = std::min(1.0, std::max(0.0, opt_.tail_fraction));
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
  for (auto& s : slots_) {
    s.lanes.resize(static_cast<size_t>(opt_.lanes));
    for (auto& l : s.lanes) l.engine = make_single_engine(opt_.engine);
  }
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

int64_t Server::held_by_others(int index, int lane) const {
  int64_t total = 0;
  for (size_t k = 0; k < slots_.size(); ++k)
    for (size_t l = 0; l < slots_[k].lanes.size(); ++l)
      if (static_cast<int>(k) != index || static_cast<int>(l) != lane) total += slots_[k].lanes[l].held;
  return total;
}

void Server::release_lane(Lane& lane) {
  lane.engine->release();
  lane.held = 0;
  lane.extent = Extent{};
  ++stats_.releases;
}

// Forms the lane's next batch from the head of the queue (FIFO, no overtaking).
bool Server::take_batch(int index, int lane, std::vector<std::unique_ptr<Request>>* batch, bool block) {
  std::unique_lock<std::mutex> lk(mu_);
  Lane& me = slots_[static_cast<size_t>(index)].lanes[static_cast<size_t>(lane)];
  const int self = index * kMaxLanes + lane;
  for (;;) {
    if (stopping_) return false;
    if (release_wanted_ && release_requester_ != self && me.held > 0) {
      // an idle lane gives its buffers back so a waiting request can fit
      release_lane(me);
      work_cv_.notify_all();
    }
    if (!queue_.empty()) {
      const int64_t others = held_by_others(index, lane);
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
        // reserve the projection now, so lanes admitting batches meanwhile count it; the
        // lane's real holding (never larger) replaces it after the upload
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
        release_wanted_ = false;  // the head moved on
        ++stats_.batches;
        stats_.max_batch_seen = std::max<int64_t>(stats_.max_batch_seen, static_cast<int64_t>(take));
        return true;
      }
      // This lane's buffers grow per dimension, so the projection above takes the maxima of
      // its earlier batches and the head (a row-heavy batch, then an nnz-heavy request):
      // when the head fits on its own next to the others, the lane gives its buffers back
      // and forms the batch from nothing.
      if (me.held > 0) {
        const HostLp& head = *queue_.front()->lp;
        const int64_t alone = estimate_device_bytes(head.num_rows, head.num_cols, head.nnz(), 1,
                                                    opt_.engine.precision);
        if (others + alone <= opt_.memory_budget) {
          release_lane(me);
          continue;
        }
      }
      // the head does not fit next to what the others hold: ask the idle lanes to release
      if (!release_wanted_) {
        release_wanted_ = true;
        release_requester_ = self;
        work_cv_.notify_all();
      }
    }
    if (!block) return false;
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

int Server::start_lane(Lane& lane) {
  Engine& e = *lane.engine;
  e.clear();
  for (const auto& r : lane.batch) e.add(r->lp.get());
  lane.running = true;
  lane.open = static_cast<int32_t>(lane.batch.size());
  int st = e.upload();
  {
    std::lock_guard<std::mutex> lk(mu_);
    lane.held = e.device_bytes();
    int64_t total = 0;
    for (const auto& s : slots_)
      for (const auto& l : s.lanes) total += l.held;
    stats_.peak_held = std::max(stats_.peak_held, total);
  }
  if (!st) st = e.start();
  return st;
}

bool Server::poll_lane(Lane& lane, bool wait) {
  Engine& e = *lane.engine;
  bool finished = false;
  std::vector<int32_t> fresh;
  int st = e.poll(wait, &finished, &fresh);
  for (const int32_t k : fresh) {
    std::unique_ptr<Request>& r = lane.batch[static_cast<size_t>(k)];
    const int rs = e.result(k, &r->result);
    r->finished = now();
    if (rs) {
      r->error = rs;
      r->error_message = last_error();
    }
    --lane.open;
    complete(std::move(r));
  }
  if (!st && !finished) return !fresh.empty();
  if (st) {
    // the batch failed: every request still in it completes with the error, and the
    // engine starts over from nothing
    const std::string message = last_error();
    for (auto& r : lane.batch) {
      if (!r) continue;
      r->finished = now();
      r->error = st;
      r->error_message = message;
      complete(std::move(r));
    }
    e.release();
  }
  lane.batch.clear();
  lane.running = false;
  lane.open = 0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    lane.held = e.device_bytes();
    if (st) lane.extent = Extent{};
  }
  work_cv_.notify_all();
  return true;
}

void Server::run_slot(int index) {
  Slot& slot = slots_[static_cast<size_t>(index)];
  const int lanes = static_cast<int>(slot.lanes.size());
  uint64_t order = 0;
  auto failed_start = [&](Lane& lane, int st) {
    // an upload or start that failed: the batch completes with the error, as a failed run does
    const std::string message = last_error();
    for (auto& r : lane.batch) {
      r->finished = now();
      r->error = st;
      r->error_message = message;
      complete(std::move(r));
    }
    lane.batch.clear();
    lane.running = false;
    lane.engine->release();
    std::lock_guard<std::mutex> lk(mu_);
    lane.held = 0;
    lane.extent = Extent{};
    work_cv_.notify_all();
  };
  for (;;) {
    int active = 0, newest = -1, idle = -1;
    for (int l = 0; l < lanes; ++l) {
      Lane& ln = slot.lanes[static_cast<size_t>(l)];
      if (ln.running) {
        ++active;
        if (newest < 0 || ln.started > slot.lanes[static_cast<size_t>(newest)].started) newest = l;
      } else if (idle < 0) {
        idle = l;
      }
    }
    // admission: a slot with nothing running waits for work; one whose newest batch is in its
    // tail takes the queue head into an idle lane when it fits
    if (active == 0) {
      Lane& ln = slot.lanes[static_cast<size_t>(idle)];
      if (!take_batch(index, idle, &ln.batch, true)) return;  // stopping
      ln.started = ++order;
      if (const int st = start_lane(ln)) failed_start(ln, st);
      continue;
    }
    if (idle >= 0) {
      const Lane& nl = slot.lanes[static_cast<size_t>(newest)];
      const bool tail = static_cast<double>(nl.open) <= opt_.tail_fraction * static_cast<double>(nl.batch.size());
      if (tail) {
        Lane& ln = slot.lanes[static_cast<size_t>(idle)];
        if (take_batch(index, idle, &ln.batch, false)) {
          ln.started = ++order;
          if (const int st = start_lane(ln)) failed_start(ln, st);
          continue;
        }
      }
    }
    {
      // idle lanes of a busy slot answer a release request too
      std::lock_guard<std::mutex> lk(mu_);
      if (release_wanted_)
        for (int l = 0; l < lanes; ++l) {
          Lane& ln = slot.lanes[static_cast<size_t>(l)];
          if (!ln.running && ln.held > 0 && release_requester_ != index * kMaxLanes + l) {
            release_lane(ln);
            work_cv_.notify_all();
          }
        }
    }
    if (active == 1) {
      poll_lane(slot.lanes[static_cast<size_t>(newest)], true);
    } else {
      bool progressed = false;
      for (int l = 0; l < lanes; ++l)
        if (slot.lanes[static_cast<size_t>(l)].running) progressed |= poll_lane(slot.lanes[static_cast<size_t>(l)], false);
      if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
  }
}

}  // namespace sk
