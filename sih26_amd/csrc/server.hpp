// Continuous-batching request loop over the batched PDHG engine, on one HIP device.
//
// Requests (one LP each) enter a FIFO queue. Each of `slots` worker threads owns `lanes`
// engines, each with its own stream; whenever a slot has nothing running it takes, from the
// head of the queue, every waiting request that fits its batch limits and the device-memory
// budget, uploads them as one block-diagonal batch on an idle lane and starts it. Slots run
// concurrently, so the batches overlap on the device. While a batch runs, every instance that
// stops is read back and completed at once (it retires mid-batch). Once a batch is in its tail
// (at most `tail_fraction` of its instances still running) the slot admits the queue head into
// its next idle lane, so the capacity the retired instances freed serves new requests while
// the tail finishes; a slot runs at most `lanes` batches at once.
//
// Admission by memory footprint: the server accounts the device memory every slot holds
// (engines keep their buffers between batches, growing each one as needed) and admits a batch
// only if the slot's projected holding, the estimate for the per-dimension maxima of the
// batch and of what the slot ran before, plus what the other slots hold stays within the
// budget. A request larger than the whole budget is rejected at
// submission; idle slots give their memory back when the queue head does not fit.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "lp.hpp"

namespace sk {

struct ServerOptions {
  EngineOptions engine;
  int32_t slots = 2;           // concurrent batches (engines, each with its own streams)
  int32_t max_batch = 256;     // instances per batch
  int64_t memory_budget = 0;   // device bytes all slots may hold; 0: 90% of free memory
  int32_t lanes = 1;           // batches one slot may run at once (1: one batch per slot)
  double tail_fraction = 0.25; // a batch is in its tail when at most this share of its
                               // instances is still running
};

struct Request {
  int64_t id = 0;
  std::unique_ptr<HostLp> lp;
  int64_t footprint = 0;   // estimate_device_bytes of this LP alone
  double arrival = 0.0;    // server clock, seconds
  double admitted = 0.0;   // its batch was formed
  double finished = 0.0;   // its answer was read back
  int32_t slot = -1;
  int32_t batch_size = 0;
  int32_t error = 0;       // kStatus* of the batch, when it failed
  std::string error_message;
  InstanceResult result;
};

struct ServerStats {
  int64_t submitted = 0;
  int64_t completed = 0;
  int64_t rejected = 0;
  int64_t failed = 0;
  int64_t batches = 0;
  int64_t max_queue = 0;
  int64_t max_batch_seen = 0;
  int64_t memory_budget = 0;
  int64_t peak_held = 0;      // largest total device memory held by the slots
  int64_t releases = 0;       // idle slots that gave their memory back
};

class Server {
 public:
  // Called on a slot thread for every request that completes (or fails), with its LP still
  // alive; the server frees the LP after the hook returns.
  using Hook = std::function<void(Request&)>;

  Server(const ServerOptions& options, Hook on_complete);
  ~Server();  // finishes the running batches, drops the queue, joins the slot threads
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  int open();  // checks the device and starts the slot threads
  // Takes ownership of `lp` when it accepts the request (`lp` is then null); a refused LP
  // stays with the caller. `arrival` < 0 means now (server clock).
  int submit(std::unique_ptr<HostLp>& lp, double arrival, int64_t* id);
  // Blocks until every submitted request has completed, or `timeout` seconds passed;
  // returns true when idle.
  bool wait_idle(double timeout);
  double now() const;
  ServerStats stats() const;

 private:
  // Batch totals; a slot's engine buffers grow per dimension, so the largest totals it has run
  // since its last release bound what it holds.
  struct Extent {
    int64_t rows = 0, cols = 0, nnz = 0, instances = 0;
  };
  // One engine of a slot and the batch it runs. `held` and `extent` are shared with the other
  // slot threads (under mu_); the batch and its counts belong to the slot's own thread.
  struct Lane {
    std::unique_ptr<Engine> engine;
    int64_t held = 0;   // device bytes its engine holds
    Extent extent;      // per-dimension maxima of its batches since the last release
    bool running = false;
    std::vector<std::unique_ptr<Request>> batch;
    int32_t open = 0;   // instances of the batch still running
    uint64_t started = 0;  // order in which the slot's lanes took their batches
  };
  struct Slot {
    std::vector<Lane> lanes;
  };
  void run_slot(int index);
  // Forms the next batch of lane `lane` of slot `index`; with `block`, waits until one can
  // be formed. Returns false when none was formed (or the server stops).
  bool take_batch(int index, int lane, std::vector<std::unique_ptr<Request>>* batch, bool block);
  int start_lane(Lane& lane);
  // Reads back what stopped; returns whether anything did. A failed batch completes every
  // request still in it with the error.
  bool poll_lane(Lane& lane, bool wait);
  void complete(std::unique_ptr<Request> r);
  int64_t held_by_others(int index, int lane) const;
  void release_lane(Lane& lane);  // under mu_

  ServerOptions opt_;
  Hook hook_;
  std::chrono::steady_clock::time_point t0_;
  mutable std::mutex mu_;
  std::condition_variable work_cv_;   // queue changed, memory freed, or stopping
  std::condition_variable idle_cv_;   // a request completed
  std::deque<std::unique_ptr<Request>> queue_;
  std::vector<Slot> slots_;
  std::vector<std::thread> threads_;
  bool stopping_ = false;
  bool release_wanted_ = false;
  int release_requester_ = -1;  // slot * kMaxLanes + lane
  int64_t next_id_ = 0;
  int64_t outstanding_ = 0;
  ServerStats stats_;
};

}  // namespace sk
