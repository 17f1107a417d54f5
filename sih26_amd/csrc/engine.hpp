// Host driver of the batched restarted-PDHG engine on one HIP device.
//
// A batch is any number of LPs ("instances") laid out block-diagonally in device memory. The
// solve runs entirely on the device: scaling, power iteration, the PDHG loop (captured as a
// hipGraph of `steps_per_block` iterations plus one evaluation stage) and the final report.
// The host only reads a finished-instance counter between graph launches.
#pragma once

#include <cfloat>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "device_types.hpp"
#include "lp.hpp"

namespace sk {

struct EngineOptions {
  double tolerance = 1e-4;        // pdhg_tolerance
  int64_t iteration_limit = -1;   // -1: 1,000,000 (pdhg.cpp:312-313)
  double time_limit = DBL_MAX;    // seconds
  int32_t restart = 1;            // pdhg_restart
  int32_t stop_at_request = 0;    // pdhg_stop_at_request
  int32_t absolute_clause = 1;    // meets_request also needs absolute primal <= 1e-7
  int64_t random_seed = 0;        // power iteration uses std::mt19937(seed + 1)
  int32_t precision = 0;          // 0: FP32 iterates + FP64 evaluation; 1: FP64 throughout
  int32_t steps_per_block = 40;   // PDHG steps per captured graph
  int32_t eval_stride = 20;       // evaluation stage after every this many steps
  int32_t use_graph = 1;
  int32_t stage = -1;             // staging order of SpMV tiles: -1 default (1), 0 CSR, 1 sorted
  int32_t tile_nnz = 0;           // row-block capacity; 0 picks per instance
  int32_t streams = 0;            // concurrent sub-batches: 0 automatic, 1 one stream, n >= 2
  int32_t column_order = 0;       // 0 automatic, 1 most-referenced columns first, 2 the LP's order
  // bit 0: never run the resident solve (k_resident, otherwise automatic when every instance
  // is small); bit 1: never let the primal launch make the previous step's decision
  int32_t step_fusion = 0;
  // A x-bar over column slices, one per XCD (each slice's x-bar entries stay in that XCD's
  // L2), for a batch of one instance: 0 automatic (instances of at least 2^22 columns, whose
  // x-bar is several times an L2), 1 on, 2 off. A sliced instance keeps the LP's column order.
  int32_t column_slices = 0;
  // set by a split batch for its parts: how many sub-batches run side by side (they share the
  // device's resident workgroups, so the one-round capacity rule of upload does not apply)
  int32_t split_parts = 1;
};

// A warm start held by an engine until its next upload (Engine::warm_start).
struct WarmStart {
  bool set = false;
  std::vector<double> x, y;  // LP order; NaN keeps an element's cold start
  double omega = 0.0, eta = 0.0;
};

struct InstanceResult {
  int32_t status = 0;
  int32_t converged = 0;
  int64_t iterations = 0;
  int64_t restarts = 0;
  int64_t evaluations = 0;
  int64_t attempts = 0;
  double spectral_norm = 0.0;
  double omega = 0.0;
  double eta = 0.0;
  Residuals final_residuals{};
  double dual_bound = 0.0;
  std::vector<double> col_value, col_dual, row_dual, row_activity;
};

struct EngineTimings {
  double setup_ms = 0.0;   // device scaling + power iteration + initial state
  double loop_ms = 0.0;    // PDHG graph launches until every instance stopped
  double report_ms = 0.0;  // final evaluation and download
  int64_t graph_launches = 0;
};

class Engine {
 public:
  virtual ~Engine() = default;
  // Instances are borrowed: they must stay alive until upload() returns.
  virtual int32_t add(const HostLp* lp) = 0;
  // Start instance `index` (already added, not yet uploaded) from x (cols, LP order) and y
  // (rows, the engine's sign: row_dual = -sense y) instead of proj(0) and 0; a NaN entry keeps
  // that element's cold start. omega and eta > 0 replace the initial primal weight and step
  // size. Takes effect at the next upload().
  virtual int warm_start(int32_t index, const double* x, const double* y, double omega,
                         double eta) = 0;
  virtual int upload() = 0;  // host CSR/tiles + device allocation + H2D (not part of a solve)
  virtual int solve() = 0;   // everything on the device, from the uploaded originals
  virtual int result(int32_t index, InstanceResult* out) const = 0;
  // Forget the instances added so far (a serving slot taking its next batch); device buffers
  // are kept and reused by the next upload() when large enough.
  virtual void clear() = 0;
  // Non-blocking solve for a serving loop: start() enqueues the set-up and the first blocks
  // and returns; each poll() returns at once, appends the instances that stopped since the
  // last call to `newly_done` (result() answers for them from then on, while the rest of the
  // batch keeps running) and sets `finished` once every instance is done and the device is
  // idle. solve() is start() followed by poll() until finished, minus the early read-backs.
  // With `wait`, poll() first blocks until the oldest block in flight has completed.
  virtual int start() = 0;
  virtual int poll(bool wait, bool* finished, std::vector<int32_t>* newly_done) = 0;
  // Device memory the engine holds, and a way to give it back (an idle serving slot).
  virtual int64_t device_bytes() const = 0;
  virtual void release() = 0;
  virtual const EngineTimings& timings() const = 0;
  virtual int32_t size() const = 0;
  // Launch one of the step kernels `reps` times in timing mode between two HIP events on the
  // engine's stream; returns the mean duration in ms. kernel: 0 primal, 1 dual (A), 2 A^T;
  // 3, 4, 5: the same in normal mode (tiles of stopped instances exit at once).
  virtual int time_kernel(int kernel, int reps, double* mean_ms) = 0;
  // Algorithmic bytes of one launch of that kernel over the whole batch.
  virtual double kernel_bytes(int kernel) const = 0;
  virtual int64_t total_nnz() const = 0;
  virtual int64_t total_rows() const = 0;
  virtual int64_t total_cols() const = 0;
  virtual int stage_choice() const = 0;
  // Sub-batches the batch runs as (each on its own stream).
  virtual int32_t streams() const = 0;
  // how the uploaded batch runs: 0 separate step launches; 1 the resident solve (k_resident),
  // one launch for many steps and evaluations of every instance
  virtual int32_t fused() const = 0;
  // column slices A x-bar runs over (EngineOptions::column_slices), 0 when unsliced
  virtual int32_t column_slices() const = 0;
  // the row-block tile capacity the uploaded batch runs with (the largest over sub-batches)
  virtual int32_t tile_capacity() const = 0;

  // Probes for tests, valid once a solve (or start()) has run the device set-up. They return
  // what the device holds or computes, in the LP's own row and column order.
  // probe_spmv: the step kernels' sparse product (step_product) on instance `index` alone:
  // which 1: out (rows) = A-hat in (cols), as k_step_dual forms A-hat x-bar; which 2: out
  // (cols) = A-hat^T in (rows), as k_step_aty forms A-hat^T y'. Inputs are rounded to the
  // iterate precision first, as the iterates are.
  virtual int probe_spmv(int32_t index, int which, const double* in, double* out) = 0;
  // probe_scaling: the device scaling's D_r (rows), D_c (cols) and the scaled FP64 values in
  // the LP's CSC order (nnz); any pointer may be null.
  virtual int probe_scaling(int32_t index, double* row_scale, double* col_scale,
                            double* values) = 0;
  // The LP column held in each device column of instance `index` (cols entries).
  virtual int column_order(int32_t index, int32_t* device_to_lp) const = 0;
};

// An engine for a batch: with options.streams other than 1 the batch is split into
// sub-batches that run concurrently, each on its own stream (split.cpp).
std::unique_ptr<Engine> make_engine(const EngineOptions& options);

// Row-block tile capacity when EngineOptions::tile_nnz is 0. A step launch over a batch runs
// its tiles as a few rounds of the workgroups the device holds at once (2,048: 8 per CU);
// measured on the 256-LP batch, 2,560 nonzeros per tile gives ~1.9 rounds and step launches
// ~8% shorter than 2,048 (~2.4 rounds, a long partly-occupied last round).
constexpr int32_t kDefaultTileNnz = 2560;
// The capacity for a batch whose CSR(A^T) tiles at the default capacity are fewer than one
// round of resident workgroups, and for one whose tiles are fewer than a quarter of a round
// (EngineImpl::upload).
constexpr int32_t kOneRoundTileNnz = 1536;
constexpr int32_t kQuarterRoundTileNnz = 1024;

// Column slices (EngineOptions::column_slices): one per XCD, used automatically for a batch of
// one instance of at least kSliceMinCols columns. Slice s holds the columns
// [s * n / kColumnSlices, (s + 1) * n / kColumnSlices) of the LP's column order. Measured on the
// 2M x 5M LP as a batch of its 8 column slices (each slice's tiles on one XCD): A x-bar 228 us
// against 262 us in the most-referenced-first order and 359 us in the LP's order, with the dual
// step's per-row work done eight times over.
constexpr int32_t kColumnSlices = 8;
constexpr int32_t kSliceMinCols = 1 << 22;
// The same batches run A^T y' over kRowSlices slices of A's rows (y' entries), two XCDs per
// slice. Measured as a batch of the 2M x 5M LP's 4 row slices: A^T y' 247 us against 289 us
// unsliced (LP column order), with the interaction's per-column work done four times over; 8
// slices (1 MB of y' each, but 40M partial sums): 322 us.
constexpr int32_t kRowSlices = 4;

// One block-diagonal batch on one stream.
std::unique_ptr<Engine> make_single_engine(const EngineOptions& options);

// The calling thread's HIP device, and selecting one (status codes of status.hpp).
int current_device(int* device);
int use_device(int device);

// Device bytes an engine allocates for a batch with these totals (rows, columns, nonzeros
// and instances summed over the batch), rounded up; used for admission by memory footprint.
int64_t estimate_device_bytes(int64_t rows, int64_t cols, int64_t nnz, int64_t instances,
                              int32_t precision);

// Free and total memory of the current HIP device.
int query_device_memory(int64_t* free_bytes, int64_t* total_bytes);

// Injected faults (sk_fault_inject), process-wide, each firing once: the engine's next device
// allocation or kernel launch that the countdown selects fails, and the next upload puts a NaN
// into the cost of one instance's first column.
bool fault_allocation_fails();
bool fault_launch_fails();
int64_t fault_take_nan_instance();
// True for each of the next n triangular solves of the interior point (kind 4): its blocks give
// up at their first wait, so the solve takes the rocBLAS path.
bool fault_solve_gives_up();

// Thread-local description of the last failure inside the library.
void set_last_error(const std::string& message);
const char* last_error();

}  // namespace sk
