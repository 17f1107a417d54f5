// Plain-old-data types shared by the HIP kernels and the host driver of the batched
// restarted-PDHG engine. Every LP in a batch ("instance") owns contiguous ranges of the
// block-diagonal row, column and nonzero arrays, and contiguous ranges of each tile table.
#pragma once

#include <cstdint>

namespace sk {

// Launch geometry shared by every kernel: one 256-lane workgroup (four wave64s) per tile.
constexpr int kBlock = 256;
constexpr int kWave = 64;
// Elementwise tiles cover this many consecutive columns or rows.
constexpr int kChunk = 1024;
// Largest nonzero count a row-block tile may stage in LDS; a single row above it is
// processed by the whole workgroup without staging ("vector" tile).
constexpr int kMaxTileNnz = 4096;
// Most rows a row-block tile holds: a short-row tile gives each lane of a 256-lane workgroup
// up to two rows, a long-row tile each wavefront up to two.
constexpr int kMaxTileRows = 2 * kBlock;
// Extra elements allocated after the staging-order matrix arrays: the step kernels load them
// in aligned groups of four, and the last group may reach three elements past the end.
constexpr int kStagePad = 4;

// A tile is a half-open range [begin, end) of rows (CSR tables) or of elements (chunk
// tables) that belongs to exactly one instance. `cap` is the staging capacity the tile was
// built for (row-block tables) so a kernel knows whether it is a stream or a vector tile, and
// [p0, p1) is its nonzero range (row-block tables; zero for chunks), so a kernel can issue
// its matrix loads as soon as the tile record arrives. `base` is the first gather index of
// the tile's instance (its first column for CSR(A), first row for CSR(A^T)): packed index
// words hold gather indices relative to it. `kind` is the row-block schedule (kTile*).
struct Tile {
  int32_t begin;
  int32_t end;
  int32_t inst;
  int32_t cap;
  int32_t p0;
  int32_t p1;
  int32_t base;
  int32_t kind;
};

// Row-block schedules. Each row's sum has one fixed order that depends only on its length, so
// the result of a row is the same in every tiling (tile capacity and neighbours do not matter):
// a row of at most kLongRow nonzeros is summed by one lane in CSR order; a longer row by the
// 64 lanes of a wavefront, lane l taking nonzeros l, l+64, ... in order, then a butterfly.
constexpr int kLongRow = 64;
enum : int32_t {
  kTileShort = 0,   // up to kMaxTileRows short rows, one lane each
  kTileLong = 1,    // up to kMaxTileRows / kWave long rows, one wavefront each
  kTileVector = 2   // one row with more than `cap` nonzeros, read straight from memory
};

// Per-tile flags, rewritten for every tile of an instance whenever its state changes (set-up,
// the finalize kernels, and the host after it edits a record), so a step or evaluation kernel
// learns whether its tile has work, and which iterate buffers it uses, from a load that does
// not wait for the instance record.
enum : uint32_t {
  kFlagStepping = 1u,  // not stopped and no evaluation pending
  kFlagParity = 2u,    // the current iterate is buffer 1
  kFlagAccept = 4u,    // the last step was accepted and is not yet in the running sums
  kFlagEval = 8u,      // an evaluation is pending
  kFlagReport = 16u,   // the final (report) evaluation is pending
  kFlagCommit = 32u    // the evaluation's commit stage has work (best copy or restart)
};

// pdhg.cpp:81-132 (Residuals)
struct Residuals {
  double primal;
  double dual;
  double gap;
  double primal_objective;
  double dual_objective;
  double absolute_primal;
  double absolute_dual;
  double gap_as_verified;
  double complementarity;
};

// Where an instance's data lives in the batch arrays and tile tables.
struct InstDesc {
  int32_t row0, rows, col0, cols;
  int32_t ta0, ta1;  // row blocks of CSR(A)
  int32_t tt0, tt1;  // row blocks of CSR(A^T) (one row per column of A)
  int32_t tc0, tc1;  // column chunks
  int32_t tr0, tr1;  // row chunks
};

// Statuses, numbered as sankhya::SolveStatus (include/sankhya/model.hpp:62-80).
enum : int32_t {
  kStNotSolved = 0,
  kStOptimal = 1,
  kStFeasible = 5,
  kStIterationLimit = 6,
  kStTimeLimit = 7,
  kStNumericalError = 9,
  kStModelError = 10
};

// Mutable per-instance solver state; lives in device memory for the whole solve and is
// written only by the per-instance finalize kernels (and the set-up kernels).
struct InstState {
  // step control (pdhg.cpp:368-373)
  double eta;
  double omega;
  double spectral_norm;
  double eta_ceiling;
  double restart_kkt;
  double best_worst;
  // a warm start's step size (0: the reference's 1 / ||A-hat||, pdhg.cpp:368)
  double eta_start;
  // the step-size rule's shrink and grow factors at `iteration` (pdhg.cpp:472-474)
  double rule_shrink;
  double rule_grow;
  // termination options and problem norms (pdhg.cpp:285-315)
  double tolerance;
  double cost_norm;
  double bound_norm;
  // diagnostics of the last step / evaluation
  double movement;
  double interaction;
  double dx_norm;
  double dy_norm;
  int64_t iteration;
  int64_t averaged;
  int64_t last_restart;
  int64_t restarts;
  int64_t evaluations;
  int64_t iteration_limit;
  int64_t attempts;
  int32_t parity;          // current iterate is buffer `parity`; the candidate is 1 - parity
  int32_t pending_accept;  // the last step was accepted and not yet added to the sums
  int32_t eval_pending;    // stepping pauses until the evaluation stage has run
  int32_t done;
  int32_t status;
  int32_t converged;
  int32_t use_restarts;
  int32_t stop_at_request;
  int32_t absolute_clause;  // meets_request also demands absolute primal <= 1e-7
  int32_t chosen_avg;       // last evaluation picked the running average
  int32_t copy_best;        // commit stage copies the chosen candidate into best
  int32_t do_restart;       // commit stage resets sums and restart point
  int32_t report_pending;   // final evaluation of `best` still to run
  int32_t nonfinite;        // a non-finite step quantity was seen
  int32_t norm_trivial;     // no rows, columns or nonzeros: spectral norm is 1 by definition
  int32_t pad1;
  Residuals current;
  Residuals average;
  Residuals best;
  Residuals report;
};

}  // namespace sk
