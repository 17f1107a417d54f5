// Presolve reductions and postsolve (src/presolve/presolve.hpp:36-127; presolve.cpp).
//
// The reductions the reference's solve() applies ahead of every LP engine (solve.cpp:415-443),
// on the host: the records, the order of the passes and every test and tolerance are the
// reference's, so the reduced LP is the one the reference hands its engine. The reductions are
// sequential by nature (a reduction changes the bounds the next test in the same pass reads),
// so they run on one host thread, in O(nnz) per pass; only the reduced LP goes to the device.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lp.hpp"

namespace sk::presolve {

// presolve.hpp:43-64. The values are those of include/sk_gpu.h's SK_PRESOLVE_*.
enum class Kind : int32_t {
  kEmptyRow = 0,
  kRedundantRow = 1,
  kFixedColumn = 2,
  kEmptyColumn = 3,
  kSingletonRow = 4,
  kForcingRow = 5,  // in the vocabulary; no pass produces it (postsolve prices it at zero)
  kFreeColumnSingleton = 6,
  kDoubletonEquation = 7,
};
constexpr int kKinds = 8;

// presolve.hpp:66-91: what a reduction did, so postsolve can undo it.
struct Record {
  Kind kind = Kind::kEmptyRow;
  int32_t index = -1;  // the original row or column the record is about
  double value = 0.0;  // the value a removed column takes
  double coefficient = 0.0;
  double row_lower = 0.0, row_upper = 0.0;
  int32_t column = -1;  // a singleton row's column; the eliminated column of the last two kinds
  double substituted_rhs = 0.0;
  int32_t partner_column = -1;  // doubleton: the column kept
  double partner_coefficient = 0.0;
  double eliminated_cost = 0.0;  // the eliminated column's cost when it was eliminated
};

// presolve.hpp:94-117.
struct Result {
  HostLp reduced;
  std::vector<Record> records;  // in the order applied; postsolve replays them backwards
  std::vector<int32_t> col_to_original, row_to_original;
  int32_t original_rows = 0, original_cols = 0;
  int64_t original_nonzeros = 0;
  int32_t passes = 0;
  bool proved_infeasible = false;
  std::string message;
};

// presolve.cpp:204-808. `is_integer` holds one flag per column or is empty (every column
// continuous); `feasibility` is the primal_feasibility_tolerance option (default 1e-7).
Result run(const HostLp& lp, const std::vector<uint8_t>& is_integer, double feasibility);

// presolve.cpp:812-1697 without the quality measures (the caller recomputes them on the
// original LP): a point of the reduced LP in the Solution's signs (col_dual = c - A^T row_dual)
// back to the original LP. Output arrays: original num_cols, num_rows, num_cols.
void postsolve(const Result& result, const HostLp& original, const double* col_value,
               const double* row_dual, const double* col_dual, double* out_col_value,
               double* out_row_dual, double* out_col_dual);

}  // namespace sk::presolve
