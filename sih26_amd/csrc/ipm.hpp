// Interior-point finish of a first-order answer: Mehrotra's predictor-corrector on the normal
// equations (reference src/ipm/ipm.cpp; include/sankhya/ipm.hpp:10-35), warm-started from a
// PDHG point as the reference's polish does (src/core/solve.cpp:184-243).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lp.hpp"

namespace sk {

// The options the polish reads (src/util/options.cpp:342-393, 183).
struct IpmOptions {
  int64_t iteration_limit = 50;            // polish_iteration_limit
  double start_margin = 0.1;               // polish_start_margin
  double max_seconds = 30.0;               // polish_max_seconds
  int64_t max_factor_nonzeros = 50000000;  // polish_max_factor_nonzeros (-1: no cap)
  int32_t scaling = 1;                     // scaling
};

// A starting point in the original model's units and sign convention (the Solution's):
// x (n), row duals (m), reduced costs (n).
struct WarmPoint {
  const double* col_value = nullptr;
  const double* row_dual = nullptr;
  const double* col_dual = nullptr;
};

struct IpmOutcome {
  int32_t status = 0;  // SolveStatus numbering (device_types.hpp)
  std::string message;
  int64_t iterations = 0;
  int64_t factorizations = 0;
  int64_t floored_pivots = 0;
  double seconds = 0.0;
  // original units, Solution sign convention; empty when the status claims no point
  std::vector<double> col_value, row_dual, col_dual;
};

// The interior point on `lp`, cold when `warm` is null. Returns a library status code; the
// method's own outcome (optimal, feasible, a limit, declined, numerical failure) is in `out`.
int solve_ipm(const HostLp& lp, const IpmOptions& options, const WarmPoint* warm,
              IpmOutcome* out);

// Ruiz and Pock-Chambolle equilibration (reference src/la/scaling.cpp:111-159): the row and
// column multipliers after `ruiz_iterations` max rounds and one sum round, and the scaled
// values in the LP's CSC order.
void equilibrate(const HostLp& lp, int32_t ruiz_iterations, std::vector<double>* row,
                 std::vector<double>* column, std::vector<double>* values);

}  // namespace sk
