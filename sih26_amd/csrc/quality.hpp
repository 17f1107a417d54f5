// Solution quality measures (src/core/model.cpp:228-422).
#pragma once

#include "lp.hpp"

namespace sk {

struct Quality {
  double objective = 0.0;
  double primal_infeasibility = 0.0;
  double primal_infeasibility_scaled = 0.0;
  double dual_infeasibility = 0.0;
  double dual_infeasibility_scaled = 0.0;
  double complementarity_violation = 0.0;
  double absolute_gap = 0.0;
  double relative_gap = 0.0;
};

// x: num_cols; row_dual: num_rows; col_dual: num_cols, with the Solution's sign convention.
void recompute_quality(const HostLp& lp, const double* x, const double* row_dual,
                       const double* col_dual, double dual_bound, Quality* q);

}  // namespace sk
