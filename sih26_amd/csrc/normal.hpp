// Dense normal equations of the interior-point finish on one HIP device:
//   M = A diag(theta) A^T + diag(shift),  M = L L^T,
// assembled by a HIP kernel from the CSR and CSC of A, factorised by a blocked right-looking
// Cholesky whose diagonal-block kernel floors every pivot below `floor` (the regularisation of
// the reference's sparse LDL^T, src/la/ldl.hpp:25-30), with the off-diagonal panel solves and
// the trailing updates in rocBLAS (dtrsm, dsyrk). Column-major, lower triangle.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

namespace sk {

class NormalEquations {
 public:
  NormalEquations();
  ~NormalEquations();
  NormalEquations(const NormalEquations&) = delete;
  NormalEquations& operator=(const NormalEquations&) = delete;

  // Device bytes the dense factor of an m-row system takes.
  static int64_t factor_bytes(int32_t m);

  // A (m x n) by rows and by columns; copied to the device.
  int init(int32_t m, int32_t n, const std::vector<int32_t>& row_ptr,
           const std::vector<int32_t>& row_idx, const std::vector<double>& row_val,
           const std::vector<int32_t>& col_ptr, const std::vector<int32_t>& col_idx,
           const std::vector<double>& col_val);
  // Assembles M from theta (n) and shift (m) and factorises it; `floored` receives the
  // number of pivots raised to `floor`.
  int factor(const double* theta, const double* shift, double floor, int64_t* floored);
  // rhs (m, host) <- M^-1 rhs with the current factor.
  int solve(double* rhs);

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace sk
