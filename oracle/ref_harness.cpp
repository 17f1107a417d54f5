// Thin extern "C" driver over the reference's own scaling and sparse kernels
// (src/la/scaling.cpp, src/la/sparse.cpp), compiled from the sources where they lie by
// oracle/Makefile into oracle/_ref/. TEST INFRASTRUCTURE ONLY: tests use it to check the C
// restatement in sk_oracle.c bit for bit. Nothing here restates an algorithm; it only builds
// the reference's Model/SparseMatrix from plain arrays and calls the reference functions.
#include <cstdint>
#include <cstring>
#include <vector>

#include "sankhya/model.hpp"
#include "scaling.hpp"

namespace {

sankhya::SparseMatrix make_matrix(int32_t rows, int32_t cols, const int32_t* col_starts,
                                  const int32_t* row_indices, const double* values) {
  sankhya::SparseMatrix a(rows, cols);
  a.reserve(static_cast<std::size_t>(col_starts[cols]));
  for (int32_t j = 0; j < cols; ++j) {
    for (int32_t p = col_starts[j]; p < col_starts[j + 1]; ++p) {
      a.add_entry(row_indices[p], j, values[p]);
    }
  }
  a.finalize();
  return a;
}

}  // namespace

extern "C" {

// Returns the number of nonzeros kept by SparseMatrix::finalize (so a caller can see whether
// the reference dropped or merged anything), and the scaled outputs of build_scaling.
int32_t ref_build_scaling(int32_t rows, int32_t cols, const int32_t* col_starts,
                          const int32_t* row_indices, const double* values,
                          const double* col_lower, const double* col_upper,
                          const double* row_lower, const double* row_upper,
                          const double* min_cost, int32_t ruiz_iterations, double* row_scale,
                          double* col_scale, double* scaled_values, double* scaled_cost,
                          double* scaled_col_lower, double* scaled_col_upper,
                          double* scaled_row_lower, double* scaled_row_upper, double* min_abs,
                          double* max_abs) {
  sankhya::Model model;
  model.col_cost.assign(min_cost, min_cost + cols);
  model.col_lower.assign(col_lower, col_lower + cols);
  model.col_upper.assign(col_upper, col_upper + cols);
  model.col_type.assign(static_cast<std::size_t>(cols), sankhya::VarType::kContinuous);
  model.row_lower.assign(row_lower, row_lower + rows);
  model.row_upper.assign(row_upper, row_upper + rows);
  model.matrix = make_matrix(rows, cols, col_starts, row_indices, values);
  const std::vector<double> cost(min_cost, min_cost + cols);
  const sankhya::Scaling s = sankhya::build_scaling(model, cost, ruiz_iterations);
  std::memcpy(row_scale, s.row.data(), sizeof(double) * s.row.size());
  std::memcpy(col_scale, s.column.data(), sizeof(double) * s.column.size());
  const auto& v = s.matrix.values();
  std::memcpy(scaled_values, v.data(), sizeof(double) * v.size());
  std::memcpy(scaled_cost, s.cost.data(), sizeof(double) * s.cost.size());
  std::memcpy(scaled_col_lower, s.col_lower.data(), sizeof(double) * s.col_lower.size());
  std::memcpy(scaled_col_upper, s.col_upper.data(), sizeof(double) * s.col_upper.size());
  std::memcpy(scaled_row_lower, s.row_lower.data(), sizeof(double) * s.row_lower.size());
  std::memcpy(scaled_row_upper, s.row_upper.data(), sizeof(double) * s.row_upper.size());
  *min_abs = s.min_abs;
  *max_abs = s.max_abs;
  return static_cast<int32_t>(v.size());
}

double ref_spectral_norm(int32_t rows, int32_t cols, const int32_t* col_starts,
                         const int32_t* row_indices, const double* values, int32_t iterations,
                         uint32_t seed) {
  const sankhya::SparseMatrix a = make_matrix(rows, cols, col_starts, row_indices, values);
  return sankhya::estimate_spectral_norm(a, iterations, seed);
}

void ref_multiply(int32_t rows, int32_t cols, const int32_t* col_starts,
                  const int32_t* row_indices, const double* values, const double* x, double* y) {
  const sankhya::SparseMatrix a = make_matrix(rows, cols, col_starts, row_indices, values);
  a.multiply(x, y);
}

void ref_transpose_multiply(int32_t rows, int32_t cols, const int32_t* col_starts,
                            const int32_t* row_indices, const double* values, const double* x,
                            double* y) {
  const sankhya::SparseMatrix a = make_matrix(rows, cols, col_starts, row_indices, values);
  a.transpose_multiply(x, y);
}

}  // extern "C"
