/*
 * FP64 CPU oracle for the restarted-PDHG LP path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's CPU algorithm,
 * used as the checker by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg. The product library (sih26_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * Ajaysingh78/SIH26 repository root).
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   ok_project, ok_cpu_primal_step ...... closed-form KATs of tests/unit/test_gpu.cpp:107-140,222-243
 *   ok_cpu_spmv_csc, ok_multiply,
 *   ok_transpose_multiply ............... 2x3 KAT test_gpu.cpp:142-177; dense matmul brute force;
 *                                         bit-exact against oracle/_ref (reference sparse.cpp)
 *   ok_set_threads ...................... a solve on 1 and on 4 threads gives the same bits
 *   ok_mt19937_raw ...................... the standard's 10,000th output (4123659995)
 *   ok_mt19937_uniform .................. through ok_spectral_norm's bit-exact match with
 *                                         oracle/_ref, whose start vector these draws are
 *   ok_cpu_dot, ok_cpu_axpy ............. KATs test_gpu.cpp:86-105; FP64 brute force
 *   ok_compare_vectors .................. KATs test_gpu.cpp:183-216
 *   ok_build_scaling .................... bit-exact against oracle/_ref (reference scaling.cpp);
 *                                         Ruiz/PC invariants on random matrices
 *   ok_spectral_norm .................... bit-exact against oracle/_ref; >= sigma_max (numpy SVD)
 *   ok_evaluate ......................... closed-form KKT terms on hand LPs
 *   ok_solve_pdhg ....................... closed-form LPs of test_pdhg.cpp:95-146; KKT-backwards
 *                                         instances with analytic optima; Netlib published optima
 *   ok_recompute_quality ................ hand-computed violations on small fabricated points
 */
#ifndef SK_ORACLE_H
#define SK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Solve statuses, numbered as sankhya::SolveStatus (include/sankhya/model.hpp:62-80). */
enum {
  OK_NOT_SOLVED = 0,
  OK_OPTIMAL = 1,
  OK_INFEASIBLE = 2,
  OK_UNBOUNDED = 3,
  OK_INFEASIBLE_OR_UNBOUNDED = 4,
  OK_FEASIBLE = 5,
  OK_ITERATION_LIMIT = 6,
  OK_TIME_LIMIT = 7,
  OK_NODE_LIMIT = 8,
  OK_NUMERICAL_ERROR = 9,
  OK_MODEL_ERROR = 10
};

/* A frozen LP in the reference's Model layout (model.hpp:131-225): CSC matrix with rows
 * sorted inside each column, two-sided rows, bounds where |v| >= 1e30 or inf is absent. */
typedef struct {
  int32_t num_rows;
  int32_t num_cols;
  int32_t sense; /* +1 minimise, -1 maximise (sense_multiplier, model.hpp:207-209) */
  double objective_offset;
  const int32_t* col_starts; /* num_cols + 1 */
  const int32_t* row_indices; /* nnz */
  const double* values;       /* nnz */
  const double* col_cost;
  const double* col_lower;
  const double* col_upper;
  const double* row_lower;
  const double* row_upper;
} ok_model;

/* pdhg.cpp:81-132 */
typedef struct {
  double primal;
  double dual;
  double gap;
  double primal_objective;
  double dual_objective;
  double absolute_primal;
  double absolute_dual;
  double gap_as_verified;
  double complementarity;
} ok_residuals;

/* The options solve_pdhg reads (pdhg.cpp:304-315; defaults options.cpp:83-89,100-106,
 * 284-290,328-334,395-418). */
typedef struct {
  double tolerance;        /* pdhg_tolerance, default 1e-4 */
  int64_t iteration_limit; /* -1 -> 1,000,000 */
  double time_limit;       /* seconds, default DBL_MAX */
  int32_t restart;         /* pdhg_restart, default 1 */
  int32_t stop_at_request; /* pdhg_stop_at_request, default 0 */
  int64_t random_seed;     /* default 0; the power iteration uses seed + 1 */
  int32_t absolute_primal; /* default 1: meets_request also needs absolute primal <= 1e-7
                              (pdhg.cpp:119-122); 0 stops on the relative KKT error alone, the
                              rule the GPU engine's FP32 mode is measured with */
} ok_options;

/* Outputs of solve_pdhg plus Solution::recompute_quality (model.cpp:228-422). Vectors are
 * caller-allocated: col_value/col_dual num_cols, row_dual/row_activity num_rows. */
typedef struct {
  int32_t status;
  int32_t converged;
  int64_t iterations;
  int64_t restarts;
  int64_t evaluations;
  double objective;
  double dual_bound;
  double solve_seconds;
  double setup_seconds; /* scaling + power iteration */
  double spectral_norm;
  ok_residuals final_residuals;
  double primal_infeasibility;
  double primal_infeasibility_scaled;
  double dual_infeasibility;
  double dual_infeasibility_scaled;
  double complementarity_violation;
  double absolute_gap;
  double relative_gap;
  double* col_value;
  double* col_dual;
  double* row_dual;
  double* row_activity;
} ok_solution;

void ok_default_options(ok_options* options);

/* types.hpp:32-39 */
int ok_is_finite_bound(double v);
/* pdhg.cpp:63-67 */
double ok_project(double value, double lower, double upper);

/* sparse.cpp:202-212, 231-234: y = A x by column scatter, skipping x_j == 0 */
void ok_multiply(int32_t rows, int32_t cols, const int32_t* col_starts,
                 const int32_t* row_indices, const double* values, const double* x, double* y);
/* Threads of the A^T x gather (sparse.cpp:216-220); default 1. Results do not depend on it. */
void ok_set_threads(int32_t threads);

/* sparse.cpp:214-229, 236-239: y = A^T x by column gather */
void ok_transpose_multiply(int32_t rows, int32_t cols, const int32_t* col_starts,
                           const int32_t* row_indices, const double* values, const double* x,
                           double* y);

/* gpu_ops_cpu.cpp:21-124 (the reference's CPU references for its GPU operators) */
void ok_cpu_axpy(int32_t n, double alpha, const double* x, double* y);
double ok_cpu_dot(int32_t n, const double* x, const double* y);
void ok_cpu_primal_step(int32_t n, double tau, const double* x, const double* cost,
                        const double* at_y, const double* col_lower, const double* col_upper,
                        double* x_next, double* extrapolated);
void ok_cpu_spmv_csc(int32_t num_rows, int32_t num_cols, const int32_t* col_starts,
                     const int32_t* row_indices, const double* values, const double* x,
                     double* y, double alpha, int32_t transpose);
/* returns passed; writes max_abs, max_rel, nonfinite */
int32_t ok_compare_vectors(int32_t n, const double* test_vec, const double* ref_vec, double tol,
                           double* max_abs, double* max_rel, int32_t* nonfinite);

/* scaling.cpp:111-159: Ruiz x ruiz_iterations then one Pock-Chambolle round. Outputs the
 * accumulated row/column multipliers, the scaled values (same CSC pattern), scaled cost and
 * bounds, and min/max |scaled entry|. */
void ok_build_scaling(const ok_model* model, const double* min_space_cost, int32_t ruiz_iterations,
                      double* row_scale, double* col_scale, double* scaled_values,
                      double* scaled_cost, double* scaled_col_lower, double* scaled_col_upper,
                      double* scaled_row_lower, double* scaled_row_upper, double* min_abs,
                      double* max_abs);

/* scaling.cpp:161-194: power iteration on A^T A with std::mt19937(seed) start */
double ok_spectral_norm(int32_t rows, int32_t cols, const int32_t* col_starts,
                        const int32_t* row_indices, const double* values, int32_t iterations,
                        uint32_t seed);

/* The MT19937 + libstdc++ uniform_real_distribution<double>(-1, 1) draws used above,
 * exposed so the restatement of the generator can be checked on its own. */
void ok_mt19937_uniform(uint32_t seed, int32_t count, double* out);
void ok_mt19937_raw(uint32_t seed, int32_t count, uint32_t* out);

/* pdhg.cpp:143-263 (evaluate). activity (m) and reduced (n) are outputs. */
ok_residuals ok_evaluate(const ok_model* model, const double* min_cost, double bound_norm,
                         double cost_norm, const double* x, const double* y, double* activity,
                         double* reduced);
int32_t ok_meets_request(const ok_residuals* r, double tolerance, int32_t absolute_primal);
int32_t ok_meets_project_standard(const ok_residuals* r);

/* pdhg.cpp:267-726 followed by model.cpp:228-422. Returns the status. */
int32_t ok_solve_pdhg(const ok_model* model, const ok_options* options, ok_solution* solution);

/* model.cpp:228-422 on an already-filled solution (col_value, row_dual, col_dual). */
void ok_recompute_quality(const ok_model* model, ok_solution* solution);

#ifdef __cplusplus
}
#endif

#endif
