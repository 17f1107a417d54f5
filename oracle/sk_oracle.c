/*
 * FP64 CPU oracle for the restarted-PDHG LP path. TEST INFRASTRUCTURE ONLY: the checker used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never part of the
 * product library. See sk_oracle.h for the parity pin of every function.
 *
 * Compiled without FMA contraction (-ffp-contract=off) so that every product and sum is
 * rounded exactly where the reference's x86-64 build rounds it.
 */
#define _POSIX_C_SOURCE 200809L
#include "sk_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---- small helpers --------------------------------------------------------------------- */

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static double dmax(double a, double b) { return a < b ? b : a; } /* std::max(a, b) */
static double dmin(double a, double b) { return b < a ? b : a; } /* std::min(a, b) */

static void* xcalloc(size_t count, size_t size) {
  void* p = calloc(count ? count : 1, size);
  if (!p) abort();
  return p;
}

void ok_default_options(ok_options* o) {
  o->tolerance = 1e-4;       /* tolerances.hpp:159 kPdhgLoose */
  o->iteration_limit = -1;   /* options.cpp:100-106 */
  o->time_limit = DBL_MAX;   /* options.cpp:30, 83-89 */
  o->restart = 1;            /* options.cpp:328-334 */
  o->stop_at_request = 0;    /* options.cpp:404-418 */
  o->random_seed = 0;        /* options.cpp:284-290 */
  o->absolute_primal = 1;    /* pdhg.cpp:119-122 */
}

/* types.hpp:32-39: absent iff inf or |v| >= 1e30 */
int ok_is_finite_bound(double v) { return !(isinf(v) || fabs(v) >= 1e30); }

/* pdhg.cpp:63-67 */
double ok_project(double value, double lower, double upper) {
  if (ok_is_finite_bound(lower) && value < lower) return lower;
  if (ok_is_finite_bound(upper) && value > upper) return upper;
  return value;
}

/* ---- sparse products (sparse.cpp:202-239) ---------------------------------------------- */

void ok_multiply(int32_t rows, int32_t cols, const int32_t* cs, const int32_t* ri,
                 const double* v, const double* x, double* y) {
  for (int32_t i = 0; i < rows; ++i) y[i] = 0.0;
  for (int32_t j = 0; j < cols; ++j) {
    const double xj = 1.0 * x[j]; /* alpha = 1 */
    if (xj == 0.0) continue;
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) y[ri[p]] += v[p] * xj;
  }
}

/* Threads of the A^T x gather, the one loop the reference parallelises (sparse.cpp:216-220:
 * OpenMP, static schedule; each y[j] is written by one thread, so the result does not depend
 * on the thread count). 1 = no parallel region at all. */
static int g_threads = 1;

void ok_set_threads(int32_t threads) { g_threads = threads > 1 ? threads : 1; }

static void transpose_multiply_add(int32_t cols, const int32_t* cs, const int32_t* ri,
                                   const double* v, const double* x, double* y) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int32_t j = 0; j < cols; ++j) {
    double dot = 0.0;
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) dot += v[p] * x[ri[p]];
    y[j] += 1.0 * dot;
  }
}

void ok_transpose_multiply(int32_t rows, int32_t cols, const int32_t* cs, const int32_t* ri,
                           const double* v, const double* x, double* y) {
  (void)rows;
  for (int32_t j = 0; j < cols; ++j) y[j] = 0.0;
  transpose_multiply_add(cols, cs, ri, v, x, y);
}

/* ---- gpu_ops_cpu.cpp:21-124 ------------------------------------------------------------- */

void ok_cpu_axpy(int32_t n, double alpha, const double* x, double* y) {
  if (n <= 0 || !x || !y) return;
  for (int32_t i = 0; i < n; ++i) y[i] += alpha * x[i];
}

double ok_cpu_dot(int32_t n, const double* x, const double* y) {
  if (n <= 0 || !x || !y) return 0.0;
  double sum = 0.0;
  for (int32_t i = 0; i < n; ++i) sum += x[i] * y[i];
  return sum;
}

void ok_cpu_primal_step(int32_t n, double tau, const double* x, const double* cost,
                        const double* at_y, const double* lo, const double* hi, double* x_next,
                        double* extrapolated) {
  for (int32_t j = 0; j < n; ++j) {
    const double grad = cost[j] + at_y[j];
    const double clamped = ok_project(x[j] - tau * grad, lo[j], hi[j]);
    x_next[j] = clamped;
    extrapolated[j] = 2.0 * clamped - x[j];
  }
}

void ok_cpu_spmv_csc(int32_t num_rows, int32_t num_cols, const int32_t* cs, const int32_t* ri,
                     const double* v, const double* x, double* y, double alpha,
                     int32_t transpose) {
  if (num_rows <= 0 || num_cols <= 0 || !cs || !ri || !v || !x || !y) return;
  if (!transpose) {
    for (int32_t i = 0; i < num_rows; ++i) y[i] = 0.0;
    for (int32_t j = 0; j < num_cols; ++j) {
      const double xj = alpha * x[j];
      if (xj == 0.0) continue;
      for (int32_t p = cs[j]; p < cs[j + 1]; ++p) y[ri[p]] += v[p] * xj;
    }
  } else {
    for (int32_t j = 0; j < num_cols; ++j) {
      double dot = 0.0;
      for (int32_t p = cs[j]; p < cs[j + 1]; ++p) dot += v[p] * x[ri[p]];
      y[j] = alpha * dot;
    }
  }
}

int32_t ok_compare_vectors(int32_t n, const double* t, const double* r, double tol,
                           double* max_abs_out, double* max_rel_out, int32_t* nonfinite) {
  int32_t passed = 1;
  double max_abs = 0.0, max_rel = 0.0;
  *nonfinite = 0;
  *max_abs_out = 0.0;
  *max_rel_out = 0.0;
  if (n <= 0) return 1;
  if (!t || !r) return 0;
  for (int32_t i = 0; i < n; ++i) {
    if (!isfinite(t[i]) || !isfinite(r[i])) {
      *nonfinite = 1;
      return 0;
    }
    const double abs_diff = fabs(t[i] - r[i]);
    if (abs_diff > max_abs) max_abs = abs_diff;
    const double denom = dmax(1.0, dmax(fabs(t[i]), fabs(r[i])));
    const double rel = abs_diff / denom;
    if (rel > max_rel) max_rel = rel;
    if (abs_diff > tol && rel > tol) passed = 0;
  }
  *max_abs_out = max_abs;
  *max_rel_out = max_rel;
  return passed;
}

/* ---- scaling (scaling.cpp:14-159) ------------------------------------------------------- */

/* scaling.cpp:14-40: value * row_pass[i] * column_pass[j], multipliers accumulate */
static void apply_pass(int32_t rows, int32_t cols, const int32_t* cs, const int32_t* ri,
                       double* values, const double* row_pass, const double* col_pass,
                       double* row, double* column) {
  for (int32_t j = 0; j < cols; ++j) {
    const double cj = col_pass[j];
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) values[p] = values[p] * row_pass[ri[p]] * cj;
  }
  for (int32_t i = 0; i < rows; ++i) row[i] *= row_pass[i];
  for (int32_t j = 0; j < cols; ++j) column[j] *= col_pass[j];
}

/* scaling.cpp:44-72 (use_sum = 0) and 76-100 (use_sum = 1) */
static void equilibrate_round(int32_t rows, int32_t cols, const int32_t* cs, const int32_t* ri,
                              double* values, double* row, double* column, int use_sum) {
  double* row_acc = (double*)xcalloc((size_t)rows, sizeof(double));
  double* col_acc = (double*)xcalloc((size_t)cols, sizeof(double));
  for (int32_t j = 0; j < cols; ++j) {
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) {
      const double magnitude = fabs(values[p]);
      const int32_t i = ri[p];
      if (use_sum) {
        row_acc[i] += magnitude;
        col_acc[j] += magnitude;
      } else {
        row_acc[i] = dmax(row_acc[i], magnitude);
        col_acc[j] = dmax(col_acc[j], magnitude);
      }
    }
  }
  double* row_pass = (double*)xcalloc((size_t)rows, sizeof(double));
  double* col_pass = (double*)xcalloc((size_t)cols, sizeof(double));
  for (int32_t i = 0; i < rows; ++i) row_pass[i] = row_acc[i] > 0.0 ? 1.0 / sqrt(row_acc[i]) : 1.0;
  for (int32_t j = 0; j < cols; ++j) col_pass[j] = col_acc[j] > 0.0 ? 1.0 / sqrt(col_acc[j]) : 1.0;
  apply_pass(rows, cols, cs, ri, values, row_pass, col_pass, row, column);
  free(row_acc);
  free(col_acc);
  free(row_pass);
  free(col_pass);
}

/* scaling.cpp:105-107 */
static double scale_bound(double bound, double multiplier) {
  return ok_is_finite_bound(bound) ? bound * multiplier : bound;
}

void ok_build_scaling(const ok_model* model, const double* min_cost, int32_t ruiz_iterations,
                      double* row, double* column, double* values, double* cost,
                      double* col_lower, double* col_upper, double* row_lower,
                      double* row_upper, double* min_abs, double* max_abs) {
  const int32_t rows = model->num_rows, cols = model->num_cols;
  const int32_t* cs = model->col_starts;
  const int32_t* ri = model->row_indices;
  const int32_t nnz = cs[cols];
  for (int32_t i = 0; i < rows; ++i) row[i] = 1.0;
  for (int32_t j = 0; j < cols; ++j) column[j] = 1.0;
  memcpy(values, model->values, sizeof(double) * (size_t)nnz);

  for (int32_t pass = 0; pass < ruiz_iterations; ++pass)
    equilibrate_round(rows, cols, cs, ri, values, row, column, 0);
  equilibrate_round(rows, cols, cs, ri, values, row, column, 1);

  for (int32_t j = 0; j < cols; ++j) {
    const double dc = column[j];
    cost[j] = min_cost[j] * dc;
    col_lower[j] = scale_bound(model->col_lower[j], 1.0 / dc);
    col_upper[j] = scale_bound(model->col_upper[j], 1.0 / dc);
  }
  for (int32_t i = 0; i < rows; ++i) {
    const double dr = row[i];
    row_lower[i] = scale_bound(model->row_lower[i], dr);
    row_upper[i] = scale_bound(model->row_upper[i], dr);
  }
  *max_abs = 0.0;
  *min_abs = 0.0;
  int first = 1;
  for (int32_t p = 0; p < nnz; ++p) {
    const double magnitude = fabs(values[p]);
    *max_abs = dmax(*max_abs, magnitude);
    if (first || magnitude < *min_abs) {
      *min_abs = magnitude;
      first = 0;
    }
  }
}

/* ---- std::mt19937 and uniform_real_distribution<double> -------------------------------- */
/* MT19937 (Matsumoto & Nishimura 1998) with the C++11 [rand.eng.mers] seeding, followed by
 * libstdc++'s generate_canonical<double, 53>: two 32-bit draws g0 + g1 * 2^32, divided by
 * 2^64 in double, clamped below 1; then a + u * (b - a). */
typedef struct {
  uint32_t mt[624];
  int index;
} mt19937;

static void mt_seed(mt19937* g, uint32_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->index = 624;
}

static uint32_t mt_next(mt19937* g) {
  if (g->index >= 624) {
    for (int i = 0; i < 624; ++i) {
      const uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      g->mt[i] = v;
    }
    g->index = 0;
  }
  uint32_t y = g->mt[g->index++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

static double mt_uniform(mt19937* g, double a, double b) {
  const double r = 4294967296.0; /* max - min + 1 */
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += (double)mt_next(g) * tmp;
    tmp *= r;
  }
  double u = sum / tmp;
  if (u >= 1.0) u = nextafter(1.0, 0.0);
  return u * (b - a) + a;
}

void ok_mt19937_raw(uint32_t seed, int32_t count, uint32_t* out) {
  mt19937 g;
  mt_seed(&g, seed);
  for (int32_t k = 0; k < count; ++k) out[k] = mt_next(&g);
}

void ok_mt19937_uniform(uint32_t seed, int32_t count, double* out) {
  mt19937 g;
  mt_seed(&g, seed);
  for (int32_t k = 0; k < count; ++k) out[k] = mt_uniform(&g, -1.0, 1.0);
}

/* scaling.cpp:161-194 */
double ok_spectral_norm(int32_t rows, int32_t cols, const int32_t* cs, const int32_t* ri,
                        const double* values, int32_t iterations, uint32_t seed) {
  if (rows == 0 || cols == 0 || cs[cols] == 0) return 1.0;
  mt19937 g;
  mt_seed(&g, seed);
  double* v = (double*)xcalloc((size_t)cols, sizeof(double));
  double* av = (double*)xcalloc((size_t)rows, sizeof(double));
  for (int32_t j = 0; j < cols; ++j) v[j] = mt_uniform(&g, -1.0, 1.0);
  double norm = 0.0;
  for (int32_t step = 0; step < iterations; ++step) {
    double length = 0.0;
    for (int32_t j = 0; j < cols; ++j) length += v[j] * v[j];
    length = sqrt(length);
    if (length == 0.0) {
      free(v);
      free(av);
      return 1.0;
    }
    for (int32_t j = 0; j < cols; ++j) v[j] /= length;
    ok_multiply(rows, cols, cs, ri, values, v, av);
    ok_transpose_multiply(rows, cols, cs, ri, values, av, v);
    double next = 0.0;
    for (int32_t j = 0; j < cols; ++j) next += v[j] * v[j];
    norm = sqrt(sqrt(next));
  }
  free(v);
  free(av);
  return norm > 0.0 ? norm * 1.01 : 1.0;
}

/* ---- KKT evaluation (pdhg.cpp:81-263) --------------------------------------------------- */

double residual_worst(const ok_residuals* r) { return dmax(dmax(r->primal, r->dual), r->gap); }

int32_t ok_meets_request(const ok_residuals* r, double tolerance, int32_t absolute_primal) {
  return r->primal <= tolerance && r->dual <= tolerance && r->gap <= tolerance &&
         (!absolute_primal || r->absolute_primal <= 1e-7);
}

int32_t ok_meets_project_standard(const ok_residuals* r) {
  return r->absolute_primal <= 1e-7 && r->absolute_dual <= 1e-7 && r->gap_as_verified <= 1e-9 &&
         r->complementarity <= 1e-6;
}

ok_residuals ok_evaluate(const ok_model* model, const double* cost, double bound_norm,
                         double cost_norm, const double* x, const double* y, double* activity,
                         double* reduced) {
  const int32_t rows = model->num_rows, cols = model->num_cols;
  const double* rl = model->row_lower;
  const double* ru = model->row_upper;
  const double* cl = model->col_lower;
  const double* cu = model->col_upper;
  ok_residuals r;
  memset(&r, 0, sizeof(r));

  for (int32_t i = 0; i < rows; ++i) activity[i] = 0.0;
  if (rows > 0)
    ok_multiply(rows, cols, model->col_starts, model->row_indices, model->values, x, activity);
  double primal_violation = 0.0;
  for (int32_t i = 0; i < rows; ++i) {
    const double a = activity[i];
    double violation = 0.0;
    if (ok_is_finite_bound(rl[i])) violation = dmax(violation, rl[i] - a);
    if (ok_is_finite_bound(ru[i])) violation = dmax(violation, a - ru[i]);
    primal_violation += violation * violation;
  }
  r.absolute_primal = sqrt(primal_violation);
  r.primal = r.absolute_primal / (1.0 + bound_norm);

  for (int32_t j = 0; j < cols; ++j) reduced[j] = cost[j];
  if (rows > 0)
    transpose_multiply_add(cols, model->col_starts, model->row_indices, model->values, y,
                           reduced);
  double dual_violation = 0.0, bound_contribution = 0.0;
  for (int32_t j = 0; j < cols; ++j) {
    const double d = reduced[j];
    if (d > 0.0) {
      if (ok_is_finite_bound(cl[j])) bound_contribution += d * cl[j];
      else dual_violation += d * d;
    } else if (d < 0.0) {
      if (ok_is_finite_bound(cu[j])) bound_contribution += d * cu[j];
      else dual_violation += d * d;
    }
  }
  double support = 0.0;
  for (int32_t i = 0; i < rows; ++i) {
    const double yi = y[i];
    if (yi > 0.0) {
      if (ok_is_finite_bound(ru[i])) support += yi * ru[i];
      else dual_violation += yi * yi;
    } else if (yi < 0.0) {
      if (ok_is_finite_bound(rl[i])) support += yi * rl[i];
      else dual_violation += yi * yi;
    }
  }
  r.absolute_dual = sqrt(dual_violation);
  r.dual = r.absolute_dual / (1.0 + cost_norm);

  for (int32_t i = 0; i < rows; ++i) {
    if (rl[i] == ru[i]) continue;
    const double lower_slack = ok_is_finite_bound(rl[i]) ? activity[i] - rl[i] : INFINITY;
    const double upper_slack = ok_is_finite_bound(ru[i]) ? ru[i] - activity[i] : INFINITY;
    r.complementarity = dmax(r.complementarity, fabs(y[i]) * dmin(lower_slack, upper_slack));
  }
  for (int32_t j = 0; j < cols; ++j) {
    if (cl[j] == cu[j]) continue;
    const double lower_slack = ok_is_finite_bound(cl[j]) ? x[j] - cl[j] : INFINITY;
    const double upper_slack = ok_is_finite_bound(cu[j]) ? cu[j] - x[j] : INFINITY;
    r.complementarity =
        dmax(r.complementarity, fabs(reduced[j]) * dmin(lower_slack, upper_slack));
  }
  double primal_objective = 0.0;
  for (int32_t j = 0; j < cols; ++j) primal_objective += cost[j] * x[j];
  r.primal_objective = primal_objective;
  r.dual_objective = bound_contribution - support;
  const double absolute_gap = fabs(r.primal_objective - r.dual_objective);
  r.gap = absolute_gap / (1.0 + fabs(r.primal_objective) + fabs(r.dual_objective));
  r.gap_as_verified = absolute_gap / dmax(1.0, fabs(r.primal_objective));
  return r;
}

/* ---- Solution::recompute_quality (model.cpp:228-422), LP case --------------------------- */

void ok_recompute_quality(const ok_model* model, ok_solution* s) {
  const int32_t n = model->num_cols, m = model->num_rows;
  const int32_t* cs = model->col_starts;
  const int32_t* ri = model->row_indices;
  const double* av = model->values;
  const double* x = s->col_value;
  double* activity = s->row_activity;
  for (int32_t i = 0; i < m; ++i) activity[i] = 0.0;
  if (m > 0) ok_multiply(m, n, cs, ri, av, x, activity);

  s->primal_infeasibility = 0.0;
  s->primal_infeasibility_scaled = 0.0;
  double* row_scale = (double*)xcalloc((size_t)m, sizeof(double));
  for (int32_t i = 0; i < m; ++i) row_scale[i] = 1.0;
  for (int32_t j = 0; j < n; ++j) {
    if (x[j] == 0.0) continue;
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p)
      row_scale[ri[p]] = dmax(row_scale[ri[p]], fabs(av[p] * x[j]));
  }
#define RECORD(violation, scale)                                                          \
  do {                                                                                     \
    const double v_ = (violation);                                                         \
    if (v_ > 0.0) {                                                                        \
      s->primal_infeasibility = dmax(s->primal_infeasibility, v_);                         \
      s->primal_infeasibility_scaled =                                                     \
          dmax(s->primal_infeasibility_scaled, v_ / dmax(1.0, (scale)));                   \
    }                                                                                      \
  } while (0)
  for (int32_t j = 0; j < n; ++j) {
    if (ok_is_finite_bound(model->col_lower[j])) RECORD(model->col_lower[j] - x[j], fabs(x[j]));
    if (ok_is_finite_bound(model->col_upper[j])) RECORD(x[j] - model->col_upper[j], fabs(x[j]));
  }
  for (int32_t i = 0; i < m; ++i) {
    if (ok_is_finite_bound(model->row_lower[i]))
      RECORD(model->row_lower[i] - activity[i], row_scale[i]);
    if (ok_is_finite_bound(model->row_upper[i]))
      RECORD(activity[i] - model->row_upper[i], row_scale[i]);
  }
#undef RECORD
  /* Model::evaluate_objective, model.cpp:99-106 */
  double linear = model->objective_offset;
  for (int32_t j = 0; j < n; ++j) linear += model->col_cost[j] * x[j];
  s->objective = linear;

  s->dual_infeasibility = 0.0;
  s->dual_infeasibility_scaled = 0.0;
  s->complementarity_violation = 0.0;
  const double sense = (double)model->sense;
  double dual_norm = 0.0;
  for (int32_t i = 0; i < m; ++i) dual_norm = dmax(dual_norm, fabs(s->row_dual[i]));
#define RECORD_DUAL(violation, scale)                                                     \
  do {                                                                                     \
    const double v_ = (violation);                                                         \
    if (v_ > 0.0) {                                                                        \
      s->dual_infeasibility = dmax(s->dual_infeasibility, v_);                             \
      s->dual_infeasibility_scaled = dmax(s->dual_infeasibility_scaled, v_ / (scale));     \
    }                                                                                      \
  } while (0)
#define RECORD_COMPL(mult, slack, mscale, pscale)                                         \
  do {                                                                                     \
    const double mu_ = (mult), sl_ = (slack);                                              \
    if (!(mu_ == 0.0 || !ok_is_finite_bound(sl_))) {                                       \
      const double product_ = mu_ * sl_;                                                   \
      s->complementarity_violation = dmax(s->complementarity_violation, product_);         \
      s->dual_infeasibility_scaled = dmax(s->dual_infeasibility_scaled,                    \
                                          product_ / dmax(1.0, (mscale) * (pscale)));      \
    }                                                                                      \
  } while (0)
  for (int32_t j = 0; j < n; ++j) {
    double implied = model->col_cost[j];
    double terms = fabs(model->col_cost[j]);
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p) {
      const double term = av[p] * s->row_dual[ri[p]];
      implied -= term;
      terms = dmax(terms, fabs(term));
    }
    RECORD_DUAL(fabs(implied - s->col_dual[j]), dmax(1.0, terms));
  }
  for (int32_t j = 0; j < n; ++j) {
    const double d = sense * s->col_dual[j];
    const double xv = x[j];
    const double lo = model->col_lower[j], hi = model->col_upper[j];
    if (lo == hi) continue;
    double scale = fabs(model->col_cost[j]);
    for (int32_t p = cs[j]; p < cs[j + 1]; ++p)
      scale = dmax(scale, fabs(av[p] * s->row_dual[ri[p]]));
    scale = dmax(1.0, scale);
    if (d > 0.0 && !ok_is_finite_bound(lo)) RECORD_DUAL(d, scale);
    if (d < 0.0 && !ok_is_finite_bound(hi)) RECORD_DUAL(-d, scale);
    double distance = INFINITY;
    if (ok_is_finite_bound(lo)) distance = dmin(distance, fabs(xv - lo));
    if (ok_is_finite_bound(hi)) distance = dmin(distance, fabs(xv - hi));
    RECORD_COMPL(fabs(d), distance, scale, dmax(1.0, fabs(xv)));
  }
  for (int32_t i = 0; i < m; ++i) {
    const double yv = sense * s->row_dual[i];
    const double a = activity[i];
    const double lo = model->row_lower[i], hi = model->row_upper[i];
    if (lo == hi) continue;
    const double price_scale = dmax(1.0, dual_norm);
    if (yv > 0.0 && !ok_is_finite_bound(lo)) RECORD_DUAL(yv, price_scale);
    if (yv < 0.0 && !ok_is_finite_bound(hi)) RECORD_DUAL(-yv, price_scale);
    double distance = INFINITY;
    if (ok_is_finite_bound(lo)) distance = dmin(distance, fabs(a - lo));
    if (ok_is_finite_bound(hi)) distance = dmin(distance, fabs(a - hi));
    RECORD_COMPL(fabs(yv), distance, price_scale, row_scale[i]);
  }
#undef RECORD_DUAL
#undef RECORD_COMPL
  free(row_scale);
  s->absolute_gap = fabs(s->objective - s->dual_bound);
  s->relative_gap = s->absolute_gap / dmax(1.0, fabs(s->objective));
}

/* ---- the engine (pdhg.cpp:267-726) ------------------------------------------------------ */

int32_t ok_solve_pdhg(const ok_model* model, const ok_options* options, ok_solution* sol) {
  const double t0 = now_seconds();
  const int32_t rows = model->num_rows, cols = model->num_cols;
  const int32_t* cs = model->col_starts;
  const int32_t* ri = model->row_indices;
  const size_t n = (size_t)cols, m = (size_t)rows;
  const double sense = (double)model->sense;
  const int32_t nnz = cs[cols];

  /* pdhg.cpp:285-301: the unscaled problem in minimise space */
  double* cost = (double*)xcalloc(n, sizeof(double));
  for (size_t j = 0; j < n; ++j) cost[j] = sense * model->col_cost[j];
  double cost_square = 0.0;
  for (size_t j = 0; j < n; ++j) cost_square += cost[j] * cost[j];
  const double cost_norm = sqrt(cost_square);
  double bound_square = 0.0;
  for (size_t i = 0; i < m; ++i) {
    const double b = ok_is_finite_bound(model->row_lower[i])
                         ? model->row_lower[i]
                         : (ok_is_finite_bound(model->row_upper[i]) ? model->row_upper[i] : 0.0);
    bound_square += b * b;
  }
  const double bound_norm = sqrt(bound_square);

  /* pdhg.cpp:304-307 */
  double* srow = (double*)xcalloc(m, sizeof(double));
  double* scol = (double*)xcalloc(n, sizeof(double));
  double* sval = (double*)xcalloc((size_t)nnz, sizeof(double));
  double* scost = (double*)xcalloc(n, sizeof(double));
  double* sl = (double*)xcalloc(n, sizeof(double));
  double* su = (double*)xcalloc(n, sizeof(double));
  double* srl = (double*)xcalloc(m, sizeof(double));
  double* sru = (double*)xcalloc(m, sizeof(double));
  double min_abs = 0.0, max_abs = 0.0;
  ok_build_scaling(model, cost, 10, srow, scol, sval, scost, sl, su, srl, sru, &min_abs, &max_abs);
  const double spectral_norm =
      ok_spectral_norm(rows, cols, cs, ri, sval, 30, (uint32_t)options->random_seed + 1u);
  sol->spectral_norm = spectral_norm;
  sol->setup_seconds = now_seconds() - t0;

  const double tolerance = options->tolerance;
  const double time_limit = options->time_limit;
  const int64_t iteration_limit = options->iteration_limit < 0 ? 1000000 : options->iteration_limit;
  const int use_restarts = options->restart != 0;
  const int stop_at_request = options->stop_at_request != 0;

  double* x = (double*)xcalloc(n, sizeof(double));
  double* y = (double*)xcalloc(m, sizeof(double));
  for (size_t j = 0; j < n; ++j) x[j] = ok_project(0.0, sl[j], su[j]);
  double* x_next = (double*)xcalloc(n, sizeof(double));
  double* y_next = (double*)xcalloc(m, sizeof(double));
  double* extrapolated = (double*)xcalloc(n, sizeof(double));
  double* at_y = (double*)xcalloc(n, sizeof(double));
  double* a_x = (double*)xcalloc(m, sizeof(double));
  double* x_sum = (double*)xcalloc(n, sizeof(double));
  double* y_sum = (double*)xcalloc(m, sizeof(double));
  int64_t averaged = 0;
  double* x_restart = (double*)xcalloc(n, sizeof(double));
  double* y_restart = (double*)xcalloc(m, sizeof(double));
  memcpy(x_restart, x, n * sizeof(double));
  memcpy(y_restart, y, m * sizeof(double));
  double* activity = (double*)xcalloc(m, sizeof(double));
  double* reduced = (double*)xcalloc(n, sizeof(double));
  double* cur_x = (double*)xcalloc(n, sizeof(double));
  double* cur_y = (double*)xcalloc(m, sizeof(double));
  double* avg_x = (double*)xcalloc(n, sizeof(double));
  double* avg_y = (double*)xcalloc(m, sizeof(double));
  double* best_x = (double*)xcalloc(n, sizeof(double));
  double* best_y = (double*)xcalloc(m, sizeof(double));
  double* dx = (double*)xcalloc(n, sizeof(double));
  double* adx = (double*)xcalloc(m, sizeof(double));
  memcpy(best_x, x, n * sizeof(double)); /* pdhg.cpp:377-378 (scaled start, as the reference) */
  memcpy(best_y, y, m * sizeof(double));

  double eta = spectral_norm > 0.0 ? 1.0 / spectral_norm : 1.0;
  double omega = 1.0;
  int64_t iteration = 0, restarts = 0, last_restart = 0, evaluations = 0;
  double restart_kkt = INFINITY;
  ok_residuals best;
  memset(&best, 0, sizeof(best));
  best.primal = best.dual = best.gap = INFINITY;
  int converged = 0;

  for (;;) {
    if (iteration >= iteration_limit) break;
    if (now_seconds() - t0 > time_limit) break;
    const double tau = eta / omega;
    const double sigma = eta * omega;

    /* pdhg.cpp:392-400 */
    for (size_t j = 0; j < n; ++j) at_y[j] = 0.0;
    if (rows > 0) ok_transpose_multiply(rows, cols, cs, ri, sval, y, at_y);
    for (size_t j = 0; j < n; ++j) {
      const double gradient = scost[j] + at_y[j];
      x_next[j] = ok_project(x[j] - tau * gradient, sl[j], su[j]);
      extrapolated[j] = 2.0 * x_next[j] - x[j];
    }
    /* pdhg.cpp:403-408 */
    if (rows > 0) ok_multiply(rows, cols, cs, ri, sval, extrapolated, a_x);
    for (size_t i = 0; i < m; ++i) {
      const double v = y[i] + sigma * a_x[i];
      y_next[i] = v - sigma * ok_project(v / sigma, srl[i], sru[i]);
    }
    /* pdhg.cpp:413-438 */
    double movement = 0.0;
    for (size_t j = 0; j < n; ++j) {
      const double d = x_next[j] - x[j];
      movement += 0.5 * omega * d * d;
    }
    for (size_t i = 0; i < m; ++i) {
      const double d = y_next[i] - y[i];
      movement += 0.5 * d * d / omega;
    }
    double interaction = 0.0;
    if (rows > 0) {
      for (size_t j = 0; j < n; ++j) dx[j] = x_next[j] - x[j];
      ok_multiply(rows, cols, cs, ri, sval, dx, adx);
      for (size_t i = 0; i < m; ++i) interaction += (y_next[i] - y[i]) * adx[i];
      interaction = fabs(interaction);
    }
    /* pdhg.cpp:454-496 */
    const int no_information = interaction <= 0.0;
    const double limit = no_information ? INFINITY : movement / interaction;
    const double exponent = (double)(iteration + 1 > 2 ? iteration + 1 : 2);
    const double shrink = 1.0 - pow(exponent, -0.3);
    const double grow = 1.0 + pow(exponent, -0.6);
    const double proposed = dmin(shrink * limit, grow * eta);
    if (eta <= limit) {
      double* t = x;
      x = x_next;
      x_next = t;
      t = y;
      y = y_next;
      y_next = t;
      for (size_t j = 0; j < n; ++j) x_sum[j] += x[j];
      for (size_t i = 0; i < m; ++i) y_sum[i] += y[i];
      ++averaged;
      ++iteration;
    }
    const double eta_ceiling = 1.0e3 / dmax(spectral_norm, 1e-12);
    if (!no_information) {
      /* std::clamp(proposed, 1e-12, eta_ceiling) */
      eta = proposed < 1e-12 ? 1e-12 : (eta_ceiling < proposed ? eta_ceiling : proposed);
    }

    /* pdhg.cpp:502-545 */
    if (iteration == 0) continue;
    if (iteration % 40 != 0 && !no_information) continue;
    ++evaluations;
    for (size_t j = 0; j < n; ++j) cur_x[j] = x[j] * scol[j];
    for (size_t i = 0; i < m; ++i) cur_y[i] = y[i] * srow[i];
    const ok_residuals current =
        ok_evaluate(model, cost, bound_norm, cost_norm, cur_x, cur_y, activity, reduced);
    ok_residuals chosen = current;
    int chosen_is_average = 0;
    ok_residuals average;
    if (averaged > 0) {
      const double count = (double)averaged;
      for (size_t j = 0; j < n; ++j) avg_x[j] = (x_sum[j] / count) * scol[j];
      for (size_t i = 0; i < m; ++i) avg_y[i] = (y_sum[i] / count) * srow[i];
      average = ok_evaluate(model, cost, bound_norm, cost_norm, avg_x, avg_y, activity, reduced);
      if (residual_worst(&average) < residual_worst(&current)) {
        chosen = average;
        chosen_is_average = 1;
      }
    }
    const double* chosen_x = chosen_is_average ? avg_x : cur_x;
    const double* chosen_y = chosen_is_average ? avg_y : cur_y;
    if (residual_worst(&chosen) < residual_worst(&best)) {
      best = chosen;
      memcpy(best_x, chosen_x, n * sizeof(double));
      memcpy(best_y, chosen_y, m * sizeof(double));
    }
    /* pdhg.cpp:585-597 */
    if (ok_meets_request(&chosen, tolerance, options->absolute_primal) &&
        (stop_at_request || ok_meets_project_standard(&chosen))) {
      best = chosen;
      memcpy(best_x, chosen_x, n * sizeof(double));
      memcpy(best_y, chosen_y, m * sizeof(double));
      converged = 1;
      break;
    }
    /* pdhg.cpp:599-643 */
    if (use_restarts) {
      const double kkt = residual_worst(&chosen);
      const int64_t since = iteration - last_restart;
      const int sufficient = kkt <= 0.2 * restart_kkt;
      const int64_t artificial_floor = (int64_t)(0.36 * (double)iteration);
      const int artificial = since >= (artificial_floor > 40 ? artificial_floor : 40);
      if (sufficient || artificial) {
        double dx2 = 0.0, dy2 = 0.0;
        for (size_t j = 0; j < n; ++j) {
          const double d = x[j] - x_restart[j];
          dx2 += d * d;
        }
        for (size_t i = 0; i < m; ++i) {
          const double d = y[i] - y_restart[i];
          dy2 += d * d;
        }
        const double dx_norm = sqrt(dx2), dy_norm = sqrt(dy2);
        if (dx_norm > 1e-12 && dy_norm > 1e-12) {
          const double theta = 0.5;
          omega = exp(theta * log(dy_norm / dx_norm) + (1.0 - theta) * log(omega));
          omega = omega < 1e-6 ? 1e-6 : (1e6 < omega ? 1e6 : omega);
        }
        memset(x_sum, 0, n * sizeof(double));
        memset(y_sum, 0, m * sizeof(double));
        averaged = 0;
        memcpy(x_restart, x, n * sizeof(double));
        memcpy(y_restart, y, m * sizeof(double));
        restart_kkt = kkt;
        last_restart = iteration;
        ++restarts;
      }
    }
  }

  /* pdhg.cpp:647-716 */
  for (size_t j = 0; j < n; ++j) sol->col_value[j] = best_x[j];
  const ok_residuals final_r =
      ok_evaluate(model, cost, bound_norm, cost_norm, best_x, best_y, activity, reduced);
  for (size_t j = 0; j < n; ++j) sol->col_dual[j] = sense * reduced[j];
  for (size_t i = 0; i < m; ++i) sol->row_dual[i] = sense * (-best_y[i]);
  sol->iterations = iteration;
  sol->restarts = restarts;
  sol->evaluations = evaluations;
  sol->converged = converged;
  sol->final_residuals = final_r;
  const double elapsed = now_seconds() - t0;
  const int verifiable = converged && ok_meets_project_standard(&final_r);
  if (verifiable) sol->status = OK_OPTIMAL;
  else if (converged) sol->status = OK_FEASIBLE;
  else sol->status = elapsed > time_limit ? OK_TIME_LIMIT : OK_ITERATION_LIMIT;
  if (verifiable) sol->dual_bound = sense * final_r.dual_objective + model->objective_offset;
  else sol->dual_bound = model->sense < 0 ? INFINITY : -INFINITY;
  ok_recompute_quality(model, sol);
  sol->solve_seconds = now_seconds() - t0;

  free(cost); free(srow); free(scol); free(sval); free(scost); free(sl); free(su); free(srl);
  free(sru); free(x); free(y); free(x_next); free(y_next); free(extrapolated); free(at_y);
  free(a_x); free(x_sum); free(y_sum); free(x_restart); free(y_restart); free(activity);
  free(reduced); free(cur_x); free(cur_y); free(avg_x); free(avg_y); free(best_x); free(best_y);
  free(dx); free(adx);
  return sol->status;
}
