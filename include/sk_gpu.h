/*
 * sk_gpu.h - C ABI of the MI355X restarted-PDHG LP engine (library: libsk_gpu.so).
 *
 * Conventions follow the reference's C API (include/sankhya/sankhya.h:16-28): plain C types,
 * opaque caller-owned handles freed with their _free or _destroy call, int status codes, a thread-local
 * last-error string, and no exception or C++ type crosses the boundary. Handles are not
 * internally synchronised; distinct handles may be used from distinct threads.
 *
 * Status codes are numbered as sankhya::gpu::GpuStatus (include/sankhya/gpu_runtime.hpp:19-30)
 * with one addition, SK_GPU_INVALID_ARGUMENT. Solve statuses are numbered as
 * sankhya::SolveStatus (include/sankhya/model.hpp:62-80).
 */
#ifndef SK_GPU_H
#define SK_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (gpu_runtime.hpp:19-30) ---------------------------------------------- */
#define SK_GPU_SUCCESS 0
#define SK_GPU_UNAVAILABLE 1        /* kCudaUnavailable: the HIP runtime cannot be used */
#define SK_GPU_NO_DEVICE 2          /* kNoCompatibleDevice */
#define SK_GPU_INIT_FAILED 3        /* kInitFailed */
#define SK_GPU_OUT_OF_MEMORY 4      /* kOutOfMemory */
#define SK_GPU_LAUNCH_FAILED 5      /* kKernelLaunchFailed */
#define SK_GPU_COPY_FAILED 6        /* kCopyFailed */
#define SK_GPU_SYNC_FAILED 7        /* kSyncFailed */
#define SK_GPU_INVALID_STATE 8      /* kInvalidDeviceState */
#define SK_GPU_NON_FINITE 9         /* kNonFiniteOutput */
#define SK_GPU_INVALID_ARGUMENT 10  /* malformed model or argument; no reference counterpart */

/* ---- solve statuses (model.hpp:62-80) ---------------------------------------------------- */
#define SK_SOLVE_NOT_SOLVED 0
#define SK_SOLVE_OPTIMAL 1
#define SK_SOLVE_INFEASIBLE 2
#define SK_SOLVE_FEASIBLE 5
#define SK_SOLVE_ITERATION_LIMIT 6
#define SK_SOLVE_TIME_LIMIT 7
#define SK_SOLVE_NUMERICAL_ERROR 9
#define SK_SOLVE_MODEL_ERROR 10

/* ---- runtime (replaces include/sankhya/gpu_runtime.hpp:32-105) --------------------------- */

typedef struct {
  int32_t device_id;
  char name[256];
  uint64_t total_memory_bytes;
  int32_t gcn_arch_major; /* replaces compute_capability_major */
  int32_t gcn_arch_minor; /* replaces compute_capability_minor */
  int32_t multiprocessor_count;
  int32_t warp_size; /* 64 on CDNA */
  int32_t max_threads_per_block;
  char gcn_arch_name[64];
} sk_gpu_device_info;

/* gpu_runtime.hpp:32-56 to_string(GpuStatus) */
const char* sk_gpu_status_string(int32_t status);
/* gpu_runtime.hpp:70-71 is_cuda_available: 1 when a HIP device can be used */
int32_t sk_gpu_is_available(void);
/* gpu_runtime.hpp:73-74 get_gpu_device_count */
int32_t sk_gpu_device_count(void);
/* gpu_runtime.hpp:76-77 get_device_info */
int32_t sk_gpu_get_device_info(int32_t device_id, sk_gpu_device_info* out_info);
/* gpu_runtime.hpp:79-80 set_device */
int32_t sk_gpu_set_device(int32_t device_id);
/* gpu_runtime.hpp:82-83 synchronize_device */
int32_t sk_gpu_synchronize(void);
/* gpu_runtime.hpp:85-86 get_last_error_string (thread-local; valid until the next call) */
const char* sk_gpu_last_error(void);
/* Fault injection, for tests of the failure paths from the host (no counterpart in the
 * reference, whose tests force failures only through a missing device and the limits). kind 1:
 * the n-th device allocation an engine makes from now on fails with SK_GPU_OUT_OF_MEMORY; kind
 * 2: the n-th kernel launch an engine makes from now on fails with SK_GPU_LAUNCH_FAILED (the
 * launch is not made, so no kernel runs in a bad state); kind 3: instance n of the next batch
 * upload gets a NaN cost in its first column, so its first step is non-finite and it ends with
 * SK_SOLVE_NUMERICAL_ERROR while the other instances run on; kind 4: the next n triangular
 * solves of the interior point give up at their first wait (a bound of zero polls), so each is
 * redone by the rocBLAS fallback; kind 0 clears every injection. Process-wide; kinds 1-3 fire
 * once. */
int32_t sk_fault_inject(int32_t kind, int64_t n);
/* How many times injections of `kind` (1-4) have taken effect in this process (no counterpart
 * in the reference; for the same tests). */
int64_t sk_fault_fired(int32_t kind);
/* gpu_runtime.hpp:90-105 allocate/free/copy/memset */
int32_t sk_gpu_malloc(void** dev_ptr, size_t bytes);
int32_t sk_gpu_free(void* dev_ptr);
int32_t sk_gpu_copy_h2d(void* dev_dst, const void* host_src, size_t bytes);
int32_t sk_gpu_copy_d2h(void* host_dst, const void* dev_src, size_t bytes);
int32_t sk_gpu_memset(void* dev_ptr, int32_t value, size_t bytes);

/* ---- device operators (replace include/sankhya/gpu_ops.hpp:50-62; launch_gpu_*) ----------
 * Device pointers, FP64, asynchronous on the library's stream unless stated; call
 * sk_gpu_synchronize before reading results. */

/* gpu_ops.hpp:52-53 launch_gpu_axpy: y += alpha x */
int32_t sk_gpu_axpy(int32_t n, double alpha, const double* d_x, double* d_y);
/* gpu_ops.hpp:55-57 launch_gpu_dot: synchronous, deterministic two-stage FP64 reduction */
int32_t sk_gpu_dot(int32_t n, const double* d_x, const double* d_y, double* host_result);
/* No counterpart in the reference. sk_gpu_dot for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same two-stage reduction, partition and bits as sk_gpu_dot, but its partial sums live in a
 * fixed 1 MiB device array of the library (one per device, held until the process ends) and
 * it calls no hipMalloc, hipFree or allocating routine. n is at most 268,435,456; a larger n
 * returns SK_GPU_INVALID_ARGUMENT. Calls that overlap on one device race on that array. */
int32_t sk_gpu_dot_noalloc(int32_t n, const double* d_x, const double* d_y, double* host_result);
/* gpu_ops.hpp:59-62 launch_gpu_pdhg_primal_step:
 *   x_next = proj_[lo,hi](x - tau (cost + at_y)), extrapolated = 2 x_next - x */
int32_t sk_gpu_pdhg_primal_step(int32_t n, double tau, const double* d_x, const double* d_cost,
                                const double* d_at_y, const double* d_col_lower,
                                const double* d_col_upper, double* d_x_next,
                                double* d_extrapolated);
/* gpu_ops.hpp:85-91 execute_pdhg_primal_step, the validated dispatcher (gpu_ops_cpu.cpp:126-207),
 * on HOST arrays (synchronous): device buffers, copies in, the primal step, a device
 * synchronisation, copies out, then the reference's check that every output is finite. Each
 * failure class returns its status as the reference's does: SK_GPU_UNAVAILABLE without a
 * device, SK_GPU_OUT_OF_MEMORY, SK_GPU_COPY_FAILED, SK_GPU_LAUNCH_FAILED, SK_GPU_SYNC_FAILED,
 * SK_GPU_NON_FINITE. Unlike the reference it does not then compute the step on the CPU: this
 * library has no CPU path, so on any status but success the outputs are unspecified and a
 * caller that wants the fallback runs its own CPU step (the reference keeps
 * cpu_pdhg_primal_step, gpu_ops_cpu.cpp:37-49, for exactly that). */
int32_t sk_gpu_execute_pdhg_primal_step(int32_t n, double tau, const double* x, const double* cost,
                                        const double* at_y, const double* col_lower,
                                        const double* col_upper, double* x_next,
                                        double* extrapolated);
/* gpu_ops.hpp:64-67 launch_gpu_spmv_csc on a CSC matrix.
 *   transpose = 1: y = alpha A^T x, one wavefront per column (deterministic).
 *   transpose = 0: y = alpha A x by column scatter with FP64 atomics, as the reference's
 *   k_spmv_csc_scatter (cuda_kernels.cu:96-112); summation order is not deterministic. The
 *   engine never uses this form: it keeps a CSR of A instead (see sk_gpu_spmv_csr). */
int32_t sk_gpu_spmv_csc(int32_t num_rows, int32_t num_cols, const int32_t* d_col_starts,
                        const int32_t* d_row_indices, const double* d_values, const double* d_x,
                        double* d_y, double alpha, int32_t transpose);
/* Deterministic y = alpha A x on a CSR matrix, one wavefront per row. New (the reference has
 * no CSR operator); equals sk_gpu_spmv_csc(transpose=1) applied to the CSC of A^T. */
int32_t sk_gpu_spmv_csr(int32_t num_rows, int32_t num_cols, const int32_t* d_row_starts,
                        const int32_t* d_col_indices, const double* d_values, const double* d_x,
                        double* d_y, double alpha);

/* gpu_ops.hpp:73-78, 85-87 NumericalComparisonResult / compare_vectors (host function):
 * fails on any non-finite entry, or where |a-b| > tol and |a-b| / max(1,|a|,|b|) > tol. */
typedef struct {
  int32_t passed;
  int32_t contains_nan_or_inf;
  double max_abs_diff;
  double max_rel_diff;
} sk_gpu_comparison;
int32_t sk_gpu_compare_vectors(int32_t n, const double* test_vec, const double* ref_vec,
                               double tol, sk_gpu_comparison* out);

/* ---- LP models ---------------------------------------------------------------------------
 * An LP in the reference Model layout (model.hpp:131-225): CSC matrix (rows ascending inside
 * each column, no duplicates), cost, column bounds, two-sided row bounds; a bound is absent
 * when infinite or |v| >= 1e30 (types.hpp:32-39). sense: +1 minimise, -1 maximise. */
typedef struct sk_lp sk_lp;

int32_t sk_lp_create(int32_t num_rows, int32_t num_cols, int32_t sense, double objective_offset,
                     const int32_t* col_starts, const int32_t* row_indices,
                     const double* values, const double* col_cost, const double* col_lower,
                     const double* col_upper, const double* row_lower, const double* row_upper,
                     sk_lp** out);
/* Seeded generator (no counterpart in the library; the nearest in the reference is the Python
 * script bench/runners/generate_large_lp.py:49-206). config 1..5 are the five workloads: 1 = 500x1000 blending LP (seed =
 * instance), 2 = 200k x 500k power-law LP with 2.5M nnz (seed = instance), 3 = member
 * `instance` of the 256-LP batch, 4 = serving request `instance`, 5 = 2M x 5M power-law LP
 * with 25M nnz (seed = instance). */
int32_t sk_lp_generate(int32_t config, uint64_t instance, sk_lp** out);

typedef struct {
  uint64_t seed;
  int32_t num_rows;
  int32_t num_cols;
  int32_t num_equality_rows;
  int32_t degree_kind; /* 0 fixed degree_lo, 1 uniform [lo, hi], 2 power law on [lo, hi] */
  int32_t degree_lo;
  int32_t degree_hi;
  double power_exponent;
  int64_t target_nnz; /* power law: trim/pad to exactly this many nonzeros (0: no change) */
  int32_t value_kind; /* 0 |a| uniform on [lo, hi], 1 log-uniform */
  double value_lo;
  double value_hi;
  double negative_probability;
  double upper_lo, upper_hi;
  double slack_lo, slack_hi;
  double cost_lo, cost_hi;
  int32_t fix_empty_rows;
} sk_gen_params;
/* The generator's parameters for a workload member, and a generation from any parameters. */
int32_t sk_gen_preset(int32_t config, uint64_t instance, sk_gen_params* out);
int32_t sk_lp_generate_params(const sk_gen_params* params, sk_lp** out);

/* The Model's sizes and arrays (model.hpp:131-225). */
int32_t sk_lp_dims(const sk_lp* lp, int32_t* num_rows, int32_t* num_cols, int64_t* nnz);
/* Copies the model out; every pointer may be NULL to skip that array. */
int32_t sk_lp_export(const sk_lp* lp, int32_t* sense, double* objective_offset,
                     int32_t* col_starts, int32_t* row_indices, double* values,
                     double* col_cost, double* col_lower, double* col_upper, double* row_lower,
                     double* row_upper);
void sk_lp_free(sk_lp* lp);
/* read_mps (include/sankhya/io.hpp:45-46; src/io/mps_reader.cpp, mps_bounds.cpp): an MPS file
 * (free format; fixed columns when the free reading fails) as an LP. The first N row is the
 * objective (its RHS is the negated constant), later N rows are dropped; RANGES and BOUNDS
 * follow the reference (first vector only; UP < 0 with no explicit lower bound sets the lower
 * bound to -inf); duplicate entries, non-finite coefficients and quadratic sections are
 * refused. On failure returns SK_GPU_INVALID_ARGUMENT and sk_gpu_last_error() gives
 * "path:line: reason". integer_columns (may be NULL) receives the number of columns the file
 * marks integer; an LP solve relaxes them. */
int32_t sk_lp_read_mps(const char* path, sk_lp** out, int32_t* integer_columns);
/* Free-format MPS of the LP, which read_mps (above, and the reference's, io.hpp:45-46) reads
 * back to the same arrays: values with 17 significant digits, names from the file the LP was
 * read from when they are blank-free and distinct, else R<i> / C<j>; a free row becomes an L
 * row with right-hand side 1e30. Entries below 1e-11 in magnitude do not survive (readers drop
 * them), and a two-sided row whose bound difference rounds comes back within one rounding of
 * its range. No counterpart in the reference library: its generator writes MPS from Python
 * (bench/runners/generate_large_lp.py:49-206) for the solver's CLI; this lets the reference's
 * binary be run on the generated workloads. */
int32_t sk_lp_write_mps(const sk_lp* lp, const char* path);
/* Instance cache (no counterpart in the reference): the LP's arrays in host byte order after
 * an 8-byte magic, with an FNV-1a checksum. sk_lp_load checks the size, the checksum and the
 * model, and refuses a file that fails any of them with SK_GPU_INVALID_ARGUMENT. */
int32_t sk_lp_save(const sk_lp* lp, const char* path);
int32_t sk_lp_load(const char* path, sk_lp** out);
/* The name of row (kind 0) or column (kind 1) `index` of an LP read from a file, or of the
 * model (kind 2, index ignored); NULL otherwise (Model::row_names, col_names, name:
 * model.hpp:135-173). Valid until sk_lp_free. */
const char* sk_lp_name(const sk_lp* lp, int32_t kind, int32_t index);

/* ---- PDHG options (the options solve_pdhg reads: pdhg.cpp:304-315) ------------------------ */
typedef struct {
  double tolerance;          /* pdhg_tolerance, default 1e-4 (tolerances.hpp:159) */
  int64_t iteration_limit;   /* iteration_limit, -1 = 1,000,000 (pdhg.cpp:312-313) */
  double time_limit;         /* time_limit in seconds, default DBL_MAX */
  int32_t restart;           /* pdhg_restart, default 1 */
  int32_t stop_at_request;   /* pdhg_stop_at_request, default 0 */
  int32_t absolute_primal;   /* pdhg_absolute_primal, default 1: the stop test also needs
                                absolute primal <= 1e-7 (pdhg.cpp:119-122); 0 stops on the
                                relative KKT error alone */
  int64_t random_seed;       /* random_seed, default 0 */
  int32_t precision;         /* pdhg_precision: 0 "fp32" (FP32 iterates, FP64 evaluation,
                                default), 1 "fp64" */
  int32_t steps_per_block;   /* PDHG steps per captured graph, default 40 */
  int32_t use_graph;         /* 1: hipGraph-captured step blocks (default) */
  int32_t spmv_stage;        /* SpMV staging order: -1 default (1), 0 CSR order, 1 sorted by gather index */
  int32_t tile_nnz;          /* row-block capacity in nonzeros: 0 the default (2,560; 1,536
                                for a batch, other than a sub-batch of a split one, whose
                                CSR(A^T) tiles at 2,560 would not fill one round of resident
                                workgroups, 1,024 below a quarter of a round), -1
                                picked by the autotuner at sk_batch_upload (each of 1,536,
                                2,048, 2,560, 3,072 and 4,096 is uploaded and the two SpMV step
                                kernels timed; the fastest stays), at most 4,096 */
  int32_t compute_quality;   /* 1: fill the recompute_quality measures (default) */
  int32_t eval_stride;       /* an evaluation stage follows every this many steps inside a
                                captured block, so an instance due for evaluation waits at most
                                this many steps (default 20) */
  int32_t streams;           /* a batch runs as this many sub-batches side by side, each on
                                its own stream: 0 automatic (3 for batches of 2^20 nonzeros and
                                more, else 1), 1 one block-diagonal launch per kernel */
  int32_t column_order;      /* device column order: 0 automatic (default: an instance of at
                                least 65,536 columns puts its 2 MiB of columns with the most
                                nonzeros first, so the most-gathered part of x stays in L2),
                                1 every instance so, 2 the LP's own order. Answers are returned
                                in the LP's order either way. */
  int32_t step_fusion;       /* bit mask, 0 (default) both on, each bit for bit the same
                                answers. Bit 0 (1) off: the resident solve, where every instance
                                has at most 4 SpMV tiles and 2 column chunks, one workgroup per
                                instance runs many steps and evaluations in one launch. Bit 1
                                (2) off: each step's decision made by the next primal launch
                                instead of a launch of its own. */
  int32_t backend;           /* backend (options.cpp:228-236): SK_BACKEND_AUTO (default), _CPU,
                                _GPU or _HYBRID. Read by sk_select_backend; a solve, batch or
                                server of this library refuses _CPU (the CPU engine is the
                                reference's) and _HYBRID (unsupported, as in the reference) */
  int32_t gpu;               /* the legacy gpu flag (options.cpp:220-227), default 0 */
  int32_t presolve;          /* presolve (options.cpp:190-201): default 0 here, since the
                                reference presolves in solve(), around solve_pdhg
                                (solve.cpp:423-443); 1 makes sk_pdhg_solve do the same:
                                sk_presolve, the engine on the reduced LP, sk_postsolve, and
                                the answer measured on the original LP */
  int32_t column_slices;     /* no counterpart in the reference: A x-bar of a solve or batch of
                                one LP as 8 partial products over column slices, one per XCD,
                                so that the x-bar entries each gathers touch stay in that XCD's
                                L2; 0 automatic (default: an LP of at least 2^22 columns), 1 on,
                                2 off. A sliced LP keeps its own column order on the device, and
                                its row sums add the slices' partials in slice order */
} sk_pdhg_options;

/* The defaults listed above (options.cpp:83-106, 220-236, 284-290, 328-334, 395-418). */
void sk_pdhg_default_options(sk_pdhg_options* options);
/* String-keyed setter over the reference's option names (src/util/options.cpp:78-455):
 * algorithm (must be "pdhg"), backend (auto, cpu, gpu or cuda, hybrid, as parse_backend_mode,
 * adaptive.hpp:43-62), gpu, pdhg_tolerance, iteration_limit,
 * time_limit, pdhg_restart, pdhg_stop_at_request, random_seed, pdhg_polish (validated; as in
 * the reference, the polish is a separate step after the first-order solve: sk_pdhg_polish),
 * plus pdhg_precision ("fp32"|"fp64"),
 * pdhg_absolute_primal and pdhg_steps_per_block. Unknown keys and malformed values are
 * rejected with SK_GPU_INVALID_ARGUMENT, as options.cpp:481-565 rejects them. */
int32_t sk_pdhg_set_option(sk_pdhg_options* options, const char* key, const char* value);

/* ---- backend selection (include/sankhya/adaptive.hpp:22-121; src/core/adaptive.cpp:32-232) --
 * The reference's adaptive CPU/GPU decision, restated for callers of this library: the same
 * modes, the same order of tests and thresholds, the same validity rules (an explicit GPU
 * request without a device is invalid, never a silent fallback; the legacy gpu flag falls
 * back to the CPU; hybrid is unsupported). A CPU decision means the reference's CPU engine:
 * this library only runs the GPU engine. */
#define SK_BACKEND_AUTO 0
#define SK_BACKEND_CPU 1
#define SK_BACKEND_GPU 2
#define SK_BACKEND_HYBRID 3

/* ProblemCharacteristics (adaptive.hpp:64-83) */
typedef struct {
  char problem_class[8]; /* "LP", "MILP", "QP" or "MIQP" */
  int32_t num_rows;
  int32_t num_cols;
  int64_t num_nonzeros;
  double density;        /* nnz / (rows cols) */
  double sparsity_ratio; /* 1 - density */
  int32_t num_integer_cols;
  int32_t has_quadratic_objective;
  double estimated_matvec_flops;  /* 2 nnz */
  double estimated_memory_bytes;  /* CSC (int32 index + FP64 value) + 4 FP64 vectors */
  int32_t gpu_available;
  int32_t gpu_device_count;
  uint64_t gpu_memory_bytes;
  char gpu_name[256];
} sk_problem_characteristics;

/* AdaptiveConfig (adaptive.hpp:85-93) */
typedef struct {
  int64_t min_gpu_nonzeros;       /* adaptive_min_nnz, default 50,000 */
  int64_t min_gpu_dimension;      /* adaptive_min_dim, default 1,000 */
  double min_gpu_density;         /* adaptive_min_density, default 1e-6 (read, as in the
                                     reference, but no test uses it) */
  double max_gpu_memory_headroom; /* 0.8 */
  int32_t explain_decision;       /* adaptive_explain, default 1 */
  int32_t reserved;
} sk_adaptive_config;

/* BackendDecision (adaptive.hpp:96-104) and its format_explanation (adaptive.cpp:15-29) */
typedef struct {
  int32_t selected_backend; /* SK_BACKEND_CPU, _GPU or _HYBRID */
  int32_t requested_mode;   /* SK_BACKEND_* */
  int32_t is_valid;         /* 0: the requested backend cannot run; the reference then returns
                               kNotSolved with primary_reason as the message (solve.cpp:359-366) */
  int32_t detail_count;
  char primary_reason[512];
  char explanation[2048];
} sk_backend_decision;

void sk_adaptive_default_config(sk_adaptive_config* config);
/* load_adaptive_config's keys (options.cpp:237-264): adaptive_min_nnz, adaptive_min_dim,
 * adaptive_min_density, adaptive_explain, with the registry's bounds. */
int32_t sk_adaptive_set_option(sk_adaptive_config* config, const char* key, const char* value);
/* analyze_problem (adaptive.cpp:32-81) of an LP; integer_columns and has_quadratic_objective
 * are what the model's source says (sk_lp_read_mps reports the integer columns), the device
 * fields are queried from the HIP runtime (device 0). */
int32_t sk_analyze_problem(const sk_lp* lp, int32_t integer_columns,
                           int32_t has_quadratic_objective, sk_problem_characteristics* out);
/* select_backend (adaptive.cpp:98-232), reading options->backend and options->gpu. */
int32_t sk_select_backend(const sk_problem_characteristics* chars, const sk_pdhg_options* options,
                          const sk_adaptive_config* config, sk_backend_decision* out);

/* Residuals (pdhg.cpp:81-132) of the reported point and Solution quality measures
 * (model.cpp:228-422). */
typedef struct {
  int32_t status;
  int32_t converged;
  int64_t iterations;
  int64_t restarts;
  int64_t evaluations;
  int64_t attempts;
  double objective;
  double dual_bound;
  double spectral_norm;
  double solve_seconds;
  double rel_primal;
  double rel_dual;
  double rel_gap;
  double abs_primal;
  double abs_dual;
  double gap_as_verified;
  double complementarity;
  double primal_objective;
  double dual_objective;
  double primal_infeasibility;
  double primal_infeasibility_scaled;
  double dual_infeasibility;
  double dual_infeasibility_scaled;
  double complementarity_violation;
  double absolute_gap;
  double relative_gap;
  char algorithm[16]; /* "pdhg-gpu" (verify.cpp:86-89 audits on "gpu") */
  char message[512];
  int64_t polish_iterations; /* interior-point iterations of a polish (solve.cpp:225-242) */
  double primal_weight;      /* the primal weight omega and step size eta the run ended with,
                                to continue it (sk_batch_warm_start, sk_pdhg_refine) */
  double step_size;
} sk_pdhg_result;

/* The drop-in for pdhg::solve_pdhg (include/sankhya/pdhg.hpp:15; src/pdhg/pdhg.cpp:267-726)
 * on one LP. Output arrays are caller-allocated (NULL skips one): col_value and col_dual
 * num_cols, row_dual and row_activity num_rows. Signs as the reference Solution:
 * col_dual = sense * (c_min + A^T y), row_dual = -sense * y (pdhg.cpp:652-661). */
int32_t sk_pdhg_solve(const sk_lp* lp, const sk_pdhg_options* options, sk_pdhg_result* result,
                      double* col_value, double* row_dual, double* col_dual,
                      double* row_activity);

/* The finish of an LP the interior point declines (its factor above the cap): the plain FP64
 * engine, warm-started from a first answer (for example the FP32 engine's at 1e-4), run to the
 * options' tolerance and stop rule (pdhg_precision is taken as fp64). `first` gives the primal
 * weight and step size to continue with; first_col_value and first_row_dual are its arrays in
 * the Solution's signs. The result's iterations are both phases together, polish_iterations
 * the second, and algorithm reads "pdhg-gpu+fp64". No counterpart in the reference, whose
 * PDHG always starts cold (pdhg.cpp:327-334) and whose polish declines such an LP
 * (solve.cpp:184-243). */
int32_t sk_pdhg_refine(const sk_lp* lp, const sk_pdhg_options* options,
                       const sk_pdhg_result* first, const double* first_col_value,
                       const double* first_row_dual, sk_pdhg_result* result, double* col_value,
                       double* row_dual, double* col_dual, double* row_activity);

/* writer.cpp:119-227 write_solution: the .sol file the reference's independent checker
 * (tools/verify_solution.py) reads. Header lines, then `columns` (name value reduced_cost
 * basis_status) and `rows` (name activity dual basis_status) blocks, every number with 17
 * significant digits; a status that claims no point (model.hpp:99-112) writes the header only.
 * Names default to C<j> and R<i>; col_names / row_names may be NULL. The arrays are those
 * sk_pdhg_solve fills (NULL writes zeros). */
int32_t sk_write_solution(const char* path, const sk_lp* lp, const sk_pdhg_result* result,
                          const double* col_value, const double* row_dual,
                          const double* col_dual, const double* row_activity,
                          const char* const* col_names, const char* const* row_names,
                          const char* model_name);

/* ---- interior-point finish (src/ipm/ipm.cpp; src/core/solve.cpp:184-243) ------------------ */
/* The options the polish reads (src/util/options.cpp:342-393 and 183). */
typedef struct {
  int64_t iteration_limit;     /* polish_iteration_limit, default 50 */
  double start_margin;         /* polish_start_margin, default 0.1 */
  double max_seconds;          /* polish_max_seconds, default 30 */
  int64_t max_factor_nonzeros; /* polish_max_factor_nonzeros, default 5e7, -1 = no cap; the
                                  factor here is dense: m (m + 1) / 2 nonzeros */
  int32_t scaling;             /* scaling, default 1 */
  int32_t reserved;
} sk_ipm_options;

/* The defaults listed above (options.cpp:342-393, 183). */
void sk_ipm_default_options(sk_ipm_options* options);
/* ipm::solve_ipm (include/sankhya/ipm.hpp:17, 33-34): Mehrotra's predictor-corrector on the
 * normal equations, assembled and factorised (dense Cholesky, pivots floored at 1e-10) on the
 * current HIP device. Started from the warm point when all three warm arrays are given (the
 * WarmStart of ipm.hpp:21-31, in the original units and the Solution's signs), else cold.
 * `result` gets the status (optimal / feasible / iteration_limit / time_limit / not_solved
 * when declined / numerical_error), iterations, objective, dual bound and quality measures;
 * the arrays are filled when the status claims a point. algorithm = "ipm". */
int32_t sk_ipm_solve(const sk_lp* lp, const sk_ipm_options* options,
                     const double* warm_col_value, const double* warm_row_dual,
                     const double* warm_col_dual, sk_pdhg_result* result, double* col_value,
                     double* row_dual, double* col_dual, double* row_activity);
/* polish_with_the_interior_point (src/core/solve.cpp:184-243): `first` (a PDHG answer and its
 * arrays) is handed to the interior point as its starting point unless it is already optimal
 * or claims no point. The polished answer replaces it only when it is optimal, or feasible
 * with smaller scaled violations; then iterations are the sum of both phases,
 * polish_iterations the second, and algorithm reads "<first>+ipm". Otherwise `first` is
 * returned with the polish's iteration count and a note in the message. */
int32_t sk_pdhg_polish(const sk_lp* lp, const sk_ipm_options* options,
                       const sk_pdhg_result* first, const double* first_col_value,
                       const double* first_row_dual, const double* first_col_dual,
                       sk_pdhg_result* result, double* col_value, double* row_dual,
                       double* col_dual, double* row_activity);

/* ---- presolve and postsolve (src/presolve/presolve.hpp:36-127; presolve.cpp:204-1697) -------
 * The reductions the reference's solve() applies ahead of every LP engine and undoes after it
 * (src/core/solve.cpp:415-443), on the host: the same eight record kinds, the same column and
 * row passes repeated to a fixed point (at most 20), the same tests and tolerances, so the
 * reduced LP is the one the reference's engine is handed; and postsolve of the primal point,
 * the row duals and the reduced costs back to the original LP (presolve.cpp:812-1697). */
#define SK_PRESOLVE_EMPTY_ROW 0
#define SK_PRESOLVE_REDUNDANT_ROW 1
#define SK_PRESOLVE_FIXED_COLUMN 2
#define SK_PRESOLVE_EMPTY_COLUMN 3
#define SK_PRESOLVE_SINGLETON_ROW 4
#define SK_PRESOLVE_FORCING_ROW 5
#define SK_PRESOLVE_FREE_COLUMN_SINGLETON 6
#define SK_PRESOLVE_DOUBLETON_EQUATION 7
typedef struct sk_presolved sk_presolved; /* a presolve's result and its records */

/* presolve::Result's counts (presolve.hpp:94-117), plus the records of each kind and the time */
typedef struct {
  int32_t original_rows;
  int32_t original_cols;
  int64_t original_nonzeros;
  int32_t reduced_rows;
  int32_t reduced_cols;
  int64_t reduced_nonzeros;
  int32_t records;
  int32_t passes;
  int32_t proved_infeasible;
  int32_t by_kind[8]; /* indexed by SK_PRESOLVE_* */
  int32_t reserved;
  double seconds;
} sk_presolve_info;

/* presolve::Record (presolve.hpp:66-91) */
typedef struct {
  int32_t kind; /* SK_PRESOLVE_* */
  int32_t index;
  double value;
  double coefficient;
  double row_lower;
  double row_upper;
  int32_t column;
  int32_t partner_column;
  double substituted_rhs;
  double partner_coefficient;
  double eliminated_cost;
} sk_presolve_record;

/* presolve::presolve (presolve.hpp:123): `lp` is left untouched. integer_columns holds one flag
 * per column or is NULL (every column continuous; integer columns get inward-rounded implied
 * bounds and are never eliminated by a doubleton); primal_feasibility_tolerance is the option of
 * that name (options.cpp:293-299, default 1e-7; <= 0 takes the default). A proof of
 * infeasibility is a success: sk_presolve_get_info says proved_infeasible and
 * sk_presolve_message names the row or column. */
int32_t sk_presolve(const sk_lp* lp, const uint8_t* integer_columns,
                    double primal_feasibility_tolerance, sk_presolved** out);
int32_t sk_presolve_get_info(const sk_presolved* presolve, sk_presolve_info* info);
/* Result::message: empty unless infeasibility was proved. Owned by the handle. */
const char* sk_presolve_message(const sk_presolved* presolve);
/* The reduced LP (Result::model) as a new handle the caller frees with sk_lp_free; names are
 * carried over when the original was read from a file. */
int32_t sk_presolve_reduced_lp(const sk_presolved* presolve, sk_lp** out);
/* Result::col_to_original and row_to_original (reduced sizes; either may be NULL). */
int32_t sk_presolve_maps(const sk_presolved* presolve, int32_t* col_to_original,
                         int32_t* row_to_original);
int32_t sk_presolve_get_record(const sk_presolved* presolve, int32_t k, sk_presolve_record* out);
/* presolve::postsolve (presolve.hpp:130-131): a point of the reduced LP (sizes of the reduced
 * LP; Solution signs, col_dual = c - A^T row_dual) back to the original LP `original` (the one
 * sk_presolve was given), into arrays of the original sizes. The caller measures the result on
 * the original LP, as the reference's postsolve does with recompute_quality. */
int32_t sk_postsolve(const sk_presolved* presolve, const sk_lp* original, const double* col_value,
                     const double* row_dual, const double* col_dual, double* out_col_value,
                     double* out_row_dual, double* out_col_dual);
void sk_presolve_free(sk_presolved* presolve);

/* ---- batches: many LPs of uneven size in one block-diagonal solve --------------------------
 * No counterpart in the reference, which solves one LP per call (src/core/solve.cpp:385-399);
 * every instance of a batch runs the engine of sk_pdhg_solve, and gets its answer bit for bit
 * when both run at the same tile capacity (sk_pdhg_options.tile_nnz; the default follows the
 * size of the batch), else the same to the request's tolerance. */
typedef struct sk_batch sk_batch;

typedef struct {
  double setup_ms;  /* device scaling, power iteration, initial state */
  double loop_ms;   /* PDHG graph launches until every instance stopped */
  double report_ms; /* final evaluation and download */
  int64_t graph_launches;
  int64_t total_rows;
  int64_t total_cols;
  int64_t total_nnz;
  int32_t instances;
  int32_t spmv_stage;
  int32_t streams;  /* sub-batches the batch ran as (sk_pdhg_options.streams) */
  int32_t tile_nnz; /* row-block capacity the batch runs with (picked by the autotuner when
                       sk_pdhg_options.tile_nnz is -1) */
  int32_t fused;    /* 1 when the batch runs as the resident solve: many steps and
                       evaluations per launch (sk_pdhg_options.step_fusion) */
  int32_t column_slices; /* column slices A x-bar runs over (sk_pdhg_options.column_slices),
                            0 when it runs unsliced */
} sk_batch_stats;

int32_t sk_batch_create(const sk_pdhg_options* options, sk_batch** out);
/* The LP is borrowed: it must outlive sk_batch_upload. Returns the instance index. */
int32_t sk_batch_add(sk_batch* batch, const sk_lp* lp, int32_t* index);
/* Host CSR build, tiles and device allocation + upload (inputs resident afterwards). */
/* Start instance `index` from col_value (num_cols) and row_dual (num_rows, the Solution's
 * signs) instead of x = proj(0), y = 0 (pdhg.cpp:327-334); a NaN entry keeps that element's
 * cold start; primal_weight and step_size > 0 replace omega = 1 and eta = 1 / ||A-hat||. Call
 * before sk_batch_upload. No counterpart in the reference, whose PDHG always starts cold. */
int32_t sk_batch_warm_start(sk_batch* batch, int32_t index, const double* col_value,
                            const double* row_dual, double primal_weight, double step_size);
int32_t sk_batch_upload(sk_batch* batch);
/* One complete solve of every instance from the uploaded data, on the device. May be called
 * again; each call starts from scratch. */
int32_t sk_batch_solve(sk_batch* batch);
int32_t sk_batch_result(sk_batch* batch, int32_t index, sk_pdhg_result* result,
                        double* col_value, double* row_dual, double* col_dual,
                        double* row_activity);
int32_t sk_batch_stats_get(const sk_batch* batch, sk_batch_stats* out);
/* Mean duration (ms) of one launch of a step kernel over the whole batch, measured with HIP
 * events on the batch's stream after a solve: kernel 0 primal step, 1 A xbar + dual step,
 * 2 A^T y + interaction, with every instance stepping; 3, 4, 5 the same kernels in normal
 * mode, where the tiles of instances that stopped exit at once. `bytes` receives the
 * algorithmic bytes of one launch with every instance stepping. A batch run as several
 * sub-batches (streams != 1) launches the kernel on every sub-batch's stream at once and
 * reports the wall time per repetition until the last is done. */
int32_t sk_batch_time_kernel(sk_batch* batch, int32_t kernel, int32_t reps, double* mean_ms,
                             double* bytes);
/* Probes of a solved batch, for parity tests: what the device computes, in the LP's own row
 * and column order, for instance `index`.
 * sk_batch_probe_spmv runs the step kernels' own sparse product (the per-tile staging,
 * gathers and fixed per-row summation order of the PDHG step) on a caller's vector, in the
 * iterate precision: which = 1 gives out (num_rows) = A-hat in (num_cols), the product the
 * dual step forms as A-hat x-bar; which = 2 gives out (num_cols) = A-hat^T in (num_rows), the
 * product the A^T step forms as A-hat^T y'. A-hat is the device-scaled matrix. The
 * counterparts are SparseMatrix::multiply / transpose_multiply (src/la/sparse.cpp:202-239)
 * applied to the reference's scaled matrix. */
int32_t sk_batch_probe_spmv(sk_batch* batch, int32_t index, int32_t which, const double* in,
                            double* out);
/* The device scaling of instance `index`: D_r (num_rows), D_c (num_cols) and the scaled FP64
 * values of A in the LP's CSC order (nnz); any pointer may be NULL. The counterpart is the
 * reference's Scaling (include/sankhya/scaling.hpp:42-58; built by src/la/scaling.cpp:111-159):
 * row_scale, col_scale and the scaled matrix's values. */
int32_t sk_batch_probe_scaling(sk_batch* batch, int32_t index, double* row_scale,
                               double* col_scale, double* scaled_values);
/* The LP column each device column of instance `index` holds (num_cols entries): the device
 * order the column_order option chose. No counterpart in the reference, whose engine keeps
 * the model's order. */
int32_t sk_batch_column_order(const sk_batch* batch, int32_t index, int32_t* device_to_lp);
void sk_batch_destroy(sk_batch* batch);

/* ---- serving: a continuous-batching request loop -------------------------------------------
 * No counterpart in the reference, which solves one LP per call (src/core/solve.cpp:385-399).
 * Requests queue in arrival order; `slots` worker threads, each owning `lanes` batch engines
 * with their own streams, take every waiting request that fits the batch and device-memory
 * limits as one block-diagonal batch, so batches overlap on the device. Each request completes
 * (retires) as soon as its own instance stops, while the rest of its batch keeps running, and
 * once a batch is in its tail its slot admits the next batch beside it. A request's answer is
 * bitwise sk_pdhg_solve's with the same options when its batch and the solo solve run at the
 * same tile capacity (a fixed sk_pdhg_options.tile_nnz, or default capacities that agree), else
 * the same to the request's tolerance. */
typedef struct sk_server sk_server;

typedef struct {
  int32_t slots;          /* concurrent batches (default 2: more shorten the queue wait
                             but slow every batch and lengthen the latency tail) */
  int32_t max_batch;      /* instances per batch (default 256) */
  int64_t memory_budget;  /* device bytes all slots may hold; 0: 90% of free memory at create */
  int32_t keep_solutions; /* 1: keep the primal/dual vectors for sk_server_result (default 1) */
  int32_t lanes;          /* batches one slot may run at once, 1..8 (default 1): with more,
                             once the slot's newest batch is in its tail the queue head is
                             admitted into the next idle lane (its own engine and stream) while
                             the tail finishes. It shortens the wait of a one-slot server; with
                             two slots the device is already full and it lengthens latency */
  double tail_fraction;   /* a batch is in its tail when at most this share of its instances
                             is still running (default 0.25) */
} sk_server_options;

typedef struct {
  int64_t id;
  double arrival;   /* seconds on the server clock (sk_server_now) */
  double admitted;  /* its batch was formed */
  double finished;  /* its answer was read back */
  double latency;   /* finished - arrival */
  int32_t slot;
  int32_t batch_size;
  int32_t error;    /* SK_GPU_* status when its batch failed, else SK_GPU_SUCCESS */
  int32_t num_rows;
  int32_t num_cols;
  int32_t reserved;
} sk_server_record;

typedef struct {
  int64_t submitted;
  int64_t completed;
  int64_t rejected;       /* larger than the memory budget */
  int64_t failed;         /* completed with an error */
  int64_t batches;
  int64_t max_queue;
  int64_t max_batch_seen;
  int64_t memory_budget;
  int64_t peak_held;      /* largest device memory held by the slots together */
  int64_t releases;       /* idle slots that gave their memory back to admit a request */
} sk_server_stats;

void sk_server_default_options(sk_server_options* out);
/* Starts the slot threads; fails with SK_GPU_NO_DEVICE without a device. */
int32_t sk_server_create(const sk_pdhg_options* options, const sk_server_options* server_options,
                         sk_server** out);
int32_t sk_server_now(const sk_server* server, double* seconds);
/* On success the server owns `lp` (do not use or free it again); `arrival` < 0 means now.
 * An LP whose estimated device footprint exceeds the budget is refused with
 * SK_GPU_OUT_OF_MEMORY, and the caller keeps it. */
int32_t sk_server_submit(sk_server* server, sk_lp* lp, double arrival, int64_t* id);
/* Ids completed since the previous call, oldest first, at most `capacity` (non-blocking). */
int32_t sk_server_completed(sk_server* server, int64_t* ids, int32_t capacity, int32_t* count);
/* Blocks until every submitted request has completed or `timeout` seconds passed. */
int32_t sk_server_wait(sk_server* server, double timeout, int32_t* idle);
/* The answer of a completed request, once: the server forgets it afterwards. Arrays as
 * sk_pdhg_solve's (NULL skips one); they must be NULL when keep_solutions is 0. */
int32_t sk_server_result(sk_server* server, int64_t id, sk_server_record* record,
                         sk_pdhg_result* result, double* col_value, double* row_dual,
                         double* col_dual, double* row_activity);
int32_t sk_server_stats_get(const sk_server* server, sk_server_stats* out);
/* Finishes the running batches, drops queued requests and frees everything. */
void sk_server_destroy(sk_server* server);

#ifdef __cplusplus
}
#endif

#endif /* SK_GPU_H */
