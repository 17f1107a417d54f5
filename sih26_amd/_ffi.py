"""ctypes declarations of the C ABI in include/sk_gpu.h (libsk_gpu.so).

The library is loaded from sih26_amd/lib/ only; if it is missing, importing the product
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libsk_gpu.so"

i32, i64, u64, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
P = C.POINTER
vp = C.c_void_p
cstr = C.c_char_p


class DeviceInfo(C.Structure):
    _fields_ = [("device_id", i32), ("name", C.c_char * 256), ("total_memory_bytes", u64),
                ("gcn_arch_major", i32), ("gcn_arch_minor", i32),
                ("multiprocessor_count", i32), ("warp_size", i32),
                ("max_threads_per_block", i32), ("gcn_arch_name", C.c_char * 64)]


class Comparison(C.Structure):
    _fields_ = [("passed", i32), ("contains_nan_or_inf", i32), ("max_abs_diff", f64),
                ("max_rel_diff", f64)]


class GenParams(C.Structure):
    _fields_ = [("seed", u64), ("num_rows", i32), ("num_cols", i32),
                ("num_equality_rows", i32), ("degree_kind", i32), ("degree_lo", i32),
                ("degree_hi", i32), ("power_exponent", f64), ("target_nnz", i64),
                ("value_kind", i32), ("value_lo", f64), ("value_hi", f64),
                ("negative_probability", f64), ("upper_lo", f64), ("upper_hi", f64),
                ("slack_lo", f64), ("slack_hi", f64), ("cost_lo", f64), ("cost_hi", f64),
                ("fix_empty_rows", i32)]


class PdhgOptions(C.Structure):
    _fields_ = [("tolerance", f64), ("iteration_limit", i64), ("time_limit", f64),
                ("restart", i32), ("stop_at_request", i32), ("absolute_primal", i32),
                ("random_seed", i64), ("precision", i32), ("steps_per_block", i32),
                ("use_graph", i32), ("spmv_stage", i32), ("tile_nnz", i32),
                ("compute_quality", i32), ("eval_stride", i32), ("streams", i32),
                ("column_order", i32), ("step_fusion", i32), ("backend", i32), ("gpu", i32),
                ("presolve", i32), ("column_slices", i32)]


class ProblemCharacteristics(C.Structure):
    _fields_ = [("problem_class", C.c_char * 8), ("num_rows", i32), ("num_cols", i32),
                ("num_nonzeros", i64), ("density", f64), ("sparsity_ratio", f64),
                ("num_integer_cols", i32), ("has_quadratic_objective", i32),
                ("estimated_matvec_flops", f64), ("estimated_memory_bytes", f64),
                ("gpu_available", i32), ("gpu_device_count", i32), ("gpu_memory_bytes", u64),
                ("gpu_name", C.c_char * 256)]


class AdaptiveConfig(C.Structure):
    _fields_ = [("min_gpu_nonzeros", i64), ("min_gpu_dimension", i64), ("min_gpu_density", f64),
                ("max_gpu_memory_headroom", f64), ("explain_decision", i32), ("reserved", i32)]


class BackendDecision(C.Structure):
    _fields_ = [("selected_backend", i32), ("requested_mode", i32), ("is_valid", i32),
                ("detail_count", i32), ("primary_reason", C.c_char * 512),
                ("explanation", C.c_char * 2048)]


class PdhgResult(C.Structure):
    _fields_ = [("status", i32), ("converged", i32), ("iterations", i64), ("restarts", i64),
                ("evaluations", i64), ("attempts", i64), ("objective", f64),
                ("dual_bound", f64), ("spectral_norm", f64), ("solve_seconds", f64),
                ("rel_primal", f64), ("rel_dual", f64), ("rel_gap", f64), ("abs_primal", f64),
                ("abs_dual", f64), ("gap_as_verified", f64), ("complementarity", f64),
                ("primal_objective", f64), ("dual_objective", f64),
                ("primal_infeasibility", f64), ("primal_infeasibility_scaled", f64),
                ("dual_infeasibility", f64), ("dual_infeasibility_scaled", f64),
                ("complementarity_violation", f64), ("absolute_gap", f64),
                ("relative_gap", f64), ("algorithm", C.c_char * 16),
                ("message", C.c_char * 512), ("polish_iterations", i64),
                ("primal_weight", f64), ("step_size", f64)]


class IpmOptions(C.Structure):
    _fields_ = [("iteration_limit", i64), ("start_margin", f64), ("max_seconds", f64),
                ("max_factor_nonzeros", i64), ("scaling", i32), ("reserved", i32)]


class BatchStats(C.Structure):
    _fields_ = [("setup_ms", f64), ("loop_ms", f64), ("report_ms", f64),
                ("graph_launches", i64), ("total_rows", i64), ("total_cols", i64),
                ("total_nnz", i64), ("instances", i32), ("spmv_stage", i32),
                ("streams", i32), ("tile_nnz", i32),
                ("fused", i32), ("column_slices", i32)]


class ServerOptions(C.Structure):
    _fields_ = [("slots", i32), ("max_batch", i32), ("memory_budget", i64),
                ("keep_solutions", i32), ("lanes", i32), ("tail_fraction", f64)]


class ServerRecord(C.Structure):
    _fields_ = [("id", i64), ("arrival", f64), ("admitted", f64), ("finished", f64),
                ("latency", f64), ("slot", i32), ("batch_size", i32), ("error", i32),
                ("num_rows", i32), ("num_cols", i32), ("reserved", i32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", i64), ("completed", i64), ("rejected", i64), ("failed", i64),
                ("batches", i64), ("max_queue", i64), ("max_batch_seen", i64),
                ("memory_budget", i64), ("peak_held", i64), ("releases", i64)]


class PresolveInfo(C.Structure):
    _fields_ = [("original_rows", i32), ("original_cols", i32), ("original_nonzeros", i64),
                ("reduced_rows", i32), ("reduced_cols", i32), ("reduced_nonzeros", i64),
                ("records", i32), ("passes", i32), ("proved_infeasible", i32),
                ("by_kind", i32 * 8), ("reserved", i32), ("seconds", f64)]


class PresolveRecord(C.Structure):
    _fields_ = [("kind", i32), ("index", i32), ("value", f64), ("coefficient", f64),
                ("row_lower", f64), ("row_upper", f64), ("column", i32),
                ("partner_column", i32), ("substituted_rhs", f64),
                ("partner_coefficient", f64), ("eliminated_cost", f64)]


# name -> (restype, argtypes)
SIGNATURES = {
    "sk_gpu_status_string": (cstr, [i32]),
    "sk_gpu_is_available": (i32, []),
    "sk_gpu_device_count": (i32, []),
    "sk_gpu_get_device_info": (i32, [i32, P(DeviceInfo)]),
    "sk_gpu_set_device": (i32, [i32]),
    "sk_gpu_synchronize": (i32, []),
    "sk_gpu_last_error": (cstr, []),
    "sk_fault_inject": (i32, [i32, i64]),
    "sk_fault_fired": (i64, [i32]),
    "sk_gpu_malloc": (i32, [P(vp), C.c_size_t]),
    "sk_gpu_free": (i32, [vp]),
    "sk_gpu_copy_h2d": (i32, [vp, vp, C.c_size_t]),
    "sk_gpu_copy_d2h": (i32, [vp, vp, C.c_size_t]),
    "sk_gpu_memset": (i32, [vp, i32, C.c_size_t]),
    "sk_gpu_axpy": (i32, [i32, f64, vp, vp]),
    "sk_gpu_dot": (i32, [i32, vp, vp, P(f64)]),
    "sk_gpu_dot_noalloc": (i32, [i32, vp, vp, P(f64)]),
    "sk_gpu_pdhg_primal_step": (i32, [i32, f64, vp, vp, vp, vp, vp, vp, vp]),
    "sk_gpu_execute_pdhg_primal_step": (i32, [i32, f64, vp, vp, vp, vp, vp, vp, vp]),
    "sk_gpu_spmv_csc": (i32, [i32, i32, vp, vp, vp, vp, vp, f64, i32]),
    "sk_gpu_spmv_csr": (i32, [i32, i32, vp, vp, vp, vp, vp, f64]),
    "sk_gpu_compare_vectors": (i32, [i32, vp, vp, f64, P(Comparison)]),
    "sk_lp_create": (i32, [i32, i32, i32, f64, vp, vp, vp, vp, vp, vp, vp, vp, P(vp)]),
    "sk_lp_generate": (i32, [i32, u64, P(vp)]),
    "sk_gen_preset": (i32, [i32, u64, P(GenParams)]),
    "sk_lp_generate_params": (i32, [P(GenParams), P(vp)]),
    "sk_lp_dims": (i32, [vp, P(i32), P(i32), P(i64)]),
    "sk_lp_export": (i32, [vp, P(i32), P(f64), vp, vp, vp, vp, vp, vp, vp, vp]),
    "sk_lp_free": (None, [vp]),
    "sk_pdhg_default_options": (None, [P(PdhgOptions)]),
    "sk_pdhg_set_option": (i32, [P(PdhgOptions), cstr, cstr]),
    "sk_adaptive_default_config": (None, [P(AdaptiveConfig)]),
    "sk_adaptive_set_option": (i32, [P(AdaptiveConfig), cstr, cstr]),
    "sk_analyze_problem": (i32, [vp, i32, i32, P(ProblemCharacteristics)]),
    "sk_select_backend": (i32, [P(ProblemCharacteristics), P(PdhgOptions), P(AdaptiveConfig),
                                P(BackendDecision)]),
    "sk_pdhg_solve": (i32, [vp, P(PdhgOptions), P(PdhgResult), vp, vp, vp, vp]),
    "sk_write_solution": (i32, [cstr, vp, P(PdhgResult), vp, vp, vp, vp, vp, vp, cstr]),
    "sk_ipm_default_options": (None, [P(IpmOptions)]),
    "sk_lp_read_mps": (i32, [cstr, P(vp), P(i32)]),
    "sk_lp_name": (cstr, [vp, i32, i32]),
    "sk_lp_write_mps": (i32, [vp, cstr]),
    "sk_lp_save": (i32, [vp, cstr]),
    "sk_lp_load": (i32, [cstr, P(vp)]),
    "sk_ipm_solve": (i32, [vp, P(IpmOptions), vp, vp, vp, P(PdhgResult), vp, vp, vp, vp]),
    "sk_pdhg_polish": (i32, [vp, P(IpmOptions), P(PdhgResult), vp, vp, vp, P(PdhgResult), vp, vp,
                             vp, vp]),
    "sk_presolve": (i32, [vp, vp, f64, P(vp)]),
    "sk_presolve_get_info": (i32, [vp, P(PresolveInfo)]),
    "sk_presolve_message": (cstr, [vp]),
    "sk_presolve_reduced_lp": (i32, [vp, P(vp)]),
    "sk_presolve_maps": (i32, [vp, vp, vp]),
    "sk_presolve_get_record": (i32, [vp, i32, P(PresolveRecord)]),
    "sk_postsolve": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "sk_presolve_free": (None, [vp]),
    "sk_batch_create": (i32, [P(PdhgOptions), P(vp)]),
    "sk_batch_add": (i32, [vp, vp, P(i32)]),
    "sk_batch_warm_start": (i32, [vp, i32, vp, vp, f64, f64]),
    "sk_batch_upload": (i32, [vp]),
    "sk_pdhg_refine": (i32, [vp, P(PdhgOptions), P(PdhgResult), vp, vp, P(PdhgResult), vp, vp,
                             vp, vp]),
    "sk_batch_solve": (i32, [vp]),
    "sk_batch_result": (i32, [vp, i32, P(PdhgResult), vp, vp, vp, vp]),
    "sk_batch_stats_get": (i32, [vp, P(BatchStats)]),
    "sk_batch_time_kernel": (i32, [vp, i32, i32, P(f64), P(f64)]),
    "sk_batch_probe_spmv": (i32, [vp, i32, i32, vp, vp]),
    "sk_batch_probe_scaling": (i32, [vp, i32, vp, vp, vp]),
    "sk_batch_column_order": (i32, [vp, i32, vp]),
    "sk_batch_destroy": (None, [vp]),
    "sk_server_default_options": (None, [P(ServerOptions)]),
    "sk_server_create": (i32, [P(PdhgOptions), P(ServerOptions), P(vp)]),
    "sk_server_now": (i32, [vp, P(f64)]),
    "sk_server_submit": (i32, [vp, vp, f64, P(i64)]),
    "sk_server_completed": (i32, [vp, P(i64), i32, P(i32)]),
    "sk_server_wait": (i32, [vp, f64, P(i32)]),
    "sk_server_result": (i32, [vp, i64, P(ServerRecord), P(PdhgResult), vp, vp, vp, vp]),
    "sk_server_stats_get": (i32, [vp, P(ServerStats)]),
    "sk_server_destroy": (None, [vp]),
}


def load(path: Path = LIB_PATH) -> C.CDLL:
    if not path.exists():
        raise ImportError(
            f"{path} is not built; run `python sih26_amd/build.py` (or __graft_entry__.build())")
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
