"""MI355X restarted-PDHG LP engine: a drop-in for the `algorithm=pdhg` path of the SANKHYA
solver, as hand-written HIP kernels for gfx950 behind a C ABI (include/sk_gpu.h).

Importing this package loads sih26_amd/lib/libsk_gpu.so and fails loudly if it is not built.
"""
from .api import (Batch, Lp, Presolved, Server, SkError, SolveResult, adaptive_config,
                  analyze_problem, device_info, gpu_available, ipm_options, lib, make_options,
                  polish, presolve, refine, select_backend, solve, solve_ipm, write_solution)

lib()  # load now: a missing library is an import error, never a silent fallback

__all__ = ["Batch", "Lp", "Presolved", "Server", "SkError", "SolveResult", "adaptive_config",
           "analyze_problem", "device_info", "gpu_available", "ipm_options", "lib",
           "make_options", "polish", "presolve", "refine", "select_backend", "solve", "solve_ipm",
           "write_solution"]
