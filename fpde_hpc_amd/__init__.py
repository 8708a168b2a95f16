"""MI355X-native solver for the Caputo L1 history-sum path of FPDE_HPC.

Host API over libfpde_hip.so (hand-written gfx950 HIP kernels + C++ host
runtime, C ABI in include/fpde_hip.h).
"""
from ._lib import FpdeError, LIB_PATH, check, lib, shutdown  # noqa: F401
from .solver import (HIST_MODES, Server, Solver, advance_batch, autotune_block, footprint, init_1d, init_3d,  # noqa: F401
                     make_spec, serving_mix, solve, solve_batch, varT_step, workload)

__all__ = ["FpdeError", "LIB_PATH", "check", "lib", "shutdown", "HIST_MODES", "Server", "Solver", "advance_batch",
           "autotune_block", "footprint", "init_1d", "init_3d", "make_spec", "serving_mix", "solve", "solve_batch",
           "varT_step", "workload"]
