"""ctypes binding of libfpde_hip.so (include/fpde_hip.h).

The library is built in-tree by fpde_hpc_amd.build; loading fails loudly when
it is missing -- there is no CPU fallback in this package.
"""
import atexit
import ctypes
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfpde_hip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class FpdeError(RuntimeError):
    pass


class Spec(ctypes.Structure):
    """fpde_spec"""
    _fields_ = [
        ("dims", ctypes.c_int), ("nodes", ctypes.c_int), ("steps", ctypes.c_int),
        ("tau", ctypes.c_double), ("alpha", ctypes.c_double), ("length", ctypes.c_double),
        ("ksat", ctypes.c_double), ("coef_mode", ctypes.c_int), ("picard_max", ctypes.c_int),
        ("picard_tol", ctypes.c_double), ("weight_mode", ctypes.c_int), ("series_eps", ctypes.c_double),
        ("do_q", ctypes.c_int), ("do_rule", ctypes.c_int), ("do_a0", ctypes.c_double),
        ("do_a1", ctypes.c_double), ("hist_mode", ctypes.c_int), ("block", ctypes.c_int),
        ("step_mode", ctypes.c_int), ("sum_alg", ctypes.c_int), ("sum_param", ctypes.c_double),
        ("taus", ctypes.c_void_p), ("var_tau", ctypes.c_int), ("split3d", ctypes.c_int),
        ("gamma", ctypes.c_double), ("fk", ctypes.c_int), ("fp", ctypes.c_double),
        ("chunk_rows", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    """fpde_stats"""
    _fields_ = [
        ("ms_total", ctypes.c_double), ("ms_history", ctypes.c_double), ("ms_step", ctypes.c_double),
        ("history_launches", ctypes.c_int64), ("step_launches", ctypes.c_int64),
        ("history_alg_bytes", ctypes.c_double), ("history_alg_flops", ctypes.c_double),
        ("device_bytes", ctypes.c_size_t), ("picard_iters_total", ctypes.c_int), ("reserved", ctypes.c_int),
        ("step_alg_bytes", ctypes.c_double), ("steps_done", ctypes.c_int64),
    ]


class ResultInfo(ctypes.Structure):
    """fpde_result_info"""
    _fields_ = [
        ("id", ctypes.c_int64), ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double),
        ("finish_s", ctypes.c_double), ("status", ctypes.c_int), ("steps_done", ctypes.c_int),
    ]


# (name, restype, argtypes) for every symbol include/fpde_hip.h declares
SIGNATURES = [
    ("fpde_last_error", ctypes.c_char_p, []),
    ("fpde_version", ctypes.c_int, []),
    ("fpde_shutdown", ctypes.c_int, []),
    ("fpde_device_count", ctypes.c_int, []),
    ("fpde_cl_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("fpde_cl_ctx_destroy", ctypes.c_int, [_vp]),
    ("fpde_cl_buf_alloc", ctypes.c_int, [_vp, ctypes.c_size_t, _vp, ctypes.POINTER(_vp)]),
    ("fpde_cl_buf_free", ctypes.c_int, [_vp]),
    ("fpde_cl_buf_size", ctypes.c_size_t, [_vp]),
    ("fpde_cl_buf_device_ptr", _vp, [_vp]),
    ("fpde_cl_buf_write", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_size_t, _vp]),
    ("fpde_cl_buf_read", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_size_t, _vp]),
    ("fpde_cl_buf_copy", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, ctypes.c_size_t]),
    ("fpde_cl_program_create", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    ("fpde_cl_program_destroy", ctypes.c_int, [_vp]),
    ("fpde_cl_kernel_create", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    ("fpde_cl_kernel_destroy", ctypes.c_int, [_vp]),
    ("fpde_cl_kernel_set_arg", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_size_t, _vp]),
    ("fpde_cl_kernel_set_buffer_arg", ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    ("fpde_cl_launch", ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_size_t)]),
    ("fpde_cl_finish", ctypes.c_int, [_vp]),
    ("fpde_reference_sum", ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                          _vp, _vp, ctypes.c_int, ctypes.c_double]),
    ("fpde_history_sum_f64", ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp]),
    ("fpde_history_sum_f64_noalloc", ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp]),
    ("fpde_space_sum", ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_double]),
    ("fpde_history_block_gemm", ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("fpde_solver_footprint", ctypes.c_size_t, [ctypes.POINTER(Spec)]),
    ("fpde_solver_create", ctypes.c_int, [ctypes.POINTER(Spec), ctypes.POINTER(_vp)]),
    ("fpde_solver_destroy", ctypes.c_int, [_vp]),
    ("fpde_solver_set_initial", ctypes.c_int, [_vp, _dp]),
    ("fpde_solver_reset", ctypes.c_int, [_vp, _dp]),
    ("fpde_solver_advance", ctypes.c_int, [_vp, ctypes.c_int]),
    ("fpde_solver_advance_graph", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    ("fpde_solver_advance_batch", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int]),
    ("fpde_autotune_block", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]),
    ("fpde_solver_sync", ctypes.c_int, [_vp]),
    ("fpde_solver_query", ctypes.c_int, [_vp]),
    ("fpde_solver_step", ctypes.c_int, [_vp]),
    ("fpde_solver_get_solution", ctypes.c_int, [_vp, _dp]),
    ("fpde_solver_write_results", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    ("fpde_solver_advance_tau", ctypes.c_int, [_vp, ctypes.c_double]),
    ("fpde_solver_copy", ctypes.c_int, [_vp, _vp]),
    ("fpde_varT_step", ctypes.c_int, [_vp, _vp, _vp, _dp, ctypes.c_double, ctypes.c_double, _dp]),
    ("fpde_solver_get_weights", ctypes.c_int, [_vp, _dp, _dp, _dp]),
    ("fpde_solver_set_profiling", ctypes.c_int, [_vp, ctypes.c_int]),
    ("fpde_solver_get_stats", ctypes.c_int, [_vp, ctypes.POINTER(Stats)]),
    ("fpde_init_1d", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _dp]),
    ("fpde_init_3d", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _dp]),
    ("fpde_ksat_for_fourier", ctypes.c_double, [ctypes.c_double] * 4),
    ("fpde_server_create", ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("fpde_server_destroy", ctypes.c_int, [_vp]),
    ("fpde_server_submit", ctypes.c_int, [_vp, ctypes.POINTER(Spec), _dp, ctypes.POINTER(ctypes.c_int64)]),
    ("fpde_server_cancel", ctypes.c_int, [_vp, ctypes.c_int64]),
    ("fpde_server_run", ctypes.c_int, [_vp, ctypes.c_int64]),
    ("fpde_server_poll", ctypes.c_int, [_vp, ctypes.c_int]),
    ("fpde_server_result", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.POINTER(ResultInfo), _dp,
                                          ctypes.c_size_t]),
    ("fpde_server_clock", ctypes.c_double, [_vp]),
]

_lib = None
# objects owning device state (Solver, Server): closed by shutdown() before
# the process exits, so no HIP call is left to interpreter finalisation
_owners = weakref.WeakSet()


def register_owner(obj):
    _owners.add(obj)


def shutdown():
    """Close every live Solver/Server, then release the library's own device
    tables (fpde_shutdown).  Registered with atexit on first load: Python's
    atexit handlers run before the interpreter tears modules down and before
    the C runtime's exit handlers (where a profiler finalises its tools)."""
    if _lib is None:
        return
    for obj in list(_owners):
        try:
            obj.close()
        except Exception:
            pass
    _lib.fpde_shutdown()


def lib():
    """Load libfpde_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FpdeError(f"{LIB_PATH} is missing: build it with `python -m fpde_hpc_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        atexit.register(shutdown)
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().fpde_last_error()
        raise FpdeError(f"fpde error {rc}: {msg.decode() if msg else ''}")
    return rc
