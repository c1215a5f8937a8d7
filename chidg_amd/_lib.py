"""ctypes binding of ``libchidg_hip.so`` (declarations: include/chidg_hip.h).

The product path has no fallback: if the library is missing this module
raises, naming the build step.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchidg_hip.so")

CDG_OK = 0
ERRORS = {1: "CDG_ERR_ARG", 2: "CDG_ERR_HIP", 3: "CDG_ERR_SINGULAR", 4: "CDG_ERR_NO_TRANSPOSE",
          5: "CDG_ERR_OOM", 6: "CDG_ERR_NAN", 7: "CDG_ERR_STATE", 8: "CDG_ERR_UNSUPPORTED"}
KIND_LOCAL, KIND_COUPLING = 0, 1
PREC_IDENTITY, PREC_ILU0, PREC_JACOBI = 0, 1, 2
FP64, FP32, BF16 = 0, 1, 2
CGS, MGS = 0, 1

EXPORTS = [
    "cdg_ctx_create", "cdg_ctx_destroy", "cdg_last_error", "cdg_ctx_mem_used", "cdg_ctx_stream",
    "cdg_ctx_wait_stream", "cdg_ctx_signal_stream",
    "cdg_mat_create", "cdg_mat_destroy", "cdg_mat_upload", "cdg_mat_download",
    "cdg_mat_fill_synthetic", "cdg_vec_fill_synthetic", "cdg_mat_stamp", "cdg_mat_set_partition", "cdg_mat_info", "cdg_mv", "cdg_mv_async", "cdg_mv_multi",
    "cdg_prec_create", "cdg_prec_update", "cdg_prec_apply", "cdg_prec_apply_async", "cdg_prec_apply_noalloc", "cdg_prec_apply_multi",
    "cdg_prec_destroy", "cdg_prec_download",
    "cdg_prec_set_schedule", "cdg_prec_tune", "cdg_mat_tune", "cdg_prec_launch_shape", "cdg_mat_launch_shape",
    "cdg_server_create", "cdg_server_submit", "cdg_server_wait", "cdg_server_footprint", "cdg_server_info",
    "cdg_server_last_error", "cdg_server_destroy", "cdg_server_register", "cdg_server_submit_system",
    "cdg_server_unregister",
    "cdg_fgmres_solve", "cdg_fgmres_solve_multi", "cdg_fgmres_default_params",
    "cdg_batch_create", "cdg_batch_solve", "cdg_batch_update", "cdg_batch_info", "cdg_batch_destroy",
    "cdg_euler_state", "cdg_euler_residual", "cdg_euler_jacobian", "cdg_euler_residual_gpu",
    "cdg_vec_create", "cdg_vec_destroy", "cdg_vec_copy", "cdg_vec_axpby", "cdg_vec_dot", "cdg_vec_norm",
]


class Request(C.Structure):
    """cdg_request: one synthetic parameter-sweep system."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("order", C.c_int32),
                ("periodic", C.c_int32), ("precision", C.c_int32), ("seed_A", C.c_uint64), ("seed_b", C.c_uint64),
                ("s_diag", C.c_double), ("s_off", C.c_double), ("nrhs", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("niter", C.c_int32), ("residual_norm", C.c_double),
                ("t_queue", C.c_double), ("t_run", C.c_double), ("t_latency", C.c_double),
                ("t_factor", C.c_double), ("slot", C.c_int32), ("factored", C.c_int32)]


class FgmresParams(C.Structure):
    _fields_ = [("nkrylov", C.c_int32), ("tol", C.c_double), ("rtol", C.c_double), ("nmax", C.c_int32),
                ("orthogonalization", C.c_int32), ("inner_fgmres", C.c_int32), ("inner_nkrylov", C.c_int32),
                ("inner_tol", C.c_double), ("inner_rtol", C.c_double), ("inner_nmax", C.c_int32),
                ("inner_orthogonalization", C.c_int32), ("update_preconditioner", C.c_int32)]


class FgmresStats(C.Structure):
    _fields_ = [("niter", C.c_int32), ("residual_norm", C.c_double), ("nrestart", C.c_int32),
                ("nreorth", C.c_int32), ("factored", C.c_int32), ("t_total", C.c_double),
                ("t_factor", C.c_double), ("t_precon", C.c_double), ("t_mv", C.c_double),
                ("t_dot", C.c_double), ("t_norm", C.c_double)]


class CdgError(RuntimeError):
    def __init__(self, code: int, what: str, msg: str = ""):
        super().__init__(f"{what} failed: {ERRORS.get(code, code)} {msg}".strip())
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # PyTorch ships its own HIP runtime with the same SONAME (libamdhip64.so.7).
    # Loading torch first makes this library bind to that runtime, so a process
    # that uses both has one HIP runtime; loaded the other way round, torch
    # would start a second runtime that cannot see the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python chidg_amd/build.py` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int32), C.c_void_p
    pp = C.POINTER(C.c_void_p)
    sig = {
        "cdg_ctx_create": ([C.c_int32, pp], C.c_int),
        "cdg_ctx_destroy": ([vp], C.c_int),
        "cdg_last_error": ([vp], C.c_char_p),
        "cdg_ctx_mem_used": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "cdg_ctx_stream": ([vp, pp], C.c_int),
        "cdg_ctx_wait_stream": ([vp, vp], C.c_int),
        "cdg_ctx_signal_stream": ([vp, vp], C.c_int),
        "cdg_mv_async": ([vp, vp, dp, dp], C.c_int),
        "cdg_prec_apply_async": ([vp, vp, vp, dp, dp], C.c_int),
        "cdg_prec_apply_noalloc": ([vp, vp, vp, dp, dp], C.c_int),
        "cdg_prec_apply_multi": ([vp, vp, vp, C.c_int32, dp, dp], C.c_int),
        "cdg_mv_multi": ([vp, vp, C.c_int32, dp, dp], C.c_int),
        "cdg_mat_create": ([vp, C.c_int32, C.c_int32, ip, ip, ip, pp], C.c_int),
        "cdg_mat_destroy": ([vp, vp], C.c_int),
        "cdg_mat_upload": ([vp, vp, dp], C.c_int),
        "cdg_mat_download": ([vp, vp, dp], C.c_int),
        "cdg_mat_fill_synthetic": ([vp, vp, C.c_uint64, C.c_double, C.c_double], C.c_int),
        "cdg_vec_fill_synthetic": ([vp, vp, C.c_uint64, C.c_double, dp], C.c_int),
        "cdg_mat_stamp": ([vp, C.POINTER(C.c_uint64)], C.c_int),
        "cdg_mat_set_partition": ([vp, vp, ip], C.c_int),
        "cdg_mat_info": ([vp, ip, ip, ip], C.c_int),
        "cdg_mv": ([vp, vp, dp, dp], C.c_int),
        "cdg_prec_create": ([vp, vp, C.c_int32, C.c_int32, pp], C.c_int),
        "cdg_prec_update": ([vp, vp, vp, ip], C.c_int),
        "cdg_prec_apply": ([vp, vp, vp, dp, dp], C.c_int),
        "cdg_prec_destroy": ([vp, vp], C.c_int),
        "cdg_prec_download": ([vp, vp, dp, dp], C.c_int),
        "cdg_fgmres_solve": ([vp, vp, vp, dp, dp, C.POINTER(FgmresParams), C.POINTER(FgmresStats)], C.c_int),
        "cdg_fgmres_solve_multi": ([vp, vp, vp, C.c_int32, dp, dp, C.POINTER(FgmresParams), C.POINTER(FgmresStats)],
                                   C.c_int),
        "cdg_batch_create": ([vp, C.c_int32, vp, vp, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
        "cdg_batch_solve": ([vp, vp, vp, vp, C.POINTER(FgmresParams), C.POINTER(FgmresStats)], C.c_int),
        "cdg_batch_update": ([vp, vp, vp], C.c_int),
        "cdg_batch_info": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)], C.c_int),
        "cdg_batch_destroy": ([vp, vp], C.c_int),
        "cdg_fgmres_default_params": ([C.POINTER(FgmresParams)], None),
        "cdg_prec_set_schedule": ([vp, vp, C.c_int32], C.c_int),
        "cdg_prec_tune": ([vp, vp, vp, C.c_int32, C.POINTER(C.c_double)], C.c_int),
        "cdg_mat_tune": ([vp, vp, C.c_int32, C.POINTER(C.c_double)], C.c_int),
        "cdg_prec_launch_shape": ([vp, ip], C.c_int),
        "cdg_mat_launch_shape": ([vp, ip], C.c_int),
        "cdg_server_create": ([C.c_int32, C.c_int32, C.c_int64, C.POINTER(FgmresParams), pp], C.c_int),
        "cdg_server_submit": ([vp, C.POINTER(Request), dp, C.POINTER(C.c_uint64)], C.c_int),
        "cdg_server_wait": ([vp, C.c_uint64, C.POINTER(Result)], C.c_int),
        "cdg_server_footprint": ([C.POINTER(Request), C.POINTER(FgmresParams), C.POINTER(C.c_int64)], C.c_int),
        "cdg_server_info": ([vp, ip, ip, ip, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "cdg_server_last_error": ([vp], C.c_char_p),
        "cdg_server_destroy": ([vp], C.c_int),
        "cdg_server_register": ([vp, vp, C.c_int32, C.POINTER(C.c_uint64)], C.c_int),
        "cdg_server_submit_system": ([vp, C.c_uint64, dp, C.c_int32, dp, C.POINTER(C.c_uint64)], C.c_int),
        "cdg_server_unregister": ([vp, C.c_uint64], C.c_int),
        "cdg_euler_state": ([C.c_int32] * 4 + [C.c_double, C.c_uint64, C.c_double, dp], C.c_int),
        "cdg_euler_residual": ([C.c_int32] * 4 + [dp, dp, C.c_int32], C.c_int),
        "cdg_euler_residual_gpu": ([vp] + [C.c_int32] * 4 + [dp, dp], C.c_int),
        "cdg_vec_create": ([vp, C.c_int64, C.POINTER(C.c_void_p)], C.c_int),
        "cdg_vec_destroy": ([vp, dp], C.c_int),
        "cdg_vec_copy": ([vp, C.c_int64, dp, dp], C.c_int),
        "cdg_vec_axpby": ([vp, C.c_int64, C.c_double, dp, C.c_double, dp], C.c_int),
        "cdg_vec_dot": ([vp, C.c_int64, dp, dp, C.POINTER(C.c_double)], C.c_int),
        "cdg_vec_norm": ([vp, C.c_int64, dp, C.POINTER(C.c_double)], C.c_int),
        "cdg_euler_jacobian": ([C.c_int32] * 4 + [dp, C.c_double, dp, dp, C.c_int32], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib
