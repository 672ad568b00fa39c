"""ctypes binding of librigidfibers.so (declared in include/rigidfibers.h).

The library is the only compute path: if it is missing this module raises at
import time instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librigidfibers.so")

RF_OK = 0
RF_ERR_INVALID_ARGUMENT = 1
RF_ERR_UNSUPPORTED = 2
RF_ERR_OUT_OF_MEMORY = 3
RF_ERR_DEVICE = 4
RF_ERR_NONFINITE = 5
RF_ERR_SINGULAR = 6
RF_WARN_NOT_CONVERGED = 100

SOLVERS = {"gmres": 0, "bicgstab": 1, "direct": 2, "magma": 2, "mixed-bf16": 3, "mixed-fp16": 4, "gmres-ir": 5}
INTEGRATIONS = {"numerical": 0, "analytical": 1}
VELOCITY_INTEGRATIONS = {"follow": 0, "numerical": 1, "analytical": 2}
INITIAL_GUESSES = {"zero": 0, "previous": 1, "extrapolated": 2}
MATRIX_FORMATS = {"auto": 0, "dense": 1, "theta": 2}
VELOCITY_OPERATORS = {"stored": 0, "pairwise": 1}

# Every symbol include/rigidfibers.h declares.
EXPORTS = (
    "rf_config_init", "rf_create", "rf_destroy", "rf_step", "rf_assemble_system", "rf_solve_system",
    "rf_update_velocities", "rf_update_fibers", "rf_advance", "rf_get_state", "rf_get_next_state",
    "rf_get_velocities", "rf_set_state", "rf_get_system", "rf_get_system_noalloc", "rf_system_size", "rf_get_solver_stats",
    "rf_get_timings", "rf_device_bytes", "rf_last_error", "rf_version",
    "rf_server_create", "rf_server_destroy", "rf_server_submit", "rf_server_wait", "rf_server_collect",
    "rf_server_get_stats", "rf_server_clock",
)


class RfConfig(C.Structure):
    _fields_ = [
        ("number_of_fibers", C.c_int32),
        ("number_of_terms_in_force_expansion", C.c_int32),
        ("number_of_quadrature_intervals", C.c_int32),
        ("number_of_quadrature_points_per_interval", C.c_int32),
        ("timestep", C.c_double),
        ("slenderness", C.c_double),
        ("external_force", C.c_double * 3),
        ("solver", C.c_int32),
        ("integration", C.c_int32),
        ("velocity_integration", C.c_int32),
        ("gmres_restart", C.c_int32),
        ("gmres_max_iterations", C.c_int32),
        ("gmres_tolerance", C.c_double),
        ("bicgstab_max_iterations", C.c_int32),
        ("bicgstab_tolerance", C.c_double),
        ("refinement_tolerance", C.c_double),
        ("refinement_max_iterations", C.c_int32),
        ("initial_guess", C.c_int32),
        ("matrix_format", C.c_int32),
        ("velocity_operator", C.c_int32),
        ("inner_tolerance", C.c_double),
        ("reserved0", C.c_int32),
        ("separate_residual", C.c_int32),
        ("confirm_residual", C.c_int32),
        ("reserved", C.c_int32 * 1),
    ]


class RfSolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("relative_residual", C.c_double)]


class RfTimings(C.Structure):
    _fields_ = [
        ("assemble_system", C.c_double),
        ("solve_system", C.c_double),
        ("update_velocities", C.c_double),
        ("update_fibers", C.c_double),
        ("total", C.c_double),
        ("gemv", C.c_double),
        ("gemv_launches", C.c_int32),
        ("gemm_launches", C.c_int32),
        ("gemm", C.c_double),
        ("gemm_flop", C.c_double),
        ("gemv_half", C.c_double),
        ("gemv_half_launches", C.c_int32),
        ("direct_fallback", C.c_int32),
        ("assemble_kernel", C.c_double),
        ("solver_restarts", C.c_int32),
        ("reserved1", C.c_int32),
    ]


class RfServerConfig(C.Structure):
    _fields_ = [
        ("base", RfConfig),
        ("memory_budget", C.c_int64),
        ("max_active", C.c_int32),
        ("eager", C.c_int32),
        ("reserved", C.c_int32 * 4),
    ]


class RfRequestInfo(C.Structure):
    _fields_ = [
        ("id", C.c_int64),
        ("fibers", C.c_int32),
        ("steps", C.c_int32),
        ("status", C.c_int32),
        ("gmres_iterations", C.c_int32),
        ("submitted", C.c_double),
        ("admitted", C.c_double),
        ("finished", C.c_double),
        ("footprint", C.c_int64),
    ]


class RfServerStats(C.Structure):
    _fields_ = [
        ("ticks", C.c_int64),
        ("fiber_steps", C.c_int64),
        ("completed", C.c_int64),
        ("queued", C.c_int64),
        ("active", C.c_int64),
        ("busy_seconds", C.c_double),
        ("peak_bytes", C.c_int64),
        ("memory_budget", C.c_int64),
    ]


class RfError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"rigidfibers error {code}: {message}")
        self.code = code


def load(path: str = LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: the HIP library has not been built "
            "(run `python -c 'import __graft_entry__ as g; g.build()'` or `make -C rigidfibers_amd/csrc`)")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    sim = C.c_void_p
    sigs = {
        "rf_config_init": ([C.POINTER(RfConfig)], C.c_int),
        "rf_create": ([C.POINTER(RfConfig), dp, dp, C.POINTER(C.c_void_p)], C.c_int),
        "rf_destroy": ([sim], None),
        "rf_step": ([sim, C.c_int64], C.c_int),
        "rf_assemble_system": ([sim], C.c_int),
        "rf_solve_system": ([sim], C.c_int),
        "rf_update_velocities": ([sim], C.c_int),
        "rf_update_fibers": ([sim, C.c_int32], C.c_int),
        "rf_advance": ([sim], C.c_int),
        "rf_get_state": ([sim, dp, dp], C.c_int),
        "rf_get_next_state": ([sim, dp, dp], C.c_int),
        "rf_get_velocities": ([sim, dp, dp], C.c_int),
        "rf_set_state": ([sim, dp, dp, C.c_int32, dp, dp, dp, dp], C.c_int),
        "rf_get_system": ([sim, C.POINTER(C.c_float), dp, dp], C.c_int),
        "rf_get_system_noalloc": ([sim, C.POINTER(C.c_float), dp, dp], C.c_int),
        "rf_system_size": ([sim], C.c_int64),
        "rf_get_solver_stats": ([sim, C.POINTER(RfSolverStats)], C.c_int),
        "rf_get_timings": ([sim, C.POINTER(RfTimings)], C.c_int),
        "rf_device_bytes": ([sim], C.c_int64),
        "rf_last_error": ([], C.c_char_p),
        "rf_version": ([], C.c_char_p),
        "rf_server_create": ([C.POINTER(RfServerConfig), C.POINTER(C.c_void_p)], C.c_int),
        "rf_server_destroy": ([C.c_void_p], None),
        "rf_server_submit": ([C.c_void_p, C.c_int32, dp, dp, C.c_int32, C.POINTER(C.c_int64)], C.c_int),
        "rf_server_wait": ([C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_int32)], C.c_int),
        "rf_server_collect": ([C.c_void_p, C.c_int64, dp, dp, dp, C.POINTER(RfRequestInfo)], C.c_int),
        "rf_server_get_stats": ([C.c_void_p, C.POINTER(RfServerStats)], C.c_int),
        "rf_server_clock": ([C.c_void_p], C.c_double),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def check(code, allow_warning=False):
    if code == RF_OK or (allow_warning and code == RF_WARN_NOT_CONVERGED):
        return code
    raise RfError(code, lib().rf_last_error().decode())
