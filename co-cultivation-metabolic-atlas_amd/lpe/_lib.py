"""ctypes binding of liblpe.so (declarations mirror include/lpe.h and
include/lpe_reduction.h).

The HIP library is the only compute path: importing this module raises if the
library has not been built, and nothing falls back to CPU arithmetic.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# LPE_LIBRARY selects another build of the same library (the host-sanitizer
# build of csrc/sanitize.sh); the in-tree build is the default.
LIB_PATH = os.environ.get("LPE_LIBRARY") or os.path.join(_HERE, "liblpe.so")

LPE_OK, LPE_EINVAL, LPE_EDEVICE, LPE_ESTATE = 0, -1, -2, -3
PREC_FP64, PREC_FP32 = 0, 1
METHOD_FIXED, METHOD_RESTARTED, METHOD_HALPERN = 0, 1, 2
STATUS_RUNNING, STATUS_OPTIMAL, STATUS_LIMIT = 0, 1, 2
STATUS_PRIMAL_INFEASIBLE, STATUS_DUAL_INFEASIBLE, STATUS_NUMERICAL = 3, 4, 5
STATUS_NAMES = {0: "running", 1: "optimal", 2: "iteration_limit", 3: "primal_infeasible",
                4: "dual_infeasible", 5: "numerical_error"}
NMETRIC = 8
METRIC_NAMES = ("primal_residual", "dual_violation", "relative_kkt_gap", "objective",
                "certificate_passed", "finite", "bound_complementarity_gap",
                "row_complementarity_gap")

EXPORTS = ("lpe_create", "lpe_update", "lpe_reset", "lpe_run", "lpe_set_accepted", "lpe_solve",
           "lpe_certify", "lpe_state", "lpe_polish", "lpe_spmv", "lpe_get_scaling",
           "lpe_get_primal_weight", "lpe_set_primal_weight", "lpe_set_metric", "lpe_set_launch", "lpe_time_iterations", "lpe_time_kernel", "lpe_destroy",
           "lpe_last_error", "lpe_version",
           "lpe_reduction_create", "lpe_reduction_get_dims", "lpe_reduction_get_plan",
           "lpe_reduction_get_pattern", "lpe_reduction_get_lift_map", "lpe_reduction_reduce",
           "lpe_reduction_expand", "lpe_reduction_compress", "lpe_reduction_lift", "lpe_reduction_destroy",
           "lpe_certify_noalloc", "lpe_certify_noalloc_capacity", "lpe_certify_noalloc_need")

RED_INFEASIBLE, RED_OVERFLOW, RED_INHOMOGENEOUS = 1, 2, 4


class LpeDesc(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("neq", C.c_int32), ("batch", C.c_int32),
                ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int32)), ("col_idx", C.POINTER(C.c_int32)),
                ("values", C.POINTER(C.c_double)),
                ("n_var", C.c_int32), ("var_pos", C.POINTER(C.c_int64)),
                ("precision", C.c_int32), ("method", C.c_int32),
                ("step_safety", C.c_double), ("theta", C.c_double), ("primal_weight", C.c_double),
                ("ruiz_iterations", C.c_int32), ("shared_c", C.c_int32)]


class LpeVectors(C.Structure):
    _fields_ = [("c", C.c_void_p), ("rhs", C.c_void_p), ("lower", C.c_void_p),
                ("upper", C.c_void_p), ("var_values", C.c_void_p)]


class LpeSolveOpts(C.Structure):
    _fields_ = [("max_iterations", C.c_int64), ("check_interval", C.c_int32),
                ("polish", C.c_int32), ("polish_kkt", C.c_double),
                ("polish_cg_iterations", C.c_int32), ("use_graph", C.c_int32),
                ("rows_per_wave", C.c_int32), ("lps_per_lane", C.c_int32),
                ("eps_certify_primal", C.c_double), ("eps_certify_dual", C.c_double),
                ("eps_certify_gap", C.c_double), ("stop_after_failed_polishes", C.c_int32),
                ("polish_kkt_start", C.c_double), ("polish_stable_checks", C.c_int32),
                ("first_check_interval", C.c_int32), ("infeasibility_interval", C.c_int32),
                ("eps_infeasible", C.c_double)]


class LpeSolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("certified", C.c_int32),
                ("polish_attempts", C.c_int32), ("restarts", C.c_int32), ("chunks", C.c_int32),
                ("seconds", C.c_double), ("pdhg_ms", C.c_double), ("check_ms", C.c_double),
                ("polish_ms", C.c_double), ("pdhg_kernel_launches", C.c_int64), ("cg_steps", C.c_int64), ("cg_lp_steps", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class LpeBatchProblem(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("neq", C.c_int32), ("batch", C.c_int32),
                ("nnz", C.c_int64),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("col_ptr", C.c_void_p),
                ("row_idx", C.c_void_p), ("t_pos", C.c_void_p), ("values", C.c_void_p),
                ("values_per_lp", C.c_int32), ("c_per_lp", C.c_int32), ("c", C.c_void_p),
                ("rhs", C.c_void_p), ("lower", C.c_void_p), ("upper", C.c_void_p)]


class LpeReductionDesc(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("neq", C.c_int32), ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int32)), ("col_idx", C.POINTER(C.c_int32)),
                ("values", C.POINTER(C.c_double)), ("rhs", C.POINTER(C.c_double)),
                ("n_var", C.c_int32), ("var_pos", C.POINTER(C.c_int64)), ("max_scale_ratio", C.c_double)]


class LpeReductionDims(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("neq", C.c_int32), ("nnz", C.c_int64),
                ("n_var", C.c_int32), ("n_eliminated", C.c_int32), ("lift_nnz", C.c_int64)]


class LpeReduced(C.Structure):
    _fields_ = [("c", C.c_void_p), ("rhs", C.c_void_p), ("lower", C.c_void_p), ("upper", C.c_void_p),
                ("var_values", C.c_void_p), ("lower_witness", C.c_void_p), ("upper_witness", C.c_void_p),
                ("status", C.c_void_p)]


class LpeError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load liblpe.so once; raise ImportError when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(the batched LP engine has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "lpe_create": ([C.POINTER(LpeDesc), C.POINTER(P)], C.c_int),
        "lpe_update": ([P, C.POINTER(LpeVectors), P], C.c_int),
        "lpe_reset": ([P, P, P, P], C.c_int),
        "lpe_run": ([P, I64, I32, I32, P], C.c_int),
        "lpe_set_accepted": ([P, P, P], C.c_int),
        "lpe_solve": ([P, C.POINTER(LpeSolveOpts), C.POINTER(LpeSolveInfo), P], C.c_int),
        "lpe_certify": ([P, P, P, P, D, D, D, P], C.c_int),
        "lpe_state": ([P, P, P, P, P], C.c_int),
        "lpe_polish": ([P, P, I32, P], C.c_int),
        "lpe_spmv": ([P, I32, I32, P, P, P], C.c_int),
        "lpe_get_scaling": ([P, P, P, P, P], C.c_int),
        "lpe_get_primal_weight": ([P, P], C.c_int),
        "lpe_set_primal_weight": ([P, P, P], C.c_int),
        "lpe_set_metric": ([P, P, P, D, P], C.c_int),
        "lpe_set_launch": ([P, I32, I32], C.c_int),
        "lpe_time_iterations": ([P, I64, I32, I32, C.POINTER(D), P], C.c_int),
        "lpe_time_kernel": ([P, I32, I32, C.POINTER(D), P], C.c_int),
        "lpe_destroy": ([P], C.c_int),
        "lpe_last_error": ([C.c_char_p, I64], C.c_int),
        "lpe_version": ([], C.c_char_p),
        "lpe_reduction_create": ([C.POINTER(LpeReductionDesc), C.POINTER(P)], C.c_int),
        "lpe_reduction_get_dims": ([P, C.POINTER(LpeReductionDims)], C.c_int),
        "lpe_reduction_get_plan": ([P, P, P, P, P, P], C.c_int),
        "lpe_reduction_get_pattern": ([P, P, P, P, P], C.c_int),
        "lpe_reduction_get_lift_map": ([P, P, P, P], C.c_int),
        "lpe_reduction_reduce": ([P, I32, I32, C.POINTER(LpeVectors), C.POINTER(LpeReduced), I32, P], C.c_int),
        "lpe_reduction_expand": ([P, I32, P, P, I32, P], C.c_int),
        "lpe_reduction_compress": ([P, I32, P, P, I32, P], C.c_int),
        "lpe_reduction_lift": ([P, P, I32, I32, P, P, P, P, P, P, P, P, I32, P], C.c_int),
        "lpe_reduction_destroy": ([P], C.c_int),
        "lpe_certify_noalloc": ([C.POINTER(LpeBatchProblem), P, P, P, D, D, D, P], C.c_int),
        "lpe_certify_noalloc_capacity": ([], I64),
        "lpe_certify_noalloc_need": ([I32, I32], I64),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    load().lpe_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code: int) -> None:
    """Translate a C status into the reference's exception types."""
    if code == LPE_OK:
        return
    message = last_error()
    if code == LPE_EINVAL:
        raise ValueError(message)
    if code == LPE_ESTATE:
        raise RuntimeError(message)
    raise LpeError(message)
