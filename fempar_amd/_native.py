"""ctypes binding of libfemhip.so (the C-ABI declared in include/femhip.h).

The library is loaded from the package directory only; there is no CPU
fallback: without the built HIP library (or without a GPU at context creation)
every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfemhip.so")

OK, ERR_INVALID, ERR_HIP, ERR_OOM, ERR_NOT_CONVERGED = 0, 1, 2, 3, 4
ERR_ZERO_DIAGONAL, ERR_STATE, ERR_ADMISSION, ERR_NO_DEVICE = 5, 6, 7, 8

BC_ALL_BOUNDARY, BC_X0_FACE, BC_NONE = 0, 1, 2
PHYS_POISSON, PHYS_ELASTICITY = 0, 1
K_CONST, K_SIN3, K_CELL, K_QPOINT = 0, 1, 2, 3
F_CONST, F_MANUFACTURED, F_CELL_SIN, F_QPOINT = 0, 1, 2, 3
G_ZERO, G_LINEAR = 0, 1
PREC_IDENTITY, PREC_JACOBI = 0, 1
PRECISION_FP64, PRECISION_FP32_REFINED, PRECISION_FP32_STRICT = 0, 1, 2

# iterative_solver_parameters.f90:37-46, in the C-ABI's numbering
STOPPING_CRITERIA = ["RES_NRMGIVEN_RHS_NRMGIVEN", "RES_NRMGIVEN_RES_NRMGIVEN", "DELTA_RHS", "DELTA_DELTA",
                     "RES_RES", "RES_RHS", "DELTA_RHS_AND_RES_RES", "DELTA_RHS_AND_RES_RHS",
                     "DELTA_DELTA_AND_RES_RES", "DELTA_DELTA_AND_RES_RHS"]

EXPORTS = [
    "femhip_ctx_create", "femhip_ctx_destroy", "femhip_last_error",
    "femhip_space_create_structured", "femhip_space_create_from_dofs", "femhip_space_info",
    "femhip_space_get_dofs", "femhip_space_interpolate_dirichlet", "femhip_space_destroy",
    "femhip_space_set_vertex_coords", "femhip_space_quadrature_points",
    "femhip_operator_create", "femhip_operator_info", "femhip_operator_get_pattern",
    "femhip_operator_set_physics", "femhip_operator_compute", "femhip_operator_get_values",
    "femhip_operator_apply", "femhip_operator_apply_noalloc", "femhip_operator_destroy",
    "femhip_solver_create", "femhip_solver_solve", "femhip_solver_get_solution", "femhip_solver_destroy",
    "femhip_solver_profile_spmv", "femhip_solver_profile_cg", "femhip_solver_error_norms",
    "femhip_solver_get_convergence_history",
    "femhip_batch_create", "femhip_batch_assemble", "femhip_batch_solve", "femhip_batch_get_solution",
    "femhip_batch_get_problem_matrix", "femhip_batch_autotune", "femhip_batch_destroy",
    "femhip_serve_create", "femhip_serve_estimate_bytes", "femhip_serve_submit", "femhip_serve_wait",
    "femhip_serve_get_solution", "femhip_serve_release", "femhip_serve_stats", "femhip_serve_now",
    "femhip_serve_destroy",
    "femhip_block_operator_create", "femhip_block_operator_info", "femhip_block_operator_get_dofs",
    "femhip_block_operator_get_block_pattern", "femhip_block_operator_set_physics", "femhip_block_operator_compute",
    "femhip_block_operator_get_block_values", "femhip_block_operator_get_rhs", "femhip_block_operator_apply",
    "femhip_block_operator_solve", "femhip_block_operator_destroy",
]

MAX_FIELDS = 4


class BlockLayout(C.Structure):
    _fields_ = [("num_fields", C.c_int32), ("fields", C.c_void_p * MAX_FIELDS),
                ("field_blocks", C.c_int32 * MAX_FIELDS), ("field_coupling", C.c_int32 * (MAX_FIELDS * MAX_FIELDS))]


class MultiFieldPhysics(C.Structure):
    _fields_ = [("kappa", C.c_double * MAX_FIELDS), ("sigma", C.c_double), ("f", C.c_double * MAX_FIELDS)]


class BatchProblem(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n", C.c_int32), ("k", C.c_double), ("f0", C.c_double)]


class BatchResult(C.Structure):
    _fields_ = [("num_iterations", C.c_int32), ("updates", C.c_int32), ("did_converge", C.c_int32),
                ("num_dofs", C.c_int32), ("true_residual", C.c_double), ("rhs_norm", C.c_double),
                ("latency_us", C.c_double)]


class FemhipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"femhip error {code}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n", C.c_int32 * 3), ("limits", C.c_double * 6), ("order", C.c_int32),
                ("num_components", C.c_int32), ("dirichlet", C.c_int32)]


class Physics(C.Structure):
    _fields_ = [("kind", C.c_int32), ("k_kind", C.c_int32), ("k0", C.c_double),
                ("k_cell", C.POINTER(C.c_double)), ("f_kind", C.c_int32), ("f0", C.c_double),
                ("g_kind", C.c_int32), ("E", C.c_double), ("nu", C.c_double), ("body", C.c_double * 3),
                ("k_qp", C.POINTER(C.c_double)), ("f_qp", C.POINTER(C.c_double))]


class CgParams(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("max_num_iterations", C.c_int32),
                ("stopping_criterion", C.c_int32), ("preconditioner", C.c_int32), ("precision", C.c_int32),
                ("inner_rtol", C.c_double), ("max_outer", C.c_int32)]


class CgResult(C.Structure):
    _fields_ = [("num_iterations", C.c_int32), ("did_converge", C.c_int32), ("outer_iterations", C.c_int32),
                ("status", C.c_int32), ("error_estimate", C.c_double), ("rhs_convergence_test", C.c_double),
                ("true_residual", C.c_double), ("rhs_norm", C.c_double), ("solve_ms", C.c_double)]


class ErrorNorms(C.Structure):
    _fields_ = [("l2_norm", C.c_double), ("h1_seminorm", C.c_double), ("h1_norm", C.c_double),
                ("linfty_norm", C.c_double)]


EXACT_SIN, EXACT_LINEAR = 0, 1


class ServeParams(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("max_active", C.c_int32), ("setup_threads", C.c_int32),
                ("cg", CgParams), ("admission", C.c_int32), ("aging_s", C.c_double)]


class ServeRequest(C.Structure):
    _fields_ = [("n", C.c_int32), ("E", C.c_double), ("nu", C.c_double), ("body", C.c_double * 3),
                ("arrival_s", C.c_double)]


class ServeResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("num_iterations", C.c_int32), ("outer_iterations", C.c_int32),
                ("did_converge", C.c_int32), ("num_dofs", C.c_int32), ("pad", C.c_int32),
                ("true_residual", C.c_double), ("rhs_norm", C.c_double), ("queue_ms", C.c_double),
                ("setup_ms", C.c_double), ("solve_ms", C.c_double), ("latency_ms", C.c_double),
                ("hbm_bytes", C.c_double), ("hbm_estimate", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("queued", C.c_int32), ("active", C.c_int32), ("peak_active", C.c_int32),
                ("running_peak", C.c_int32), ("hbm_in_use", C.c_double), ("hbm_peak", C.c_double),
                ("arena_bytes", C.c_double), ("arena_fallbacks", C.c_int64)]


_lib = None


def load():
    """Load libfemhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FemhipError(ERR_INVALID, f"{LIB_PATH} is missing: run fempar_amd.build.build() (no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    pP = C.POINTER(C.c_void_p)
    sig = {
        "femhip_ctx_create": [I32, pP], "femhip_ctx_destroy": [P], "femhip_last_error": [],
        "femhip_space_create_structured": [P, C.POINTER(MeshDesc), pP],
        "femhip_space_create_from_dofs": [P, C.POINTER(MeshDesc), P, I32, I32, pP],
        "femhip_space_info": [P, C.POINTER(I64), C.POINTER(I32), C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)],
        "femhip_space_get_dofs": [P, P], "femhip_space_interpolate_dirichlet": [P, C.POINTER(Physics), P],
        "femhip_space_destroy": [P],
        "femhip_space_set_vertex_coords": [P, P], "femhip_space_quadrature_points": [P, P],
        "femhip_operator_create": [P, P, I32, pP],
        "femhip_operator_info": [P, C.POINTER(I32), C.POINTER(I64), C.POINTER(I64)],
        "femhip_operator_get_pattern": [P, P, P], "femhip_operator_set_physics": [P, C.POINTER(Physics)],
        "femhip_operator_compute": [P, P, C.POINTER(D)], "femhip_operator_get_values": [P, P, P],
        "femhip_operator_apply": [P, P, P], "femhip_operator_apply_noalloc": [P, P, P],
        "femhip_operator_destroy": [P],
        "femhip_solver_create": [P, P, C.POINTER(CgParams), pP],
        "femhip_solver_solve": [P, P, P, C.POINTER(CgResult)], "femhip_solver_get_solution": [P, P],
        "femhip_solver_destroy": [P],
        "femhip_solver_profile_spmv": [P, I32, C.POINTER(D), C.POINTER(D)],
        "femhip_solver_profile_cg": [P, C.POINTER(D), C.POINTER(I32), C.POINTER(D)],
        "femhip_solver_error_norms": [P, I32, C.POINTER(ErrorNorms)],
        "femhip_solver_get_convergence_history": [P, I32, P, C.POINTER(I32)],
        "femhip_batch_create": [P, I32, C.POINTER(BatchProblem), C.POINTER(CgParams), pP],
        "femhip_batch_assemble": [P, C.POINTER(D)],
        "femhip_batch_solve": [P, C.POINTER(BatchResult), C.POINTER(D)],
        "femhip_batch_get_solution": [P, I32, P],
        "femhip_batch_get_problem_matrix": [P, I32, P, P, P, P, C.POINTER(I64)],
        "femhip_batch_autotune": [P, C.POINTER(I32)],
        "femhip_batch_destroy": [P],
        "femhip_serve_create": [P, C.POINTER(ServeParams), pP],
        "femhip_serve_estimate_bytes": [I32, I32],
        "femhip_serve_submit": [P, C.POINTER(ServeRequest), C.POINTER(I64)],
        "femhip_serve_wait": [P, I64, D, C.POINTER(ServeResult)],
        "femhip_serve_get_solution": [P, I64, P],
        "femhip_serve_release": [P, I64],
        "femhip_serve_stats": [P, C.POINTER(ServeStats)],
        "femhip_serve_now": [P],
        "femhip_serve_destroy": [P],
        "femhip_block_operator_create": [P, C.POINTER(BlockLayout), I32, pP],
        "femhip_block_operator_info": [P, C.POINTER(I32), P, P],
        "femhip_block_operator_get_dofs": [P, I32, P],
        "femhip_block_operator_get_block_pattern": [P, I32, I32, P, P],
        "femhip_block_operator_set_physics": [P, C.POINTER(MultiFieldPhysics)],
        "femhip_block_operator_compute": [P, C.POINTER(D)],
        "femhip_block_operator_get_block_values": [P, I32, I32, P],
        "femhip_block_operator_get_rhs": [P, I32, P],
        "femhip_block_operator_apply": [P, P, P],
        "femhip_block_operator_solve": [P, C.POINTER(CgParams), P, P, C.POINTER(CgResult)],
        "femhip_block_operator_destroy": [P],
    }
    restype = {"femhip_last_error": C.c_char_p, "femhip_serve_estimate_bytes": C.c_double,
               "femhip_serve_now": C.c_double}
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = restype.get(name, C.c_int)
    _lib = L
    return L


def check(rc: int):
    if rc != OK:
        msg = load().femhip_last_error().decode(errors="replace")
        raise FemhipError(rc, msg)


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
