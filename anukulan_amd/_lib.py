"""ctypes binding of libanukulan_hip.so (declarations in include/anukulan_hip.h).

The library is loaded from this package's directory only. There is no
fallback: if the HIP library is missing, importing this module raises, so no
caller can silently run something else in its place.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libanukulan_hip.so")

DP = POINTER(c_double)
IP = POINTER(c_int32)


class Csr(ctypes.Structure):
    _fields_ = [("id", c_uint64), ("rows", c_int32), ("cols", c_int32),
                ("start", IP), ("index", IP), ("value", DP)]


class ConvProblem(ctypes.Structure):
    _fields_ = [("c", DP), ("lower", DP), ("upper", DP), ("q", DP),
                ("row_scale", DP), ("col_scale", DP), ("num_equalities", c_int32)]


class ConvTerms(ctypes.Structure):
    _fields_ = [("primal_sq", c_double), ("primal_inf", c_double), ("q_dot_y", c_double),
                ("dual_sq", c_double), ("dual_inf", c_double), ("bound_term", c_double),
                ("c_dot_x", c_double)]


class Model(ctypes.Structure):
    _fields_ = [("rows", c_int32), ("cols", c_int32), ("nnz", c_int64),
                ("row_index", IP), ("col_index", IP), ("value", DP),
                ("objective", DP), ("row_lower", DP), ("row_upper", DP),
                ("col_lower", DP), ("col_upper", DP), ("maximize", c_int32),
                ("objective_offset", c_double)]


class MpsOptions(ctypes.Structure):
    _fields_ = [("format", c_int32), ("negative_up_bound", c_int32),
                ("infinity_threshold", c_double)]


class MpsInfo(ctypes.Structure):
    _fields_ = [("rows", c_int32), ("cols", c_int32), ("nnz", c_int64), ("hessian_nnz", c_int64),
                ("maximize", c_int32), ("fixed_format", c_int32), ("objective_offset", c_double),
                ("free_rows_dropped", c_int32), ("warnings", c_int32),
                ("objective_nonzeros", c_int64)]


class LpInfo(ctypes.Structure):
    _fields_ = [("rows", c_int32), ("cols", c_int32), ("nnz", c_int32),
                ("num_equalities", c_int32), ("objective_scale", c_double),
                ("objective_offset", c_double)]


class PdhgOptions(ctypes.Structure):
    _fields_ = [
        ("tolerance", c_double), ("gap_tolerance", c_double), ("max_iterations", c_int32),
        ("time_limit_seconds", c_double), ("termination_check_frequency", c_int32),
        ("absolute_tolerance", c_double), ("require_inf_norm_termination", c_int32),
        ("adaptive_step_size", c_int32), ("primal_weight_updates", c_int32),
        ("restarts", c_int32), ("halpern", c_int32), ("halpern_restart", c_int32),
        ("halpern_minimum_epoch", c_int32), ("polish_feasibility", c_int32),
        ("polish_first_iteration", c_int32), ("polish_iteration_factor", c_double),
        ("polish_minimum_iterations", c_int32), ("polish_maximum_iterations", c_int32),
        ("polish_on_exit", c_int32), ("reflection", c_double), ("constant_step_size", c_int32),
        ("constant_step_scale", c_double), ("restart_on_fixed_point", c_int32),
        ("pid_primal_weight", c_int32), ("primal_weight_kp", c_double),
        ("primal_weight_ki", c_double), ("primal_weight_kd", c_double),
        ("ruiz_iterations", c_int32), ("pock_chambolle", c_int32), ("reuse_products", c_int32),
        ("detect_infeasibility", c_int32), ("infeasibility_tolerance", c_double),
        ("termination", c_int32), ("warm_x", DP), ("warm_y", DP),
        ("known_matrix_norm", c_double), ("initial_step_size", c_double),
        ("initial_primal_weight", c_double), ("verbose", c_int32), ("log_frequency", c_int32),
        ("stop", IP), ("use_graphs", c_int32), ("product_order", c_int32), ("autotune", c_int32),
        ("batch_waves", c_int32), ("batch_rows", c_int32), ("batch_layout", c_int32),
        ("batch_pack", c_int32), ("multi_order", c_int32), ("multi_stream_loads", c_int32),
        ("small_lp_engine", c_int32),
    ]


class PdhgResidual(ctypes.Structure):
    _fields_ = [(name, c_double) for name in (
        "primal_residual", "dual_residual", "primal_residual_inf", "dual_residual_inf",
        "primal_objective", "dual_objective", "absolute_gap", "relative_primal",
        "relative_dual", "relative_gap", "relative_primal_inf", "relative_dual_inf")]


class PdhgResult(ctypes.Structure):
    _fields_ = [
        ("status", c_int32), ("iterations", c_int32), ("restarts", c_int32),
        ("polish_attempts", c_int32), ("polish_iterations", c_int32), ("polished", c_int32),
        ("objective", c_double), ("solve_seconds", c_double), ("matrix_norm_estimate", c_double),
        ("final_step_size", c_double), ("final_primal_weight", c_double),
        ("phase_scaling", c_double), ("phase_matrix_norm", c_double), ("phase_setup", c_double),
        ("phase_iterating", c_double), ("phase_checks", c_double), ("phase_polishing", c_double),
        ("residual", PdhgResidual), ("message", c_char * 128),
    ]


class NodeResult(ctypes.Structure):
    _fields_ = [("optimal", c_int32), ("proved_infeasible", c_int32), ("bound", c_double),
                ("relaxation", PdhgResult)]


class FinishOptions(ctypes.Structure):
    _fields_ = [("primal_tolerance", c_double), ("dual_tolerance", c_double),
                ("pivot_tolerance", c_double), ("max_iterations", c_int32),
                ("refactor_interval", c_int32), ("crossover", c_int32),
                ("interior_tolerance", c_double), ("crossover_pivot_tolerance", c_double),
                ("max_candidates", c_int32), ("crossover_refactor_interval", c_int32),
                ("dual_weight", c_double), ("ruiz_iterations", c_int32), ("threads", c_int32)]


class FinishResult(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in (
        "status", "iterations", "phase_one_iterations", "refactorizations", "bland_iterations",
        "started_warm", "candidates", "pushed", "rejected_small_pivot", "logicals_remaining",
        "stopped_on_refactorization")] + [
        ("objective", c_double), ("primal_violation", c_double), ("dual_violation", c_double),
        ("seconds", c_double), ("message", c_char * 64)]


class AdjointOptions(ctypes.Structure):
    _fields_ = [("method", c_int32), ("max_iterations", c_int32), ("tolerance", c_double),
                ("active_tolerance", c_double), ("primal_weight", c_double),
                ("active_set", c_int32)]


class AdjointInfo(ctypes.Structure):
    _fields_ = [("iterations", c_int32), ("converged", c_int32), ("residual_primal", c_double),
                ("residual_dual", c_double), ("free_columns", c_int32), ("binding_rows", c_int32),
                ("lower_columns", c_int32), ("upper_columns", c_int32),
                ("basis_completed", c_int32)]


class PresolveRules(ctypes.Structure):
    _fields_ = [("empty_rows", c_int32), ("singleton_rows", c_int32), ("redundant_rows", c_int32),
                ("fixed_columns", c_int32), ("empty_columns", c_int32),
                ("slack_column_singletons", c_int32), ("max_rounds", c_int32),
                ("tolerance", c_double), ("pivot_ratio", c_double)]


class PresolveInfo(ctypes.Structure):
    _fields_ = [("status", c_int32), ("failing_case", c_int32), ("rows", c_int32),
                ("cols", c_int32), ("nnz", c_int64), ("original_rows", c_int32),
                ("original_cols", c_int32), ("original_nnz", c_int64), ("empty_rows", c_int32),
                ("free_rows", c_int32), ("singleton_rows", c_int32), ("redundant_rows", c_int32),
                ("redundant_sides", c_int32), ("fixed_columns", c_int32),
                ("empty_columns", c_int32), ("slack_column_singletons", c_int32),
                ("rounds", c_int32)]


# name -> (restype, argtypes); every function the header declares.
SIGNATURES = {
    "anuk_create": (c_int, [c_int, POINTER(c_void_p)]),
    "anuk_destroy": (None, [c_void_p]),
    "anuk_last_error": (c_char_p, [c_void_p]),
    "anuk_name": (c_char_p, [c_void_p]),
    "anuk_synchronize": (c_int, [c_void_p]),
    "anuk_allocate": (c_int, [c_void_p, c_int32, POINTER(c_void_p)]),
    "anuk_deallocate": (c_int, [c_void_p, c_void_p]),
    "anuk_upload": (c_int, [c_void_p, DP, c_void_p, c_int32]),
    "anuk_download": (c_int, [c_void_p, c_void_p, DP, c_int32]),
    "anuk_fill": (c_int, [c_void_p, c_void_p, c_int32, c_double]),
    "anuk_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_int32]),
    "anuk_prepare": (c_int, [c_void_p, POINTER(Csr)]),
    "anuk_release": (c_int, [c_void_p]),
    "anuk_multiply": (c_int, [c_void_p, POINTER(Csr), c_void_p, c_void_p]),
    "anuk_multiply_transpose": (c_int, [c_void_p, POINTER(Csr), c_void_p, c_void_p]),
    "anuk_dot": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, DP]),
    "anuk_two_norm": (c_int, [c_void_p, c_void_p, c_int32, DP]),
    "anuk_inf_norm": (c_int, [c_void_p, c_void_p, c_int32, DP]),
    "anuk_primal_update": (c_int, [c_void_p, c_int32, c_double, c_double, c_double] + [c_void_p] * 9),
    "anuk_dual_update": (c_int, [c_void_p, c_int32, c_int32, c_double, c_double, c_double] + [c_void_p] * 10),
    "anuk_accumulate": (c_int, [c_void_p, c_int32, c_double, c_void_p, c_void_p]),
    "anuk_scale_into": (c_int, [c_void_p, c_int32, c_double, c_void_p, c_void_p]),
    "anuk_weighted_norm_squared": (c_int, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_double, DP]),
    "anuk_step_size_terms": (c_int, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                                     c_double, DP, DP]),
    "anuk_convergence_terms": (c_int, [c_void_p, c_int32, c_int32, POINTER(ConvProblem), c_void_p,
                                       c_void_p, c_void_p, c_void_p, POINTER(ConvTerms)]),
    "anuk_convergence_terms_noalloc": (c_int, [c_void_p, c_int32, c_int32, POINTER(ConvProblem), c_void_p,
                                       c_void_p, c_void_p, c_void_p, POINTER(ConvTerms)]),
    "anuk_set_profiling": (c_int, [c_void_p, c_int]),
    "anuk_profile_report": (c_int, [c_void_p, c_char_p, c_size_t, POINTER(c_size_t)]),
    "anuk_tuning_report": (c_int, [c_void_p, c_char_p, c_size_t, POINTER(c_size_t)]),
    "anuk_lp_from_model": (c_int, [POINTER(Model), POINTER(c_void_p)]),
    "anuk_lp_from_standard": (c_int, [c_int32, c_int32, c_int32, IP, IP, DP, DP, DP, DP, DP,
                                      c_double, c_double, POINTER(c_void_p)]),
    "anuk_lp_free": (None, [c_void_p]),
    "anuk_lp_info_get": (c_int, [c_void_p, POINTER(LpInfo)]),
    "anuk_lp_export": (c_int, [c_void_p, IP, IP, DP, DP, DP, DP, DP, IP, DP]),
    "anuk_lp_scale": (c_int, [c_void_p, c_int32, c_int32, POINTER(c_void_p), DP, DP]),
    "anuk_lp_scale_device": (c_int, [c_void_p, c_void_p, c_int32, c_int32, POINTER(c_void_p), DP, DP]),
    "anuk_pdhg_default_options": (None, [POINTER(PdhgOptions)]),
    "anuk_solve_pdhg": (c_int, [c_void_p, c_void_p, POINTER(PdhgOptions), POINTER(PdhgResult), DP, DP]),
    "anuk_solve_pdhg_batch": (c_int, [c_void_p, POINTER(c_void_p), c_int32, POINTER(PdhgOptions),
                                      POINTER(PdhgResult), DP, DP]),
    "anuk_solve_pdhg_multi": (c_int, [c_void_p, POINTER(c_void_p), c_int32, POINTER(PdhgOptions),
                                      POINTER(PdhgResult), DP, DP]),
    "anuk_solve_node_relaxations": (c_int, [c_void_p, c_void_p, c_int32, DP, DP, POINTER(PdhgOptions),
                                            POINTER(NodeResult), DP, DP]),
    "anuk_finish_default_options": (None, [POINTER(FinishOptions)]),
    "anuk_exact_finish": (c_int, [POINTER(c_void_p), c_int32, DP, DP, POINTER(FinishOptions),
                                  POINTER(FinishResult), DP, DP, IP]),
    "anuk_exact_finish_basis": (c_int, [POINTER(c_void_p), c_int32, DP, DP, IP, IP,
                                        POINTER(FinishOptions), POINTER(FinishResult), DP, DP, IP,
                                        IP]),
    "anuk_adjoint_default_options": (None, [POINTER(AdjointOptions)]),
    "anuk_pdhg_adjoint": (c_int, [c_void_p, POINTER(c_void_p), c_int32, POINTER(PdhgOptions),
                                  POINTER(AdjointOptions), DP, DP, DP, DP, DP, DP, DP, DP,
                                  POINTER(AdjointInfo)]),
    "anuk_estimate_matrix_norm": (c_int, [c_void_p, c_void_p, c_int32, c_double, DP]),
    "anuk_evaluate_residual": (c_int, [c_void_p, DP, DP, POINTER(PdhgResidual)]),
    "anuk_mps_default_options": (None, [POINTER(MpsOptions)]),
    "anuk_mps_read": (c_int, [c_char_p, POINTER(MpsOptions), POINTER(c_void_p)]),
    "anuk_mps_read_text": (c_int, [c_char_p, c_int64, c_char_p, POINTER(MpsOptions),
                                   POINTER(c_void_p)]),
    "anuk_mps_free": (None, [c_void_p]),
    "anuk_mps_info_get": (c_int, [c_void_p, POINTER(MpsInfo)]),
    "anuk_mps_model": (c_int, [c_void_p, POINTER(Model)]),
    "anuk_mps_matrix": (c_int, [c_void_p, POINTER(IP), POINTER(IP), POINTER(DP)]),
    "anuk_mps_hessian": (c_int, [c_void_p, POINTER(IP), POINTER(IP), POINTER(DP)]),
    "anuk_mps_col_types": (IP, [c_void_p]),
    "anuk_mps_name": (c_char_p, [c_void_p]),
    "anuk_mps_row_name": (c_char_p, [c_void_p, c_int32]),
    "anuk_mps_col_name": (c_char_p, [c_void_p, c_int32]),
    "anuk_mps_warning": (c_char_p, [c_void_p, c_int32]),
    "anuk_mps_write": (c_int, [c_char_p, POINTER(Model), c_char_p, POINTER(c_char_p),
                               POINTER(c_char_p), IP]),
    "anuk_presolve_default_rules": (None, [POINTER(PresolveRules)]),
    "anuk_presolve_stack": (c_int, [c_int32, c_int32, IP, IP, DP, c_int32, DP, DP, DP, DP, DP, DP,
                                    POINTER(PresolveRules), POINTER(c_void_p)]),
    "anuk_presolve_free": (None, [c_void_p]),
    "anuk_presolve_get_info": (c_int, [c_void_p, POINTER(PresolveInfo)]),
    "anuk_presolve_message": (c_char_p, [c_void_p]),
    "anuk_presolve_reduced": (c_int, [c_void_p, IP, IP, DP, DP, DP, DP, DP, DP, DP, IP, IP]),
    "anuk_postsolve": (c_int, [c_void_p, c_int32, DP, DP, DP, DP]),
}


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `make -C anukulan_amd` "
            "(or __graft_entry__.build()); there is no fallback implementation")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()
