"""ctypes mirror of include/sdx_fem.h (structs, enums, defaults).

The field order and types must match the C header exactly; the C library
checks sizeof of each struct against the values passed by
``sdx_abi_struct_sizes`` at load time.
"""
from __future__ import annotations

import ctypes as C

SDX_OK = 0
SDX_ERR_INVALID_ARGUMENT = -1
SDX_ERR_DEVICE = -2
SDX_ERR_OUT_OF_MEMORY = -3
SDX_ERR_STATE = -4
SDX_ERR_UNSUPPORTED = -5

CONV_NONE, CONV_CONVERGED, CONV_STOPPED, CONV_DIVERGED = 0, 1, 2, 3
NORM_RESIDUAL_L2, NORM_PRECONDITIONED_RESIDUAL_L2, NORM_RESIDUAL_PRECONDITIONER_INDUCED = 0, 1, 2
PREC_NONE, PREC_JACOBI, PREC_BLOCK_JACOBI = 0, 1, 2
LS_NONE, LS_SIMPLE, LS_ARMIJO, LS_WOLFE_WEAK, LS_WOLFE_STRONG, LS_RESIDUAL_NORM, LS_ARMIJO_OR_RESIDUAL_NORM = range(7)
PSD_NEVER, PSD_ALWAYS, PSD_IF_FAIL_RETRY, PSD_IF_FAIL_ALWAYS = 0, 1, 2, 3
LINTOL_CONSTANT, LINTOL_EW1, LINTOL_EW2, LINTOL_EW3 = 0, 1, 2, 3
PREC_MODE_FP32, PREC_MODE_MIXED, PREC_MODE_FP64 = 0, 1, 2
WARM_START_ZERO, WARM_START_LEARNED, WARM_START_HISTORY = 0, 1, 2
BSR_BLOCK_ROWMAJOR, BSR_REF_ROW_STRIP = 0, 1

FLOAT32_EPS = 1.1920928955078125e-07


class KrylovParams(C.Structure):
    _fields_ = [
        ("abs_tol", C.c_double),
        ("rel_tol", C.c_double),
        ("rel_div_tol", C.c_double),
        ("max_iter", C.c_int32),
        ("norm_type", C.c_int32),
        ("preconditioner", C.c_int32),
        ("abort_if_not_spd", C.c_int32),
        ("initial_guess_zero", C.c_int32),
        ("precision_mode", C.c_int32),
        ("refresh_every", C.c_int32),
        ("reserved", C.c_int32),
    ]


class LinearStatus(C.Structure):
    _fields_ = [
        ("num_iter_done", C.c_int32),
        ("convergence", C.c_int32),
        ("residual_norm", C.c_double),
        ("relative_residual_norm", C.c_double),
    ]


class SolverParams(C.Structure):
    _fields_ = [
        ("newton_max_iter", C.c_int32),
        ("psd_proj_mode", C.c_int32),
        ("newton_abs_tol", C.c_double),
        ("newton_rel_tol", C.c_double),
        ("rel_step_tol", C.c_double),
        ("explosion_control", C.c_int32),
        ("line_search_type", C.c_int32),
        ("abs_div_tol", C.c_double),
        ("rel_div_tol", C.c_double),
        ("line_search_max_iter", C.c_int32),
        ("linear_tolerance_strategy", C.c_int32),
        ("line_search_alpha", C.c_double),
        ("line_search_wolfe1", C.c_double),
        ("line_search_wolfe2", C.c_double),
        ("line_search_max_rel_increase", C.c_double),
        ("gradient_descent_fallback", C.c_int32),
        ("reserved0", C.c_int32),
        ("linear", KrylovParams),
        ("gravity", C.c_double * 3),
        ("warm_start", C.c_int32),
        ("schedule", C.c_int32),
    ]


class Schedule(C.Structure):
    _fields_ = [
        ("cg_burst", C.c_int32),
        ("use_graphs", C.c_int32),
        ("spmv_lanes", C.c_int32),
        ("element_block", C.c_int32),
        ("active_lists", C.c_int32),
        ("cg_groups", C.c_int32),
        ("group_iters", C.c_int32),
        ("fused_tail", C.c_int32),
        ("asm_groups", C.c_int32),
        ("fused_asm", C.c_int32),
        ("env_cg", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class TrainParams(C.Structure):
    _fields_ = [
        ("epochs", C.c_int32),
        ("batch", C.c_int32),
        ("lr", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("seed", C.c_uint64),
    ]


MLP_FEATURES, MLP_ROW, MLP_OUT, MLP_PARAMS = 15, 16, 3, 35459


class StepStats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int32),
        ("linear_iters", C.c_int32),
        ("line_search_iters", C.c_int32),
        ("convergence", C.c_int32),
        ("linear_convergence", C.c_int32),
        ("warm_start_accepted", C.c_int32),
        ("warm_start_tried", C.c_int32),
        ("fp64_refreshes", C.c_int32),
        ("warm_start_network", C.c_int32),
        ("reserved0", C.c_int32),
        ("res_norm0", C.c_double),
        ("res_norm", C.c_double),
        ("linear_rel_res", C.c_double),
        ("linear_true_rel_res", C.c_double),
        ("warm_start_rel_res", C.c_double),
        ("warm_start_history_rel_res", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SoftParams(C.Structure):
    _fields_ = [
        ("youngs_modulus", C.c_double),
        ("poisson_ratio", C.c_double),
        ("density", C.c_double),
        ("mass_damping", C.c_double),
        ("stiffness_damping", C.c_double),
        ("penalty_coefficient", C.c_double),
        ("penalty_smoothing_half_distance", C.c_double),
        ("penalty_threshold", C.c_double),
        ("max_alignment_normals", C.c_double),
        ("coulomb_friction_coefficient", C.c_double),
        ("viscous_friction_coefficient", C.c_double),
        ("friction_falloff_vel", C.c_double),
        ("normal_viscous_damping_coefficient", C.c_double),
        ("stiffness_damping_geometric_term", C.c_int32),
        ("reserved", C.c_int32),
    ]


def default_krylov_params() -> KrylovParams:
    """LinearSolverParams defaults (linear_solver_params.h:318-345) with CG + block Jacobi."""
    return KrylovParams(abs_tol=1e-9, rel_tol=1e-5, rel_div_tol=1e10, max_iter=500,
                        norm_type=NORM_PRECONDITIONED_RESIDUAL_L2, preconditioner=PREC_BLOCK_JACOBI,
                        abort_if_not_spd=0, initial_guess_zero=1, precision_mode=PREC_MODE_FP32,
                        refresh_every=0, reserved=0)


def default_solver_params() -> SolverParams:
    """NonLinearSolverParams defaults (nonlinear_solver_params.h:322-445)."""
    p = SolverParams()
    p.newton_max_iter = 4
    p.psd_proj_mode = PSD_ALWAYS
    p.newton_abs_tol = 1e-3
    p.newton_rel_tol = 1e-6
    p.rel_step_tol = 10.0 * FLOAT32_EPS
    p.explosion_control = 1
    p.line_search_type = LS_RESIDUAL_NORM
    p.abs_div_tol = 1e9
    p.rel_div_tol = 1e4
    p.line_search_max_iter = 4
    p.linear_tolerance_strategy = LINTOL_EW2
    p.line_search_alpha = 0.5
    p.line_search_wolfe1 = 1e-4
    p.line_search_wolfe2 = 0.9
    p.line_search_max_rel_increase = 0.0
    p.gradient_descent_fallback = 0
    p.linear = default_krylov_params()
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.8, 0.0
    p.warm_start = WARM_START_ZERO
    p.schedule = 0
    return p


def pinned_solver_params(newton_iters: int = 1, lin_rel_tol: float = 1e-6, lin_max_iter: int = 4000,
                         precision_mode: int = PREC_MODE_FP32, refresh_every: int = 0) -> SolverParams:
    """The parameter set both sides solve with (CG + BlockJacobi, ResidualL2,
    constant linear tolerance, fixed Newton iteration count, abs/rel Newton
    tolerances 0, PSD projection always, explosion control on)."""
    p = default_solver_params()
    p.linear_tolerance_strategy = LINTOL_CONSTANT
    p.newton_max_iter = newton_iters
    p.newton_abs_tol = 0.0
    p.newton_rel_tol = 0.0
    p.linear.norm_type = NORM_RESIDUAL_L2
    p.linear.rel_tol = lin_rel_tol
    p.linear.abs_tol = 0.0
    p.linear.max_iter = lin_max_iter
    p.linear.precision_mode = precision_mode
    p.linear.refresh_every = refresh_every
    return p


def default_soft_params(E: float = 1e5, nu: float = 0.4, density: float = 1000.0,
                        penalty: float = 1e6, mass_damping: float = 0.0, coulomb: float = 0.0,
                        viscous: float = 0.0, falloff_vel: float = 0.01,
                        normal_damping: float = 0.0, stiffness_damping: float = 0.0,
                        stiffness_damping_geometric: int = 0) -> SoftParams:
    """Soft material + ContactParams (material_params.h:77-105, contact_params.h:96-225) of the
    pinned workloads: the reference's smoothing, threshold and alignment defaults, with penalty
    1e6 and no friction unless asked for (the reference defaults to Coulomb 0.5)."""
    return SoftParams(youngs_modulus=E, poisson_ratio=nu, density=density, mass_damping=mass_damping,
                      stiffness_damping=stiffness_damping, stiffness_damping_geometric_term=stiffness_damping_geometric,
                      reserved=0, penalty_coefficient=penalty, penalty_smoothing_half_distance=0.005,
                      penalty_threshold=0.001, max_alignment_normals=0.0,
                      coulomb_friction_coefficient=coulomb, viscous_friction_coefficient=viscous,
                      friction_falloff_vel=falloff_vel, normal_viscous_damping_coefficient=normal_damping)

# Component kinds of a captured state (SDX_STATE_*), for partial restores.
STATE_DISPLACEMENTS = 1
STATE_VELOCITIES = 2
STATE_WARM_START = 4
STATE_COLLIDER_TARGETS = 8
