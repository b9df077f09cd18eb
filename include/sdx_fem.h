/*
 * sdx_fem.h - C ABI of the MI355X implicit soft-body FEM step.
 *
 * Two levels, both plain C: opaque handles, POD parameter structs, host
 * pointers plus sizes, integer status codes.  No C++ or torch types cross
 * this boundary.
 *
 *   L1  sdx_pcg_bsr3_solve: block-Jacobi PCG on a 3x3-block BSR matrix.  It
 *       replaces the CudaCG branch of LinearSolver<T>::Solve
 *       (mochi_core/include/mochi_core/solvers/linear_solver.h:480-484, 615-662).
 *   L2  sdx_world_*: soft tet actors, pinned nodes, kinematic plane/sphere
 *       colliders and one backward-Euler Newton step per call.  It replaces
 *       the soft-body path of Scene::Step
 *       (mochi_physics/src/mochi_scene.cpp:594 -> mochi_step.cpp:547-569 ->
 *       mochi_solve.cpp:1090-1160 -> newton_solver.cpp:603-876).
 *
 * Reference paths are relative to superdex_physics/libraries/mochi/.
 * Every function returns SDX_OK (0) or a negative sdx_status; the text of
 * the last error of the calling thread is returned by sdx_last_error().
 * One sdx_world must be used by one host thread at a time; it owns its
 * device buffers and its HIP stream.
 */
#ifndef SDX_FEM_H_
#define SDX_FEM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t sdx_status;
enum {
  SDX_OK = 0,
  SDX_ERR_INVALID_ARGUMENT = -1,
  SDX_ERR_DEVICE = -2,       /* a HIP call failed (message names it) */
  SDX_ERR_OUT_OF_MEMORY = -3,
  SDX_ERR_STATE = -4,        /* call not allowed in the current world state */
  SDX_ERR_UNSUPPORTED = -5,
};

/* ConvergenceStatus / LinearSolverConvergenceStatus, ordered by severity as in
 * mochi_core/include/mochi_core/solvers/nonlinear_solver_params.h:28-55 and
 * mochi_core/include/mochi_core/solvers/krylov_solver.h:85-116. */
enum {
  SDX_CONV_NONE = 0,
  SDX_CONV_CONVERGED = 1,
  SDX_CONV_STOPPED = 2,
  SDX_CONV_DIVERGED = 3,
};

/* LinearSolverConvergenceNorm (linear_solver_params.h:303-311). */
enum {
  SDX_NORM_RESIDUAL_L2 = 0,
  SDX_NORM_PRECONDITIONED_RESIDUAL_L2 = 1,
  SDX_NORM_RESIDUAL_PRECONDITIONER_INDUCED = 2,
};

/* PreconditionerType subset (linear_solver_params.h:270-283). */
enum {
  SDX_PREC_NONE = 0,
  SDX_PREC_JACOBI = 1,
  SDX_PREC_BLOCK_JACOBI = 2,
};

/* LineSearchType (nonlinear_solver_params.h:137-194; line_search.h).  The merit of
 * Simple, Armijo, WolfeWeak, WolfeStrong and ArmijoOrResidualNorm is the objective of the
 * backward-Euler step (stress, gravity, inertia, mass damping and contact energies). */
enum {
  SDX_LS_NONE = 0,
  SDX_LS_SIMPLE = 1,
  SDX_LS_ARMIJO = 2,
  SDX_LS_WOLFE_WEAK = 3,
  SDX_LS_WOLFE_STRONG = 4,
  SDX_LS_RESIDUAL_NORM = 5,
  SDX_LS_ARMIJO_OR_RESIDUAL_NORM = 6,
};

/* PsdProjectionMode (nonlinear_solver_params.h:69-75).  A Newton iteration whose linear
 * solve does not converge or whose line search finds no improvement is retried once with
 * the projected DResidual in every mode (newton_solver.cpp:736-860); IfFailAlways then
 * keeps projecting for the rest of the solve. */
enum {
  SDX_PSD_NEVER = 0,
  SDX_PSD_ALWAYS = 1,
  SDX_PSD_IF_FAIL_RETRY = 2,
  SDX_PSD_IF_FAIL_ALWAYS = 3,
};

/* LinearToleranceStrategy (nonlinear_solver_params.h:196-255): the relative tolerance of
 * each Newton iteration's linear solve (newton_solver.cpp:426-501).  Constant keeps
 * linear.rel_tol; Eisenstat-Walker choices 1 and 2 start at 0.01 and stay in
 * [100 eps, 0.01]; choice 3 is max(min(1/(2 + k), sqrt(|r|)), linear.rel_tol). */
enum {
  SDX_LINTOL_CONSTANT = 0,
  SDX_LINTOL_EW1 = 1,
  SDX_LINTOL_EW2 = 2,
  SDX_LINTOL_EW3 = 3,
};

/* Arithmetic of the CG (extension: FP32 matvec with FP64 dots + refresh). */
enum {
  SDX_PREC_MODE_FP32 = 0,     /* FP32 vectors, matvec and dots */
  SDX_PREC_MODE_MIXED = 1,    /* FP32 vectors/matvec, FP64 dots, FP64 residual refresh */
  SDX_PREC_MODE_FP64 = 2,     /* everything FP64 (reference SUPERDEX_PRECISION=fp64) */
};

/* CG initial guess (extension).  HISTORY: the least-residual combination of the
 * environment's two previous Newton steps of the same Newton iteration (the learned
 * warm start without the network's guess); LEARNED adds the network's guess to the
 * combination (needs a model, sdx_world_set_warm_start_model). */
enum {
  SDX_WARM_START_ZERO = 0,
  SDX_WARM_START_LEARNED = 1,
  SDX_WARM_START_HISTORY = 2,
};

/* BSR value layout accepted by sdx_pcg_bsr3_solve. */
enum {
  SDX_BSR_BLOCK_ROWMAJOR = 0, /* values[9*k + 3*r + c] = block k, row r, col c */
  SDX_BSR_REF_ROW_STRIP = 1,  /* BlockSparseMatrix strip: each block row is a row-major
                                 3 x (3*nnz_row) matrix (block_sparse_matrix.h:1238) */
};

/* KrylovSolverParams + LinearSolverParams subset (krylov_solver.h:43-82,
 * linear_solver_params.h:318-345). */
typedef struct sdx_krylov_params {
  double abs_tol;
  double rel_tol;
  double rel_div_tol;
  int32_t max_iter;
  int32_t norm_type;        /* SDX_NORM_* */
  int32_t preconditioner;   /* SDX_PREC_* */
  int32_t abort_if_not_spd;
  int32_t initial_guess_zero; /* 1: x is ignored and treated as 0 (InitialGuessHint::Zero) */
  int32_t precision_mode;   /* SDX_PREC_MODE_* */
  int32_t refresh_every;    /* MIXED: recompute r = b - A x in FP64 every k iterations (0: never) */
  int32_t reserved;
} sdx_krylov_params;

/* LinearSolverStatus (krylov_solver.h:106-116). */
typedef struct sdx_linear_status {
  int32_t num_iter_done;
  int32_t convergence;       /* SDX_CONV_* */
  double residual_norm;
  double relative_residual_norm;
} sdx_linear_status;

/* SolverParams (mochi_physics/include/mochi_physics/cpp_api/mochi_structs.h:70-81)
 * mapped the way GetIslandNewtonParams does (mochi_physics/src/mochi_simulation.cpp:34-135),
 * plus the extensions at the end. */
typedef struct sdx_solver_params {
  /* NonLinearSolverParams (nonlinear_solver_params.h:322-445) */
  int32_t newton_max_iter;        /* default 4 */
  int32_t psd_proj_mode;          /* SDX_PSD_* (default Always) */
  double newton_abs_tol;          /* default 1e-3 */
  double newton_rel_tol;          /* default 1e-6 */
  double rel_step_tol;            /* default 10*eps(real) */
  int32_t explosion_control;      /* default 1 */
  int32_t line_search_type;       /* SDX_LS_* (default ResidualNorm) */
  double abs_div_tol;             /* default 1e9 */
  double rel_div_tol;             /* default 1e4 */
  int32_t line_search_max_iter;   /* default 4 */
  int32_t linear_tolerance_strategy; /* SDX_LINTOL_* (default EisenstatWalker2) */
  double line_search_alpha;       /* default 0.5 */
  double line_search_wolfe1;      /* default 1e-4: Armijo and Wolfe sufficient decrease */
  double line_search_wolfe2;      /* default 0.9: Wolfe curvature, in (wolfe1, 1) */
  double line_search_max_rel_increase; /* default 0: Simple accepts merit <= m0 + |m0| x this */
  int32_t gradient_descent_fallback; /* default 0: after a failed last try, step along -residual */
  int32_t reserved0;
  /* LinearSolverParams */
  sdx_krylov_params linear;
  /* scene */
  double gravity[3];              /* default (0, -9.8, 0) */
  /* extensions */
  int32_t warm_start;             /* SDX_WARM_START_* */
  int32_t schedule;               /* -1: autotune the world's schedule at its first step (sdx_autotune);
                                     0: use the world's schedule as set (default) */
} sdx_solver_params;

/* Per-actor result of one step: NewtonSolverStatus fields that
 * StageSolverStats drops, kept here (newton_solver_status.h:40-90). */
typedef struct sdx_step_stats {
  int32_t newton_iters;
  int32_t linear_iters;           /* totalNumLinearIterDone */
  int32_t line_search_iters;      /* totalNumLSIterDone */
  int32_t convergence;            /* SDX_CONV_* of the stage */
  int32_t linear_convergence;     /* of the last linear solve */
  int32_t warm_start_accepted;    /* learned warm starts kept / tried */
  int32_t warm_start_tried;
  int32_t fp64_refreshes;
  int32_t warm_start_network;     /* 1 when the network ran for the last warm start (its gate runs
                                     the previous Newton steps alone for 8 steps of a Newton
                                     iteration index when the network's guess left more than half
                                     of the residual those leave, geometric mean over the
                                     environments); 0 otherwise */
  int32_t reserved0;
  double res_norm0;
  double res_norm;
  double linear_rel_res;          /* last linear solve ||r||/||b|| (as the stop test measured it) */
  double linear_true_rel_res;     /* last linear solve ||b - A x||/||b|| recomputed in FP64 */
  double warm_start_rel_res;      /* ||b - A x0||/||b|| of the last learned warm start x0 (the
                                     combination of the network's guess and the two previous
                                     Newton steps; 0: none tried) */
  double warm_start_history_rel_res; /* the same for the previous Newton steps alone (what the
                                     start leaves without the network; 0: none tried) */
} sdx_step_stats;

/* SoftActorParams subset (cpp_api/mochi_structs.h:174-197). */
typedef struct sdx_soft_params {
  double youngs_modulus;          /* homogeneous; per-element fields via sdx_set_element_materials */
  double poisson_ratio;
  double density;
  double mass_damping;            /* massDampingCoefficient [1/s] */
  double stiffness_damping;       /* stiffnessDampingCoefficient beta [s] (material_params.h:95-105):
                                     viscous PK2 S = (beta/dt) C0 : (E(F) - E(F_stage_start)), C0
                                     the material's rest stiffness (fem_stress_damping.h); default 0 */
  /* ContactParams (contact/contact_params.h:96-160) */
  double penalty_coefficient;     /* [Pa/m] */
  double penalty_smoothing_half_distance; /* default 0.005 m */
  double penalty_threshold;       /* default 0.001 m */
  double max_alignment_normals;   /* default 0 */
  /* Contact dissipation (contact_params.h:170-225), evaluated as the reference's default
   * ExperimentalEvalParams do (eval_params.h:73-140): C1-regularised Coulomb friction along the
   * collider normal's tangent plane, viscous friction and normal damping, all proportional to the
   * penalty force at the sample's stage-start distance and faded by normal alignment, with the
   * fitted (PSD) friction Hessian.  The colliders carry no contact parameters of their own: these
   * are the pair's values (the reference's geometric mean of equal coefficients). */
  double coulomb_friction_coefficient;       /* default 0.5 */
  double viscous_friction_coefficient;       /* [s/m], default 0 */
  double friction_falloff_vel;               /* [m/s], default 0.01 */
  double normal_viscous_damping_coefficient; /* [s/m], default 0 */
  int32_t stiffness_damping_geometric_term;  /* stiffnessDampingIncludeGeometricTerm (material_params.h:
                                                107-120): 1 adds the geometric block of the viscous
                                                tangent (exact Newton), 0 (default) keeps the
                                                material block only */
  int32_t reserved;
} sdx_soft_params;

typedef struct sdx_world sdx_world;

/* Fill a struct with the reference defaults. */
void sdx_default_solver_params(sdx_solver_params* out);
void sdx_default_soft_params(sdx_soft_params* out);
void sdx_default_krylov(sdx_krylov_params* out);
const char* sdx_last_error(void);
/* sizeof of the ABI structs (0 krylov, 1 linear status, 2 solver params,
 * 3 step stats, 4 soft params) so bindings can check their layouts. */
int32_t sdx_abi_struct_size(int32_t which);

/* ---------------- L1: linear solve (CudaIterativeSolve seam) ---------------- */

/* Solve A x = b with PCG (Polak-Ribiere beta, pcg.h:70-238) on device `device`.
 * A: n_block_rows x n_block_rows, nnzb blocks, host arrays; block_row_ptr has
 * n_block_rows+1 entries, block_col_idx nnzb entries, values 9*nnzb reals
 * laid out as `layout`.  b, x: 3*n_block_rows doubles (x is the initial guess
 * unless params->initial_guess_zero).  Replaces
 * LinearSolver<T>::CudaIterativeSolve<3> (linear_solver.h:615-662). */
sdx_status sdx_pcg_bsr3_solve(int32_t device, int32_t n_block_rows, int32_t nnzb,
                              const int32_t* block_row_ptr, const int32_t* block_col_idx,
                              const double* values, int32_t layout, const double* b, double* x,
                              const sdx_krylov_params* params, sdx_linear_status* out);

/* The same solve on float arrays: the reference's default build (real = float,
 * mochi_core/include/mochi_core/mochi_config.h:106-116) hands
 * CudaIterativeSolve<3> a BlockSparseMatrixView<float const, 3> and float b, x
 * (linear_solver.h:615-662), which bind here without a conversion pass.  The
 * arithmetic is still params->precision_mode (FP64 mode widens on the device
 * side of the copy). */
sdx_status sdx_pcg_bsr3_solve_f32(int32_t device, int32_t n_block_rows, int32_t nnzb,
                                  const int32_t* block_row_ptr, const int32_t* block_col_idx,
                                  const float* values, int32_t layout, const float* b, float* x,
                                  const sdx_krylov_params* params, sdx_linear_status* out);

/* Allocation-free variant of sdx_pcg_bsr3_solve.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks, results and status as sdx_pcg_bsr3_solve, but it
 * calls no hipMalloc / hipFree / stream creation: every device buffer lives in
 * a fixed 64 MiB per-device array that the library holds from its first call
 * on a device until the process ends (8 MiB of it stage host arrays; the rest
 * holds the system: about 9 s + 4 bytes per block plus 160 bytes per block row
 * at s = 4 (FP32 and MIXED), so a 40^3-hex cube's system of 993,961 blocks
 * fits; FP64 takes about twice as much).  A larger system is rejected with
 * SDX_ERR_INVALID_ARGUMENT.  It runs on the null stream and returns when x is
 * written.  Calls on one device share the array: calls from threads of one
 * process are serialised by the library, and each call writes every value it
 * reads before reading it. */
sdx_status sdx_pcg_bsr3_solve_noalloc(int32_t device, int32_t n_block_rows, int32_t nnzb,
                                      const int32_t* block_row_ptr, const int32_t* block_col_idx,
                                      const double* values, int32_t layout, const double* b, double* x,
                                      const sdx_krylov_params* params, sdx_linear_status* out);

/* ---------------- L2: world / step (Scene::Step seam) ---------------- */

/* Context/Scene creation (cpp_api/mochi_context.h, mochi_scene.h).
 * precision_mode chooses the storage of the state (FP64 for SDX_PREC_MODE_FP64). */
sdx_status sdx_world_create(int32_t device, int32_t precision_mode, sdx_world** out);
sdx_status sdx_world_destroy(sdx_world* world);

/* CreateTetMeshShape + CreateSoftActor (mochi_context.h:142, mochi_scene.h:99).
 * X: 3*n_verts rest positions, tets: 4*n_tets node indices (positive volume). */
sdx_status sdx_add_soft_tet_mesh(sdx_world* world, const double* X, int32_t n_verts,
                                 const int32_t* tets, int32_t n_tets, const sdx_soft_params* params,
                                 int32_t* out_actor);

/* SetSoftMaterialParamsField: per-element Young's modulus and Poisson ratio. */
sdx_status sdx_set_element_materials(sdx_world* world, int32_t actor, const double* E,
                                     const double* nu, int32_t n_tets);

/* AddBoundaryConditionNodesWorldPermanent (mochi_actor.h:202): pin all three
 * DoFs of the nodes at their current displacement. */
sdx_status sdx_pin_nodes(sdx_world* world, int32_t actor, const int32_t* nodes, int32_t n);

/* AddBoundaryConditionDofsWorld / ClearBoundaryConditions (mochi_actor.h:187-211): hold
 * DoFs (3 node + axis) at world coordinates from the next step on; with clear != 0 the
 * earlier clearable conditions are dropped first (the permanent ones of sdx_pin_nodes
 * stay).  A new set of held DoFs rebuilds the assembly maps at the next step; new values
 * for the same set only move the targets.  DoF indices must be unique and in range, values
 * finite. */
sdx_status sdx_set_boundary_conditions(sdx_world* world, int32_t actor, const int32_t* dofs,
                                       const double* values_world, int32_t n, int32_t clear);

/* Contact query (CQueryActorContactForces, mochi_actor.cpp:2320-2345: registered, its
 * results are available after the next step): with enable != 0 every step ends by
 * evaluating the contact of each soft actor with each collider acting on it at the step's
 * result.  sdx_get_contact_wrench returns for one (actor, collider) pair: out[0..2] the
 * contact force on the actor (penalty and dissipation, GetContactForceWorld), out[3..5] the
 * torque of the reaction on the collider about its position, out[6] the least signed
 * distance of the actor's contact points (samples within the detection threshold,
 * GetContactPointsWorld) or +inf when there are none.  SDX_ERR_STATE without a registered
 * query or before its first step; SDX_ERR_INVALID_ARGUMENT for a collider that does not act
 * on the actor. */
sdx_status sdx_register_contact_query(sdx_world* world, int32_t enable);
sdx_status sdx_get_contact_wrench(sdx_world* world, int32_t actor, int32_t collider, double out[7]);

/* CreateRigidActor with a plane / sphere shape (kinematic).  actor = -1: the
 * collider touches every soft actor, else only that one.  plane: point p,
 * unit normal n pointing to the free side.  sphere: centre c, radius r. */
sdx_status sdx_add_rigid_plane(sdx_world* world, const double p[3], const double n[3],
                               int32_t actor, int32_t* out_collider);
sdx_status sdx_add_rigid_sphere(sdx_world* world, const double c[3], double r, int32_t actor,
                                int32_t* out_collider);
/* SetRootTransform translation: plane point or sphere centre. */
sdx_status sdx_set_rigid_position(sdx_world* world, int32_t collider, const double pos[3]);
/* GetRootTransform translation (mochi_physics/include/mochi_physics/cpp_api/mochi_actor.h:77):
 * the collider's current target, as set or as restored with a state. */
sdx_status sdx_get_rigid_position(sdx_world* world, int32_t collider, double out[3]);
/* The same for n colliders in one call: collider colliders[k] moves to
 * positions[3 k .. 3 k + 2].  No counterpart in the reference (its rigid
 * actors are moved one by one); a batch of environments sets every target of
 * a step at once.  All indices are checked before any collider moves. */
sdx_status sdx_set_rigid_positions(sdx_world* world, int32_t n, const int32_t* colliders, const double* positions);
/* The targets of n colliders in one call: positions[3 k .. 3 k + 2] of collider colliders[k]
 * (after a state restore, every target at once).  No counterpart in the reference. */
sdx_status sdx_get_rigid_positions(sdx_world* world, int32_t n, const int32_t* colliders, double* positions);

/* Scene::Step(dt): one backward-Euler stage per actor, each actor its own
 * island and Newton solve.  stats: n_actor entries (may be NULL). */
sdx_status sdx_step(sdx_world* world, double dt, const sdx_solver_params* params,
                    sdx_step_stats* stats, int32_t n_stats);

/* State access (Actor::GetDofValues / SetNodeVelocities analogues): 3*n_verts doubles. */
sdx_status sdx_get_displacements(sdx_world* world, int32_t actor, double* out, int32_t n);
sdx_status sdx_set_displacements(sdx_world* world, int32_t actor, const double* in, int32_t n);
sdx_status sdx_get_velocities(sdx_world* world, int32_t actor, double* out, int32_t n);
sdx_status sdx_set_velocities(sdx_world* world, int32_t actor, const double* in, int32_t n);
sdx_status sdx_get_positions_world(sdx_world* world, int32_t actor, double* out, int32_t n);

/* State snapshots (Scene::CaptureState / RestoreState / ReleaseState / ReleaseAllStates /
 * IsEqualState / CaptureStateToBytes / RestoreStateFromBytes / CaptureStateToFile,
 * mochi_physics/include/mochi_physics/cpp_api/mochi_scene.h:79-95).  A state is the
 * displacements and velocities of every actor, the collider targets and the Newton-step
 * history of the warm start, so steps after a restore repeat the steps after the
 * capture.  A handle keeps its state in device memory until it is released, or until actors,
 * pins or materials change (then every handle is dropped).
 *
 * The byte form (sdx_state_size bytes) has the container layout of the reference's
 * capture::CaptureState (mochi_physics/src/mochi_capture.cpp:38-75, 456-510): a 24-byte
 * header {"mcap", uint16 context blocks, uint16 entity blocks, 8 zero bytes, uint64 bytes
 * after the header}, component blocks each led by a 24-byte header {uint64 type id, uint64 entity
 * hash, uint32 data size, uint32 entity count}, and the footer "mend".  Type ids are FNV-1a 64
 * of the component names: context "sdx::CWarmStartGate" (12 int32); per actor
 * "mochi::CDisplacementSlice<R>", "mochi::CVelocitySlice<R>" (uint64 n = 3 n_verts, then n
 * reals) and "sdx::CNewtonStepHistory<R>" (uint64 n = 24 n_verts, then n reals); per collider
 * "sdx::CColliderTarget" (target and stage-start target, 6 doubles); R = float or double.
 * The entity hashes cover the actors' meshes, materials and permanent pins and the colliders'
 * kinds, attachments and shapes: a world of another real type or topology rejects the bytes
 * with SDX_ERR_STATE, malformed bytes give SDX_ERR_INVALID_ARGUMENT. */
sdx_status sdx_capture_state(sdx_world* world, int32_t* handle);
sdx_status sdx_restore_state(sdx_world* world, int32_t handle, int32_t release_immediately);
sdx_status sdx_release_state(sdx_world* world, int32_t handle);  /* 0: every handle */
sdx_status sdx_is_equal_state(sdx_world* world, int32_t a, int32_t b, int32_t* equal);
sdx_status sdx_state_size(sdx_world* world, int64_t* bytes);
sdx_status sdx_capture_state_to_bytes(sdx_world* world, uint8_t* out, int64_t capacity);
sdx_status sdx_restore_state_from_bytes(sdx_world* world, const uint8_t* data, int64_t size);
/* Scene::CaptureStateToFile (mochi_scene.h:95): the byte form written to a file. */
sdx_status sdx_capture_state_to_file(sdx_world* world, const char* path);

/* Component kinds of a state, for the exclusions of a partial restore. */
enum {
  SDX_STATE_DISPLACEMENTS = 1,
  SDX_STATE_VELOCITIES = 2,
  SDX_STATE_WARM_START = 4,       /* Newton-step history and warm-start gate counters */
  SDX_STATE_COLLIDER_TARGETS = 8,
};
/* capture::RestorePartialState (mochi_physics/src/mochi_capture.h:47-56): restore every
 * component of the byte form except the kinds in `excluded` (SDX_STATE_* bits); the excluded
 * ones are still checked against the world. */
sdx_status sdx_restore_partial_state_from_bytes(sdx_world* world, const uint8_t* data, int64_t size,
                                                uint32_t excluded);
/* capture::IsEqualState on two byte forms (mochi_capture.h:58-66): equal values of every
 * component (NaN equal to NaN), although the bytes may differ.  Needs no world or device. */
sdx_status sdx_is_equal_state_bytes(const uint8_t* a, int64_t size_a, const uint8_t* b, int64_t size_b,
                                    int32_t* equal);
/* capture::CaptureStateToJson (mochi_capture.h:38-43): a human-readable JSON document of the
 * state, {"actors": {"Soft": {"_handle", "_name", "_type", "components": {...}}, "Soft2": ...,
 * "Rigid": ...}, "scene": {"sdx::CWarmStartGate": ...}}, with reals written to read back
 * exactly (non-finite ones as null).  *length receives the text's length; out (capacity
 * bytes, NUL-terminated) may be NULL with capacity 0 to query it, else capacity must exceed
 * the length.  The first form converts a byte form and needs no world or device. */
sdx_status sdx_state_bytes_to_json(const uint8_t* data, int64_t size, int32_t pretty, char* out, int64_t capacity,
                                   int64_t* length);
sdx_status sdx_capture_state_to_json(sdx_world* world, int32_t pretty, char* out, int64_t capacity,
                                     int64_t* length);

/* World totals: out[0..4] counts over all actors (vertices, tets, BSR blocks, boundary
 * triangles, environments); out[5], out[6] the CG matvec and vector tiles; out[7] the
 * environment groups of the grouped CG iterations under the current schedule (1: none). */
sdx_status sdx_world_info(sdx_world* world, int64_t out[8]);

/* Assembly of one actor at its current displacement, copied to the host:
 * residual (3*n_verts), BSR pattern and values in SDX_BSR_BLOCK_ROWMAJOR with
 * Dirichlet rows/columns applied, as AssembleIslandPipeline leaves them
 * (mochi_solve.cpp:595-771).  Any output pointer may be NULL.  The pattern
 * sizes are given by sdx_world_info / sdx_actor_bsr_size. */
sdx_status sdx_actor_bsr_size(sdx_world* world, int32_t actor, int32_t* n_block_rows,
                              int32_t* nnzb);
sdx_status sdx_assemble(sdx_world* world, int32_t actor, double dt, const sdx_solver_params* params,
                        double* out_residual, int32_t* out_row_ptr, int32_t* out_col_idx,
                        double* out_values);
/* Objective (merit) of one actor at its current displacement with the same stage start as
 * sdx_assemble (SnleProblem::UpdateObjective, snle_problem.cpp; the energies of
 * fem_stress.h, fem_gravity.h:93-111, fem_inertia.h:85-130 and contact_utils.h:560, :935):
 * the function whose gradient sdx_assemble's residual is, as the objective-based line
 * searches evaluate it. */
sdx_status sdx_objective(sdx_world* world, int32_t actor, double dt, const sdx_solver_params* params,
                         double* out_objective);

/* Timing of one kernel class inside sdx_step, measured with HIP events
 * recorded at the kernel's own dispatch on the world's stream.  kernel: 0 =
 * CG matvec (w = A z + z.w), 1 = CG vector update, 2 = element Hessian, 3 =
 * BSR gather, 4 = whole-solve CG (sdx_schedule.env_cg: one launch per linear
 * solve).  enable N > 0 clears the accumulators and times every N-th
 * launch of an assembly class, or every launch of the first CG burst (all
 * environments active; with environment groups, the first burst of each
 * group, whose launches cover that group's rows) of every N-th linear solve
 * (those bursts are launched eagerly, the others as the schedule says), or
 * the whole-solve launch of every N-th linear solve; 0
 * stops timing, -1 leaves the setting unchanged.  out (may be NULL): {timed
 * launches, total milliseconds}. */
sdx_status sdx_kernel_timing(sdx_world* world, int32_t kernel, int32_t enable, double out[2]);

/* CG iterations launched between host convergence checks (default 8). */
sdx_status sdx_set_cg_burst(sdx_world* world, int32_t burst);

/* Launch schedule of a world.  No counterpart in the reference (its CPU
 * path has no launch schedule); the batched autotuner searches this space. */
typedef struct sdx_schedule {
  int32_t cg_burst;       /* CG iterations per graph replay / host check, 1..256 (default 8) */
  int32_t use_graphs;     /* 1: each CG burst is captured once and replayed as a HIP graph (default 1) */
  int32_t spmv_lanes;     /* lanes per block row of the CG matvec: 4, 8 or 16 (default 16) */
  int32_t element_block;  /* threads per workgroup of the element kernel: 64, 128 or 256 (default 128) */
  int32_t active_lists;   /* 1: each CG burst first compacts the tiles of the environments still
                             active into work lists and runs persistent CG kernels over them
                             (default 1) */
  int32_t cg_groups;      /* with active_lists: the environments are split into this many
                             contiguous groups of about equal matrix size, and the first
                             group_iters CG iterations run group after group, so a group's
                             matrix and vectors are re-read from the GPU's last-level cache
                             instead of HBM; later iterations run over all environments
                             (1..256; 1: no groups; default 0: groups of about 200 MB of
                             matrix values and column indices, sdx_world_info out[7]) */
  int32_t group_iters;    /* CG iterations run group by group, rounded up to whole bursts
                             (0..4096, default 16; ignored when cg_groups is 1) */
  int32_t fused_tail;     /* with active_lists: 1 runs each CG burst whose active tiles fit the
                             smallest persistent grid (the tail of a batched solve) as one
                             cooperative launch with grid-wide barriers between the matvec and
                             update phases instead of two launches per iteration (default 0:
                             the barriers' device-scope fences cost more than the launches) */
  int32_t asm_groups;     /* assembly groups: the environments are assembled in this many
                             contiguous groups, each writing its element and contact records and
                             gathering them back before the next group reuses the same record
                             buffers (sized for the largest group), so a group's records can stay
                             in the GPU's last-level cache (0: groups of about 96 MB of records;
                             default 1: one group, the records of every environment kept, which
                             measured fastest: the element and gather kernels are bound by their
                             scattered accesses, not by HBM bandwidth) */
  int32_t fused_asm;      /* 1: every assembly of residual and matrix runs as one fused kernel over
                             row blocks (each workgroup owns a spatially compact block of up to
                             64 vertices, computes every tet and contact triangle touching it and
                             writes its BSR rows, residuals and preconditioner blocks directly, so
                             no per-element records pass through memory; its LDS sums with float
                             atomics vary in the last bits from run to run).  Residual-only
                             assemblies use the records either way, and so does a world with a
                             vertex coupled to more than about 1000 (FP64: 500) vertices, which
                             no block holds.  Default 0: element and contact records gathered
                             into the matrix (asm_groups applies), which measured about 4 times
                             faster: a block recomputes the tets it shares with its neighbours
                             (about 1.7-1.9 times the tets) at an occupancy of two workgroups
                             per CU.  The autotuner leaves it off */
  int32_t env_cg;         /* 1 or 2: each environment's whole CG solve runs in one workgroup of a
                             single launch per solve (that many workgroups per CU), which takes
                             environments from a queue, largest first, and iterates each to its
                             stop decision with z and w in LDS: no kernel boundary, grid barrier,
                             work-list compaction or host check between iterations.  Applies to
                             worlds of several environments whose largest one's rows fit the
                             LDS a workgroup may take (28 bytes per vertex in FP32 and mixed
                             precision, 56 in FP64: about 5800 / 2900 vertices in 160 KB, 2300 /
                             1160 in 64 KB); otherwise, and with 0, the CG runs as the two
                             kernels per iteration the other fields describe (default 1) */
} sdx_schedule;

sdx_status sdx_set_schedule(sdx_world* world, const sdx_schedule* schedule);
sdx_status sdx_get_schedule(sdx_world* world, sdx_schedule* out);

/* Autotuner: times sdx_step of the world's current state under candidate
 * schedules (coordinate search over matvec lanes, the whole-solve CG where
 * every environment fits it, element workgroup size, and where the solve
 * still runs as bursts work lists, CG burst length and graph replay, then
 * environment groups; the best of `trials` steps each),
 * restores the state after every trial and keeps the fastest schedule.
 * out / out_ms (may be NULL): the chosen schedule and its step time. */
sdx_status sdx_autotune(sdx_world* world, double dt, const sdx_solver_params* params, int32_t trials,
                        sdx_schedule* out, double* out_ms);

/* ---------------- learned CG warm start ----------------
 * No counterpart in the reference: its PCG starts from zero
 * (pcg.h:109-130).  A per-vertex network 15 -> 128 -> 128 -> 128 -> 3 (GELU)
 * maps features of the assembled Newton system to a guess u of the Newton
 * step; sdx_step uses it when sdx_solver_params.warm_start is
 * SDX_WARM_START_LEARNED and a model is set.  Per environment the initial
 * guess is the combination x0 = a u + c v + d w of the network's guess and
 * the environment's Newton steps v, w of the same Newton iteration one and two
 * time steps back, with the least residual |b - A x0| (normal equations over
 * every nonempty subset of the three, so a vanishing or parallel guess drops
 * out), kept only if that residual is below 0.2 |b|; else the environment
 * starts from zero, as the plain method does.  A Newton iteration (by index
 * within the step) whose guesses fewer than half of the environments kept in
 * its last warm-started solve (its third or later) starts from zero for the
 * next 8 steps, then tries again.  Inference runs in BF16 on the matrix cores, training in FP32
 * on the device.
 *
 * Feature row (16 floats, the last zero) of vertex i of environment e, with
 * b the residual, D_i the 3x3 diagonal block, d_i = trace(D_i) / 3,
 * y = b / d, dprev the Newton step of the same Newton iteration one time
 * step back (zero before the first), s_e = sqrt(max(|dprev|^2,
 * |y|^2) / n_e) over the environment:
 *   y_i / s_e (3), (D00, D11, D22, D01, D02, D12) / d_i (6), dprev_i / s_e (3),
 *   mean of y_j / s_e over the 1-ring neighbours j (3).
 * Target / output: the Newton step of vertex i divided by s_e.
 * Parameters (FP32, SDX_MLP_PARAMS values): W1[128][15] b1[128] W2[128][128]
 * b2[128] W3[128][128] b3[128] W4[3][128] b4[3], layer y = W x + b. */
#define SDX_MLP_FEATURES 15
#define SDX_MLP_ROW 16
#define SDX_MLP_OUT 3
#define SDX_MLP_PARAMS 35459

typedef struct sdx_mlp sdx_mlp;

typedef struct sdx_train_params {
  int32_t epochs;
  int32_t batch;
  float lr;      /* Adam (Kingma & Ba 2015) */
  float beta1;
  float beta2;
  float eps;
  uint64_t seed; /* row order of every epoch */
} sdx_train_params;

void sdx_default_train_params(sdx_train_params* out);
/* Random initialisation (He normal) from `seed`. */
sdx_status sdx_mlp_create(int32_t device, uint64_t seed, sdx_mlp** out);
sdx_status sdx_mlp_destroy(sdx_mlp* mlp);
sdx_status sdx_mlp_set_params(sdx_mlp* mlp, const float* params, int32_t n);
sdx_status sdx_mlp_get_params(sdx_mlp* mlp, float* params, int32_t n);
/* y[n_rows][3] = net(x[n_rows][16]); bf16 = 1: the matrix-core inference
 * kernel, 0: the FP32 training forward. */
sdx_status sdx_mlp_forward(sdx_mlp* mlp, const float* x, int64_t n_rows, int32_t bf16, float* y);
/* loss = mean over rows of |net(x) - t|^2 and its gradient (grad may be NULL). */
sdx_status sdx_mlp_loss_grad(sdx_mlp* mlp, const float* x, const float* t, int32_t n_rows, float* grad,
                             double* loss);
/* Adam training on host rows; epoch_loss (may be NULL): mean loss of each epoch. */
sdx_status sdx_mlp_train(sdx_mlp* mlp, const float* x, const float* t, int64_t n_rows,
                         const sdx_train_params* params, double* epoch_loss);
/* Adam training on the samples a world collected (same device, no copies). */
sdx_status sdx_mlp_train_collected(sdx_mlp* mlp, sdx_world* world, const sdx_train_params* params,
                                   double* epoch_loss);
/* The world takes a snapshot of the model's parameters (NULL: no model). */
sdx_status sdx_world_set_warm_start_model(sdx_world* world, sdx_mlp* mlp);
/* Collect (feature row, Newton step / s_e) pairs of `per_env` vertices per
 * environment from every linear solve of the following sdx_step calls, up
 * to `capacity` rows (capacity 0: stop and free). */
sdx_status sdx_world_collect(sdx_world* world, int64_t capacity, int32_t per_env);
/* Rows collected so far in *n_rows; copies min(n, cap_rows) rows into x
 * ([n][16]) and t ([n][3]) when they are not NULL. */
sdx_status sdx_world_collected(sdx_world* world, int64_t* n_rows, float* x, float* t, int64_t cap_rows);

#ifdef __cplusplus
}
#endif

#endif /* SDX_FEM_H_ */
