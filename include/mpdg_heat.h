/*
 * mpdg_heat.h - C ABI of the MI355X matrix-free heat-operator engine.
 *
 * Drop-in for the hot path of MeltPoolDG's diffuse heat operator:
 *   HeatDiffuseMultiPhaseOperator<3,double>          (include/meltpooldg/heat/heat_diffuse_operator.hpp:109-366)
 *   ::vmult(dst, src)                                 (source/heat/heat_diffuse_operator.cpp:179-197)
 *   ::compute_inverse_diagonal_from_matrixfree(diag)  (source/heat/heat_diffuse_operator.cpp:271-284)
 *   OperatorBase::reset_time_increment(dt)            (include/meltpooldg/core/operator_base.hpp:41-46)
 *   LinearSolver::solve(op, x, b, data, prec)         (include/meltpooldg/linear_algebra/linear_solver.hpp:26-134)
 *   JacobiPreconditioner::vmult/update                (include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:51-69)
 *
 * Conventions
 *   - Plain C types only. Vector arguments are DEVICE pointers (HBM) of
 *     n_dofs() entries in lexicographic lattice numbering
 *     I = ix + N1x*(iy + N1y*iz), N1d = n_cells[d]*degree + 1.
 *   - Homogeneous Dirichlet DoFs (lattice nodes on the marked box faces) follow
 *     the reference's constraint semantics (include/meltpooldg/utilities/
 *     constraints.hpp:49-108): vmult ignores src there and writes dst = 0;
 *     the inverse diagonal is 1 there.
 *   - Every call returns an mpdg_status; 0 is success. No C++ exception crosses
 *     the boundary. mpdg_last_error() returns a thread-local message.
 *   - A handle is single-caller: calls on one handle are serialised on its
 *     stream (the reference's vmult is likewise not re-entrant).
 */
#ifndef MPDG_HEAT_H
#define MPDG_HEAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MPDG_OK = 0,
  MPDG_ERR_ARGUMENT = 1,      /* invalid descriptor, size or pointer */
  MPDG_ERR_HIP = 2,           /* a HIP runtime call or kernel launch failed */
  MPDG_ERR_NOT_CONVERGED = 3, /* SolverControl::NoConvergence (linear_solver.hpp:124-129) */
  MPDG_ERR_UNSUPPORTED = 4,   /* degree / mode combination not built */
  MPDG_ERR_ZERO_TIME_INCREMENT = 5 /* AssertThrowZeroTimeIncrement (core/exceptions.hpp:6) */
} mpdg_status;

typedef enum { MPDG_FP64 = 0, MPDG_FP32 = 1 } mpdg_precision;

/* Material model (include/meltpooldg/core/material.hpp:46-55). */
typedef enum {
  MPDG_MATERIAL_GAS = 0,          /* constant rho, cp, k */
  MPDG_MATERIAL_GAS_LIQUID = 1,   /* blend by a nodal heaviside field, mpdg_heat_set_heaviside */
  MPDG_MATERIAL_LIQUID_SOLID = 2, /* liquid/solid by the temperature's solid fraction (FP64) */
  MPDG_MATERIAL_GAS_LIQUID_SOLID = 3 /* gas / (liquid-solid) by the heaviside, then the
                                        temperature (FP64) */
} mpdg_material_type;

/* ApparentCapacityType (include/meltpooldg/core/material_data.hpp:43-62). */
typedef enum {
  MPDG_APPARENT_CAPACITY_POLY4_BELL = 0, /* L (2/dT) 15/16 (1 - xi^2)^2, xi = 1 - 2 fs */
  MPDG_APPARENT_CAPACITY_CONSTANT = 1,   /* L / dT */
  MPDG_APPARENT_CAPACITY_QLQ = 2         /* quadratic-linear-quadratic, peak 2 L / dT */
} mpdg_apparent_capacity;

/* Extra conductivity field on top of the material's conductivity. */
typedef enum {
  MPDG_CONDUCTIVITY_MATERIAL = 0,
  MPDG_CONDUCTIVITY_SEPARABLE_SINE = 1 /* k += amp * sin(2 pi x) sin(2 pi y) sin(2 pi z) at q */
} mpdg_conductivity_field;

/* Prescribed advection velocity evaluated at the quadrature points. */
typedef enum {
  MPDG_VELOCITY_NONE = 0,
  MPDG_VELOCITY_VORTEX = 1,  /* (-sin^2(pi x) sin(2 pi y), sin(2 pi x) sin^2(pi y), w0) */
  MPDG_VELOCITY_CONSTANT = 2, /* velocity_value, e.g. a nodally interpolated constant
                                prescribed_velocity field (exact for any FE_Q degree) */
  MPDG_VELOCITY_NODAL = 3     /* a DoF-vector velocity (the operator's `velocity`,
                                include/meltpooldg/heat/heat_diffuse_operator.hpp:137-139)
                                given by mpdg_heat_set_velocity: FE_Q(degree) nodal
                                values of each component on the temperature's lattice,
                                read plain and interpolated at q (heat_diffuse_operator.cpp:
                                975-980). FP64 cell loop only (no FP32 handle, no mixed
                                precision, no batches), not with the separable-sine
                                conductivity */
} mpdg_velocity_field;

typedef struct {
  double density;
  double specific_heat_capacity;
  double thermal_conductivity;
} mpdg_phase;

typedef struct {
  int n_cells[3];              /* cells per direction of the box mesh */
  int degree;                  /* FE_Q degree p, 1..6; QGauss(p+1) */
  double box_lo[3], box_hi[3];
  int precision;               /* mpdg_precision of the cell loop */
  int material;                /* mpdg_material_type */
  mpdg_phase gas, liquid;      /* MaterialData gas / liquid */
  int two_phase_sharp;         /* TwoPhaseFluidPropertiesTransitionType::sharp */
  int volume_specific_heat;    /* use_volume_specific_thermal_capacity_for_phase_interpolation */
  /* liquid_solid / gas_liquid_solid (MaterialData, include/meltpooldg/core/
   * material_data.hpp): the solid phase and the solid fraction's law, mushy zone
   * clamp((liquidus - T) / (liquidus - solidus), 0, 1) or, when
   * solid_liquid_sharp (SolidLiquidPropertiesTransitionType::sharp), 1 below
   * the solidus temperature. */
  mpdg_phase solid;
  double solidus_temperature, liquidus_temperature;
  int solid_liquid_sharp;
  /* latent heat of fusion > 0: its apparent capacity (Heat::ApparentCapacity,
   * include/meltpooldg/heat/apparent_capacity.hpp:59) is added to c_p in the
   * mushy zone, of type mpdg_apparent_capacity; not with volume_specific_heat */
  double latent_heat_of_fusion;
  int apparent_capacity_type;
  int conductivity_field;      /* mpdg_conductivity_field */
  double conductivity_amplitude;
  int velocity;                /* mpdg_velocity_field */
  double velocity_w0;          /* vortex: z velocity */
  double velocity_value[3];    /* constant: the velocity vector */
  unsigned dirichlet_faces;    /* bit 2d: face x_d = lo, bit 2d+1: x_d = hi; 0x3F = all */
  /* Robin faces, bits as for dirichlet_faces (the colorized boundary ids of the
   * box): convection alpha (T - T_inf) and radiation emissivity sigma
   * (T^4 - T_inf^4), sigma = 5.670374419e-8 (HeatData::convection / ::radiation,
   * source/heat/heat_data.cpp:55-73; tangent_local_boundary_operation and
   * rhs_boundary_loop, source/heat/heat_diffuse_operator.cpp:719-781,
   * 1088-1129). All zero: no Robin face. */
  unsigned convection_faces;
  double convection_coefficient;
  double convection_temperature_infinity;
  unsigned radiation_faces;
  double emissivity;
  double radiation_temperature_infinity;
  void *stream;                /* hipStream_t the handle launches on; NULL = default stream */
  /* Neumann faces (bits as for dirichlet_faces): a constant heat flux vector q
   * per face (neumann_flux[face] = {qx, qy, qz}); create_rhs adds -(q . n)
   * integrated over the face, n the outward normal (the neumann_bc functions of
   * rhs_boundary_loop, source/heat/heat_diffuse_operator.cpp:735-767, here
   * constant in space). Not on a Robin face (the reference's AssertThrow,
   * :744-747); no effect on the tangent. */
  unsigned neumann_faces;
  double neumann_flux[6][3];
} mpdg_heat_desc;

typedef struct mpdg_heat mpdg_heat;

/* Krylov method (include/meltpooldg/linear_algebra/linear_solver_data.hpp:16-25). */
typedef enum { MPDG_SOLVER_CG = 0, MPDG_SOLVER_GMRES = 1 } mpdg_solver_type;
typedef enum { MPDG_PRECONDITIONER_IDENTITY = 0, MPDG_PRECONDITIONER_DIAGONAL = 3 } mpdg_preconditioner_type;

typedef struct {
  int solver_type;          /* mpdg_solver_type; LinearSolverData::solver_type */
  int preconditioner_type;  /* mpdg_preconditioner_type; Diagonal = JacobiPreconditioner */
  int max_iterations;       /* LinearSolverData::max_iterations (default 10000) */
  double rel_tolerance;     /* LinearSolverData::rel_tolerance */
  double abs_tolerance;     /* LinearSolverData::abs_tolerance (default 1e-20) */
  int gmres_restart;        /* GMRES basis size (30) */
  int mixed_precision;      /* 1: FP32 cell loop inside FP64 defect correction */
  double inner_rel_tolerance; /* mixed: relative reduction of each inner FP32 solve */
  int check_interval;       /* iterations between host convergence checks (0 = auto, at most 64) */
} mpdg_solver_params;

typedef struct {
  int iterations;           /* SolverControl::last_step(); mixed: total inner iterations */
  int outer_iterations;     /* mixed precision: defect-correction steps */
  double initial_residual;  /* ||b - A x0||_2 */
  double final_residual;    /* ||b - A x||_2 at exit (FP64, recomputed for mixed) */
  double seconds;           /* wall time of the solve including the diagonal update */
} mpdg_solver_report;

/* ---- handle ------------------------------------------------------------ */
/* Replaces the operator construction + ScratchData::build (source/core/scratch_data.cpp:190-268). */
mpdg_status mpdg_heat_create(const mpdg_heat_desc *desc, mpdg_heat **out);
mpdg_status mpdg_heat_destroy(mpdg_heat *h);
/* Total lattice DoFs N (deal.II n_dofs()) and free DoFs. */
int64_t mpdg_heat_n_dofs(const mpdg_heat *h);
int64_t mpdg_heat_n_free_dofs(const mpdg_heat *h);

/* Numbering. The engine's vectors use the lexicographic lattice of the box's
 * support points, I = ix + N1x (iy + N1y iz). A caller whose DoFs are numbered
 * otherwise (deal.II's DoFHandler, e.g. from DoFTools::map_dofs_to_support_points)
 * maps them once: mpdg_lattice_index_of_points gives, for each of the n_dofs()
 * support points (x, y, z triples, host memory), its lattice index (host
 * output); the points must be the box's support points, each once, or
 * MPDG_ERR_ARGUMENT is returned. Host computation, no GPU needed. Then
 * mpdg_heat_to_lattice / mpdg_heat_from_lattice move device vectors between
 * the caller's numbering and the lattice with that index (device memory,
 * n_dofs() entries, on the handle's stream). */
mpdg_status mpdg_lattice_index_of_points(const mpdg_heat_desc *desc, const double *points,
                                         int64_t count, uint32_t *lattice_index);
mpdg_status mpdg_heat_to_lattice(mpdg_heat *h, double *lattice, const double *dofs,
                                 const uint32_t *lattice_index);
mpdg_status mpdg_heat_from_lattice(mpdg_heat *h, double *dofs, const double *lattice,
                                   const uint32_t *lattice_index);

/* OperatorBase::reset_time_increment (operator_base.hpp:41-46); dt = 0 is rejected. */
mpdg_status mpdg_heat_set_dt(mpdg_heat *h, double dt);
/* level_set_as_heaviside registration (heat_diffuse_operator.hpp:137-143): device
 * pointer to n_dofs() doubles, read plain (no constraints) and interpolated at q.
 * The engine copies it (and converts to the cell-loop precision). */
mpdg_status mpdg_heat_set_heaviside(mpdg_heat *h, const double *heaviside);
/* The same, for a level set on FE_Q_iso_Q1(p) (p linear sub-elements per cell and
 * direction): replaces the reference's do_level_set_temperature_gradient_interpolation
 * path, the interpolation matrix of reinit (heat_diffuse_operator.cpp:834-838,
 * DoFTools::create_dof_interpolation_matrix, source/utilities/dof_tools.cpp:20-49)
 * applied per cell before the heaviside is evaluated at q (:982-993,
 * interpolate_dof_values, dof_tools.cpp:79-103). heaviside_iso_q1: device pointer
 * to n_dofs() doubles in lattice order, node (i,j,k) at the equidistant sub-grid
 * point (box_lo + (i, j, k) h / p). The engine interpolates it to the GLL support
 * points of FE_Q(p) (the iso field's trilinear interpolant in the sub-cell holding
 * each point) and keeps that copy; degrees 1 and 2 share their support points. */
mpdg_status mpdg_heat_set_heaviside_iso_q1(mpdg_heat *h, const double *heaviside_iso_q1);

/* The nodal velocity of MPDG_VELOCITY_NODAL (the DoF vector the reference operator
 * reads through vel_eval.read_dof_values_plain, heat_diffuse_operator.cpp:975-980):
 * device pointer to 3 n_dofs() doubles, component-blocked [v_x | v_y | v_z], each
 * component in lattice numbering. The engine copies it on the handle's stream:
 * call again when the field changes. Required before vmult, inverse diagonal,
 * create_rhs and solve of such a handle. */
mpdg_status mpdg_heat_set_velocity(mpdg_heat *h, const double *velocity);

/* The temperature the tangent is linearised at: the operator's reference to the
 * current solution (constructor argument temperature_in,
 * include/meltpooldg/heat/heat_diffuse_operator.hpp:128,186), which
 * tangent_local_boundary_operation reads with the Dirichlet entries resolved
 * to 0 (heat_diffuse_operator.cpp:1096-1099). Device pointer of n_dofs()
 * doubles, kept (not copied) and read by every later vmult, inverse diagonal
 * and solve; required when radiation_faces != 0, NULL clears it.
 * mpdg_heat_newton_solve points it at its temperature argument. */
mpdg_status mpdg_heat_set_temperature(mpdg_heat *h, const double *temperature);
/* The old temperature of a temperature-dependent material's tangent: the
 * operator's reference to temperature_old (heat_diffuse_operator.hpp:129,187),
 * which tangent_local_cell_operation reads plain with do_solidification
 * (heat_diffuse_operator.cpp:1003-1014). The linearisation temperature
 * (mpdg_heat_set_temperature) is then also read plain by the cells. Device
 * pointer of n_dofs() doubles, kept (not copied); required by vmult, inverse
 * diagonal and solve for MPDG_MATERIAL_LIQUID_SOLID / _GAS_LIQUID_SOLID, NULL
 * clears it. mpdg_heat_newton_solve points it at its temperature_old argument. */
mpdg_status mpdg_heat_set_temperature_old(mpdg_heat *h, const double *temperature_old);

/* HeatDiffuseMultiPhaseOperator::vmult (heat_diffuse_operator.cpp:179-197), FP64 vectors:
 * the cell loop, then the Robin faces (tangent_boundary_loop :243-269). */
mpdg_status mpdg_heat_vmult(mpdg_heat *h, double *dst, const double *src);
/* Same operator with FP32 vectors and FP32 cell loop (inner operator of the mixed solve). */
mpdg_status mpdg_heat_vmult_f32(mpdg_heat *h, float *dst, const float *src);
/* compute_inverse_diagonal_from_matrixfree (heat_diffuse_operator.cpp:271-284), FP64,
 * with the Robin faces' part (the boundary operation, :352-356). */
mpdg_status mpdg_heat_inverse_diagonal(mpdg_heat *h, double *inv_diag);
/* HeatDiffuseMultiPhaseOperator::create_rhs(dst, src = temperature_old)
 * (heat_diffuse_operator.cpp:783-806, rhs_cell_loop :393-529), FP64:
 *   dst += -( rho_cp/dt M (T - T_old) - M_1 f + K_k T + rho_cp C_u T )
 * on the free rows; constrained rows of dst are left unchanged. T, T_old and
 * the heat source f are read plain (constrained entries as stored), as the
 * reference's read_dof_values_plain does. The reference operator holds T and
 * f by reference from its constructor (heat_diffuse_operator.hpp:186-188);
 * here they are arguments. heat_source may be NULL (f = 0). The Robin faces add
 * -(alpha (T - T_inf) + eps sigma (T^4 - T_inf^4)) integrated over the faces
 * and the Neumann faces -(q . n) (rhs_boundary_loop :719-781), T read plain. A
 * temperature-dependent material is evaluated at T(q) (get_material_parameters
 * :1133-1165). No evaporative cooling term. */
mpdg_status mpdg_heat_create_rhs(mpdg_heat *h, double *dst, const double *temperature,
                                 const double *temperature_old, const double *heat_source);
/* JacobiPreconditioner::vmult (preconditioner_jacobi.hpp:51-56): dst = inv_diag .* src. */
mpdg_status mpdg_heat_jacobi_vmult(mpdg_heat *h, double *dst, const double *src,
                                   const double *inv_diag);

/* LinearSolver::solve (linear_solver.hpp:26-134) with JacobiPreconditioner::update
 * (preconditioner_jacobi.hpp:65-69) done inside. x holds the initial guess on entry.
 * Returns MPDG_ERR_NOT_CONVERGED (report filled) where the reference throws. */
mpdg_status mpdg_heat_solve(mpdg_heat *h, double *x, const double *b,
                            const mpdg_solver_params *params, mpdg_solver_report *report);

/* LinearSolver::solve(op, x, b, data, preconditioner) exactly
 * (linear_solver.hpp:26-134): the caller's preconditioner object is applied as
 * it stands, not updated (its update is the caller's, as
 * HeatDiffuseOperation::solve calls preconditioner.update() before the solve,
 * heat_diffuse_operation.cpp:275-276). inv_diag is the inverse diagonal of a
 * JacobiPreconditioner (preconditioner_jacobi.hpp:51-69; device pointer of
 * n_dofs() doubles, e.g. from mpdg_heat_inverse_diagonal, read only), or NULL
 * for PreconditionIdentity; params->preconditioner_type is not read. A
 * diagonal computed once can serve many solves (a time step's Newton
 * increments). Otherwise as mpdg_heat_solve. */
mpdg_status mpdg_heat_solve_preconditioned(mpdg_heat *h, double *x, const double *b,
                                           const double *inv_diag,
                                           const mpdg_solver_params *params,
                                           mpdg_solver_report *report);

/* SolverControl::enable_history_data / get_history_data (deal.II; the
 * ReductionControl of linear_solver.hpp:36-38): from the next FP64 CG solve
 * on, history[k] = ||r_k||_2 for k = 0 (the initial residual) .. iterations,
 * entries past `length` dropped. history is a caller's device buffer, kept
 * (not copied); length 0 (or NULL) stops recording. GMRES and the inner FP32
 * solves of the mixed-precision solve do not record. */
mpdg_status mpdg_heat_set_residual_history(mpdg_heat *h, double *history, int length);

/* Default parameters: LinearSolverData defaults with CG + Diagonal. */
void mpdg_solver_params_default(mpdg_solver_params *p);

/* Thread-local description of the last error on this thread. */
const char *mpdg_last_error(void);

/* ---- batches of independent solves ---------------------------------------
 * No single counterpart in the reference (one HeatDiffuseOperation per
 * simulation); each member is the operator of mpdg_heat_create and each solve
 * follows mpdg_heat_solve (CG + Jacobi, optional FP32 defect correction). All
 * members are solved together: one batched vmult launch per degree and colour,
 * segmented vector kernels, per-solve convergence flags, and the iteration
 * chunk captured as one hipGraph and replayed. With mixed precision every
 * member runs its own defect correction on the device (its FP64 residual, stop
 * test and FP32 inner restart happen when its own inner solve stops), so a
 * member finishes when it converges, not when the slowest one does. Members
 * must share the material kind, conductivity field and velocity field; meshes
 * and degrees may differ. Members with Robin faces are rejected
 * (MPDG_ERR_UNSUPPORTED). */
typedef struct mpdg_heat_batch mpdg_heat_batch;
typedef struct {
  int status;               /* 0 converged, 1 not converged */
  int iterations;           /* PCG iterations (mixed: total inner iterations) */
  int outer_iterations;     /* mixed precision: defect-correction steps */
  double initial_residual;
  double final_residual;    /* FP64 ||b - A x|| at exit (mixed), PCG residual otherwise */
  double seconds;           /* device wall clock from the start of the batch solve to this member's stop */
} mpdg_batch_report;
mpdg_status mpdg_heat_batch_create(const mpdg_heat_desc *descs, int count, mpdg_heat_batch **out);
mpdg_status mpdg_heat_batch_destroy(mpdg_heat_batch *b);
mpdg_status mpdg_heat_batch_set_dt(mpdg_heat_batch *b, double dt);
int64_t mpdg_heat_batch_n_dofs(const mpdg_heat_batch *b, int member);
/* x[i], rhs[i]: device pointers of member i; reports: count entries (may be NULL). */
mpdg_status mpdg_heat_batch_solve(mpdg_heat_batch *b, double *const *x, const double *const *rhs,
                                  const mpdg_solver_params *params, mpdg_batch_report *reports);

/* ---------------------------------------------------------------------------
 * Heat time step: Newton on the residual (create_rhs) with the tangent solve.
 * ------------------------------------------------------------------------- */

/* NonlinearSolverData (include/meltpooldg/linear_algebra/nonlinear_solver_data.hpp:10-20). */
typedef struct {
  int max_nonlinear_iterations;          /* 10 */
  double field_correction_tolerance;     /* 1e-10 */
  double residual_tolerance;             /* 1e-9 */
  int max_nonlinear_iterations_alt;      /* 0: further increments with the _alt tolerances */
  double field_correction_tolerance_alt; /* 1e-9 */
  double residual_tolerance_alt;         /* 1e-8 */
  int use_relative_tolerances;           /* 0: absolute l2 norms of rhs and increment */
} mpdg_newton_params;
void mpdg_newton_params_default(mpdg_newton_params *p);

typedef struct {
  int iterations;          /* increments solved before convergence was detected */
  int linear_iterations;   /* Krylov iterations summed over the increments */
  double initial_residual; /* ||rhs|| of the first increment */
  double residual;         /* last ||rhs|| (divided by the first when relative) */
  double update;           /* last ||dT|| (divided by ||T|| when relative) */
  double seconds;
} mpdg_newton_report;

/* HeatDiffuseOperation::solve (heat_diffuse_operation.cpp:261-298) with
 * NewtonRaphsonSolver::solve / is_converged / solve_increment
 * (newton_raphson_solver.cpp:22-80, 118-175) and the heat operation's callbacks
 * (heat_diffuse_operation.cpp:82-116). On entry `temperature` holds the start
 * value (the predictor, Dirichlet values set); on return the time step's
 * solution. The operator's linearisation temperature (mpdg_heat_set_temperature)
 * is `temperature` itself, and the Jacobi preconditioner is computed once, at
 * the start value (preconditioner.update() before newton.solve,
 * heat_diffuse_operation.cpp:275-276). Per increment: rhs = create_rhs(T, T_old, f); J dT = rhs solved by
 * mpdg_heat_solve with `linear` from dT = 0; converged when ||rhs|| <
 * residual_tolerance and ||dT|| < field_correction_tolerance (that last dT is
 * not applied), otherwise T += dT (dT is zero on constrained DoFs). Returns
 * MPDG_ERR_NOT_CONVERGED where the reference throws ExcNewtonDidNotConverge.
 * newton / linear / report may be NULL (defaults; the heat defaults of the
 * reference's linear solver are GMRES + Diagonal, heat_data.cpp:11-12). */
mpdg_status mpdg_heat_newton_solve(mpdg_heat *h, double *temperature,
                                   const double *temperature_old, const double *heat_source,
                                   const mpdg_newton_params *newton,
                                   const mpdg_solver_params *linear, mpdg_newton_report *report);

/* compute_linear_predictor (linear_algebra/predictor_linear.cpp:76-96):
 * predictor = (1 + dt/dt_old) old - (dt/dt_old) old_old, or old if |dt_old| < 1e-12. */
mpdg_status mpdg_heat_linear_predictor(mpdg_heat *h, double *predictor, const double *old_vec,
                                       const double *old_old_vec, double dt, double dt_old);

/* AffineConstraints::distribute for the operator's homogeneous Dirichlet
 * constraints (the predictor step of heat_diffuse_operation.cpp:248-249): sets
 * the constrained entries of v to 0. */
mpdg_status mpdg_heat_distribute_constraints(mpdg_heat *h, double *v);

/* VectorTools::compute_norm, L2 (utilities/vector_tools.cpp:170-200): L2 norm of
 * the FE field v with the operator's Gauss quadrature, entries read plain. */
mpdg_status mpdg_heat_l2_norm(mpdg_heat *h, const double *v, double *norm);

/* Allocation-free variant of mpdg_heat_l2_norm. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops). It calls no hipMalloc / hipFree and
 * no allocating library routine; the reduction scratch (one partial per
 * workgroup and a ticket) is a static device array of 4 MiB per GPU, held
 * until the process ends. The result is written to `norm`, a caller's DEVICE
 * double, asynchronously on the handle's stream (no synchronisation); it is
 * bitwise the value mpdg_heat_l2_norm returns. Meshes needing more than 2^19
 * reduction workgroups (row segments x rows x layers) are refused with
 * MPDG_ERR_UNSUPPORTED. All calls on one GPU share the scratch: calls that
 * overlap on one GPU (other streams, other threads) race. The handle's device
 * must be current. */
mpdg_status mpdg_heat_l2_norm_noalloc(mpdg_heat *h, const double *v, double *norm);

/* Krylov iteration chunks as hipGraphs (no counterpart in the reference, whose
 * solvers run on the host): with enable != 0 (the default) a solve runs on a
 * capture stream of the handle's own, ordered after and before the handle's
 * stream by events, and replays captured graphs of its chunks (PCG:
 * check_interval iterations; GMRES: a few Arnoldi steps of a restart cycle,
 * and the cycle's end), each captured the second time it is seen with the
 * same vectors and operator state. Results are bitwise those of the eager
 * launches; 0 launches every kernel eagerly on the handle's stream. */
mpdg_status mpdg_heat_set_graphs(mpdg_heat *h, int enable);

/* Launch tuning of the vmult kernels, one set per precision (no counterpart
 * in the reference, whose cell loop is tuned by deal.II's
 * MatrixFree::AdditionalData). Results do not depend on it beyond the order
 * of floating-point summation. */
typedef enum {
  MPDG_VMULT_LINE = 0,     /* lanes own the 1D lines of a few cells, LDS between directions */
  MPDG_VMULT_CELL = 1,     /* each lane owns a whole cell in registers (degree <= 2) */
  MPDG_VMULT_LINE_MFMA = 2, /* line form with the 1D contractions on FP32 matrix cores
                              (MPDG_FP32 tuning only, degree >= 3) */
  MPDG_VMULT_CELL_MFMA = 3  /* element-matrix form: the cell matrix, the same in every
                              cell, applied to 16 cells at a time on the matrix cores
                              (degree <= 2, gas material with a constant conductivity,
                              no or a constant velocity; tile_log2 1..4 = log2 of the
                              16-cell tile's extent in x) */
} mpdg_vmult_variant;

/* How the line forms order their colour passes. Either way every lattice node
 * receives its cell contributions in the same colour order, so both give
 * bitwise the same dst. */
typedef enum {
  MPDG_SCHEDULE_PASSES = 0,  /* one launch per colour pass (8 launches per vmult) */
  MPDG_SCHEDULE_ORDERED = 1  /* one launch per vmult: work items are taken in colour
                                order and each waits only for the neighbouring items
                                of earlier colours (line forms; the cell form always
                                runs colour passes). All workgroups of a launch are
                                resident at once. Each wait is bounded: if one ever
                                gives up, the handle's failure is sticky and returned
                                as MPDG_ERR_HIP by the synchronising calls
                                (mpdg_heat_solve, mpdg_heat_newton_solve,
                                mpdg_heat_time_vmult) and by every later call on the
                                handle, mpdg_heat_vmult included (the vmult itself does
                                not synchronise: it queues a copy of the error bit that
                                the next call reads) */,
  MPDG_SCHEDULE_SLABS = 2    /* the colour passes split into two z-slabs of work items,
                                each on its own stream (the handle's second stream,
                                forked from and joined to the launch stream by events),
                                so one slab's passes fill the other's launch ramps and
                                tails. The two slabs share one lattice plane, ordered by
                                one event between them: dst is bitwise the PASSES
                                result. Line forms; the cell form runs PASSES */
} mpdg_vmult_schedule;

typedef struct {
  int z_layers;       /* cell layers swept per vmult work item (>= 1) */
  int blocks_per_cu;  /* persistent workgroups per CU, 0 = occupancy limit (<= 16) */
  int variant;        /* mpdg_vmult_variant (see the limits above) */
  int tile_log2;      /* cell variant: log2 of the tile's cells along x (3..6); a
                         wavefront's tile is 2^tile_log2 x 2^(6 - tile_log2) cells */
  int schedule;       /* mpdg_vmult_schedule */
} mpdg_tuning;
mpdg_status mpdg_heat_get_tuning(const mpdg_heat *h, int precision, mpdg_tuning *t);
mpdg_status mpdg_heat_set_tuning(mpdg_heat *h, int precision, const mpdg_tuning *t);

/* Autotuner: times the vmult of this operator in the given precision over the
 * candidate variants, tiles, z-chunk lengths and workgroups per CU (each
 * candidate `repeats` times after one warm-up), installs the fastest and
 * returns it in *best (may be NULL). Uses the handle's solver workspace as
 * scratch vectors; no counterpart in the reference. */
mpdg_status mpdg_heat_autotune(mpdg_heat *h, int precision, int repeats, mpdg_tuning *best);

/* Kernel timing hook for measurement: average device time of the last
 * n_launches vmult launches (ms), recorded with HIP events on the handle's stream. */
mpdg_status mpdg_heat_time_vmult(mpdg_heat *h, double *dst, const double *src, int repeats,
                                 int precision, double *ms_per_vmult);

#ifdef __cplusplus
}
#endif
#endif
