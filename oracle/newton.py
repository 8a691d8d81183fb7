"""Heat time step of the reference restated on the CPU oracle (test
infrastructure only).

* compute_linear_predictor (source/linear_algebra/predictor_linear.cpp:76-96):
  predictor = (1 + dt/dt_old) old - (dt/dt_old) old_old, or old when
  |dt_old| < 1e-12.
* NewtonRaphsonSolver::solve / is_converged / solve_increment
  (source/linear_algebra/newton_raphson_solver.cpp:22-80, 118-175) with the
  heat operation's callbacks (source/heat/heat_diffuse_operation.cpp:82-116):
  rhs = create_rhs(T, T_old) (no user right-hand side), the Jacobian solve is
  LinearSolver::solve with the heat operator and the Jacobi preconditioner
  from a zero initial update, the update is applied only when not converged,
  constraints of the update are homogeneous.
* The operator reads the current solution as its linearisation point (the
  radiation tangent, heat_diffuse_operator.cpp:1096-1099; the solidification
  terms, :1003-1014, which also read the old temperature), and the Jacobi
  preconditioner is updated once per time step, at the start value
  (HeatDiffuseOperation::solve, source/heat/heat_diffuse_operation.cpp:251-253).
* NonlinearSolverData defaults (include/meltpooldg/linear_algebra/
  nonlinear_solver_data.hpp:13-19).
"""
import numpy as np


class NewtonData:
    def __init__(self, max_nonlinear_iterations=10, field_correction_tolerance=1e-10,
                 residual_tolerance=1e-9, max_nonlinear_iterations_alt=0,
                 field_correction_tolerance_alt=1e-9, residual_tolerance_alt=1e-8,
                 use_relative_tolerances=False):
        self.max_nonlinear_iterations = max_nonlinear_iterations
        self.field_correction_tolerance = field_correction_tolerance
        self.residual_tolerance = residual_tolerance
        self.max_nonlinear_iterations_alt = max_nonlinear_iterations_alt
        self.field_correction_tolerance_alt = field_correction_tolerance_alt
        self.residual_tolerance_alt = residual_tolerance_alt
        self.use_relative_tolerances = use_relative_tolerances


def linear_predictor(old, old_old, dt, dt_old):
    if abs(dt_old) < 1e-12:
        return np.array(old, dtype=float)
    frac = dt / dt_old
    return (frac + 1.0) * old - frac * old_old


def newton_time_step(pb, T, T_old, heat_source=None, data=None, solver="GMRES", rel_tol=1e-12,
                     abs_tol=1e-20, max_it=10000, restart=30):
    """Returns (T_new, report): T is the start value (the predictor)."""
    data = data or NewtonData()
    T = np.array(T, dtype=float)
    pb.set_temperature(T)
    pb.set_temperature_old(T_old)
    inv_diag = pb.inverse_diagonal()
    res_tol, corr_tol = data.residual_tolerance, data.field_correction_tolerance
    max_total = data.max_nonlinear_iterations + data.max_nonlinear_iterations_alt
    r0 = None
    lin = 0
    for i in range(max(max_total, 1)):
        if i == data.max_nonlinear_iterations:
            res_tol, corr_tol = data.residual_tolerance_alt, data.field_correction_tolerance_alt
        pb.set_temperature(T)
        rhs = pb.create_rhs(T, T_old, heat_source)
        sol = pb.solve(rhs, solver=solver, rel_tol=rel_tol, abs_tol=abs_tol, max_it=max_it,
                       inv_diag=inv_diag, restart=restart)
        du = sol["x"]
        lin += sol["iterations"]
        res = np.linalg.norm(rhs)
        if i == 0:
            r0 = res
        upd = np.linalg.norm(du)
        if data.use_relative_tolerances:
            if r0 > 0.0:
                res /= r0
            tn = np.linalg.norm(T)
            if tn > 0.0:
                upd /= tn
        if res < res_tol and upd < corr_tol:
            return T, {"iterations": i, "linear_iterations": lin, "residual": res, "update": upd,
                       "converged": True}
        T = T + du
    return T, {"iterations": max_total, "linear_iterations": lin, "converged": False}
