// mpdg_heat_operator.hpp - header-only C++ adapter over the C ABI of
// mpdg_heat.h, shaped like the reference's matrix-free operator interface so
// that host code reads as it does in MeltPoolDG:
//
//   mpdg::HeatOperator        ~ OperatorMatrixFree<3,double> / HeatDiffuseMultiPhaseOperator
//                               (include/meltpooldg/core/operator_base.hpp:189-283)
//   mpdg::JacobiPreconditioner ~ include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:25-89
//   mpdg::LinearSolverData    ~ include/meltpooldg/linear_algebra/linear_solver_data.hpp:37-51
//   mpdg::LinearSolver::solve ~ include/meltpooldg/linear_algebra/linear_solver.hpp:26-134
//
// Vectors are device spans (pointer + size) of n_dofs() doubles in lattice
// numbering. Errors surface as exceptions here, as deal.II's do in the
// reference: mpdg::NoConvergence mirrors SolverControl::NoConvergence.
#pragma once

#include "mpdg_heat.h"

#include <cstddef>
#include <stdexcept>
#include <string>

namespace mpdg
{
  struct DeviceVector
  {
    double *data;
    std::size_t size;
  };

  class Error : public std::runtime_error
  {
  public:
    Error(mpdg_status s, const std::string &what)
      : std::runtime_error(what)
      , status(s)
    {}
    mpdg_status status;
  };

  class NoConvergence : public Error
  {
  public:
    NoConvergence(int last_step, double last_residual)
      : Error(MPDG_ERR_NOT_CONVERGED,
              "Iterative method reported convergence failure in step " + std::to_string(last_step))
      , last_step(last_step)
      , last_residual(last_residual)
    {}
    int last_step;
    double last_residual;
  };

  inline void check(mpdg_status s)
  {
    if (s != MPDG_OK)
      throw Error(s, mpdg_last_error());
  }

  class HeatOperator
  {
  public:
    explicit HeatOperator(const mpdg_heat_desc &desc) { check(mpdg_heat_create(&desc, &h_)); }
    ~HeatOperator() { mpdg_heat_destroy(h_); }
    HeatOperator(const HeatOperator &) = delete;
    HeatOperator &operator=(const HeatOperator &) = delete;

    // OperatorBase::reset_time_increment
    void reset_time_increment(double dt) { check(mpdg_heat_set_dt(h_, dt)); }
    // level_set_as_heaviside registration (gas_liquid material)
    void set_heaviside(const DeviceVector &H) { check(mpdg_heat_set_heaviside(h_, H.data)); }
    // the operator's reference to the current solution (heat_diffuse_operator.hpp:128):
    // the linearisation point of the radiation tangent; the pointer is kept
    void set_temperature(const DeviceVector &T)
    {
      check_size(T);
      check(mpdg_heat_set_temperature(h_, T.data));
    }
    // the operator's reference to the old solution (heat_diffuse_operator.hpp:129),
    // read by the tangent of a temperature-dependent material; the pointer is kept
    void set_temperature_old(const DeviceVector &T_old)
    {
      check_size(T_old);
      check(mpdg_heat_set_temperature_old(h_, T_old.data));
    }
    std::size_t n_dofs() const { return (std::size_t)mpdg_heat_n_dofs(h_); }

    // OperatorMatrixFree::vmult
    void vmult(DeviceVector &dst, const DeviceVector &src) const
    {
      check_size(dst);
      check_size(src);
      check(mpdg_heat_vmult(h_, dst.data, src.data));
    }
    // OperatorMatrixFree::compute_inverse_diagonal_from_matrixfree
    void compute_inverse_diagonal_from_matrixfree(DeviceVector &diag) const
    {
      check_size(diag);
      check(mpdg_heat_inverse_diagonal(h_, diag.data));
    }
    // HeatDiffuseMultiPhaseOperator::create_rhs(dst, src = temperature_old): adds
    // the residual right-hand side to dst; temperature and heat_source are the
    // vectors the reference operator holds from its constructor.
    void create_rhs(DeviceVector &dst, const DeviceVector &src, const DeviceVector &temperature,
                    const DeviceVector *heat_source = nullptr) const
    {
      check_size(dst);
      check_size(src);
      check_size(temperature);
      if (heat_source)
        check_size(*heat_source);
      check(mpdg_heat_create_rhs(h_, dst.data, temperature.data, src.data,
                                 heat_source ? heat_source->data : nullptr));
    }

    mpdg_heat *handle() const { return h_; }

  private:
    void check_size(const DeviceVector &v) const
    {
      if (v.size != n_dofs())
        throw Error(MPDG_ERR_ARGUMENT, "vector size does not match n_dofs()");
    }
    mpdg_heat *h_ = nullptr;
  };

  class JacobiPreconditioner
  {
  public:
    JacobiPreconditioner(const HeatOperator &op, DeviceVector storage)
      : op_(op)
      , inverse_diag_(storage)
    {}
    void update() { op_.compute_inverse_diagonal_from_matrixfree(inverse_diag_); }
    const DeviceVector &get_inverse_diagonal() const { return inverse_diag_; }
    void vmult(DeviceVector &dst, const DeviceVector &src) const
    {
      check(mpdg_heat_jacobi_vmult(op_.handle(), dst.data, src.data, inverse_diag_.data));
    }

  private:
    const HeatOperator &op_;
    DeviceVector inverse_diag_;
  };

  struct LinearSolverData
  {
    mpdg_solver_type solver_type = MPDG_SOLVER_GMRES;
    mpdg_preconditioner_type preconditioner_type = MPDG_PRECONDITIONER_IDENTITY;
    int max_iterations = 10000;
    double rel_tolerance = 1e-12;
    double abs_tolerance = 1e-20;
    bool mixed_precision = false;
  };

  struct LinearSolver
  {
    // Returns the iteration count (SolverControl::last_step()); throws NoConvergence.
    static int solve(const HeatOperator &op, DeviceVector &solution, const DeviceVector &rhs,
                     const LinearSolverData &data, mpdg_solver_report *report = nullptr)
    {
      mpdg_solver_params p;
      mpdg_solver_params_default(&p);
      p.solver_type = data.solver_type;
      p.preconditioner_type = data.preconditioner_type;
      p.max_iterations = data.max_iterations;
      p.rel_tolerance = data.rel_tolerance;
      p.abs_tolerance = data.abs_tolerance;
      p.mixed_precision = data.mixed_precision ? 1 : 0;
      mpdg_solver_report rep;
      const mpdg_status s = mpdg_heat_solve(op.handle(), solution.data, rhs.data, &p, &rep);
      if (report)
        *report = rep;
      if (s == MPDG_ERR_NOT_CONVERGED)
        throw NoConvergence(rep.iterations, rep.final_residual);
      check(s);
      return rep.iterations;
    }

    // LinearSolver::solve(op, x, b, data, preconditioner) (linear_solver.hpp:26-32):
    // the Jacobi preconditioner is applied as the caller last updated it.
    static int solve(const HeatOperator &op, DeviceVector &solution, const DeviceVector &rhs,
                     const LinearSolverData &data, const JacobiPreconditioner &preconditioner,
                     mpdg_solver_report *report = nullptr)
    {
      mpdg_solver_params p;
      mpdg_solver_params_default(&p);
      p.solver_type = data.solver_type;
      p.max_iterations = data.max_iterations;
      p.rel_tolerance = data.rel_tolerance;
      p.abs_tolerance = data.abs_tolerance;
      p.mixed_precision = data.mixed_precision ? 1 : 0;
      mpdg_solver_report rep;
      const mpdg_status s =
        mpdg_heat_solve_preconditioned(op.handle(), solution.data, rhs.data,
                                       preconditioner.get_inverse_diagonal().data, &p, &rep);
      if (report)
        *report = rep;
      if (s == MPDG_ERR_NOT_CONVERGED)
        throw NoConvergence(rep.iterations, rep.final_residual);
      check(s);
      return rep.iterations;
    }
  };
} // namespace mpdg
