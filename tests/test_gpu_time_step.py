"""GPU heat time step (linear predictor, Newton with create_rhs and the
tangent solve, FE L2 norm) through the C ABI.

* The reference's own regression output of unidirectional_heat_transfer (the
  norms printed after each Newton solve, 6 significant digits) on the
  z-extruded mesh that carries the same discrete solution (see
  tests/test_time_step_oracle.py).
* The same steps on a problem with Dirichlet faces, variable conductivity, a
  heat source and Q3 elements against the oracle's restatement of the time
  step (within 1e-10 relative).
* The FE L2 norm and the predictor against the oracle.
"""
import json
import os

import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _t(a):
    import torch
    return torch.tensor(a, dtype=torch.float64, device="cuda")


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_unidirectional_heat_transfer_matches_reference_output():
    from meltpooldg_amd import HeatDiffuseOperation, HeatDiffuseOperator, NonlinearSolverData
    g = json.load(open(os.path.join(HERE, "golden", "unidirectional_heat_transfer.json")))
    s = g["setup"]
    n = 2 ** s["global_refinements"]
    op = HeatDiffuseOperator((n, 1, 1), s["degree"], box=((0.0, 0.0, 0.0), (0.1, 0.1, 1.0)),
                             dirichlet_faces=0)
    x = np.linspace(0.0, 0.1, n * s["degree"] + 1)
    T0 = _t(np.tile(x, 4))
    heat = HeatDiffuseOperation(op, T0, nlsolve=NonlinearSolverData(
        residual_tolerance=s["nlsolve"]["residual_tolerance"]))
    assert abs(heat.l2_norm() - np.sqrt(0.1 * 0.1 ** 3 / 3.0)) < 1e-15
    for step, ref in enumerate(g["l2_norms"]):
        rep = heat.advance(s["time_step"])
        norm = heat.l2_norm()
        print("step", step + 1, norm, ref, rep, flush=True)
        assert float(f"{norm:.5e}") == ref


@pytest.mark.parametrize("p", [1, 3])
def test_time_steps_match_oracle(p):
    import torch
    from meltpooldg_amd import HeatDiffuseOperation, LinearSolverData
    from oracle.newton import NewtonData, linear_predictor, newton_time_step
    from oracle.rng import constrained_mask
    case = make_case((7, 6, 5), p, k_amplitude=0.5, dt=2e-3, faces=0b011011)
    pb = oracle_problem(case)
    op = engine_operator(case)
    rng = np.random.default_rng(2)
    mask = constrained_mask(case["n"], p, case["faces"])
    T0 = rng.uniform(0.0, 1.0, pb.N)
    T0[mask] = 0.0
    f = rng.uniform(-1.0, 1.0, pb.N)
    lin = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal", rel_tolerance=1e-13)
    heat = HeatDiffuseOperation(op, _t(T0), heat_source=_t(f), linear_solver=lin)
    To, Too, dto = T0, T0, 0.0
    for step, dt in enumerate((2e-3, 2e-3, 1e-3)):
        rep = heat.advance(dt)
        pb.set_dt(dt)
        T, orep = newton_time_step(pb, linear_predictor(To, Too, dt, dto), To, f,
                                   data=NewtonData(), solver="CG", rel_tol=1e-13)
        print("step", step, rep, orep, flush=True)
        assert orep["converged"]
        assert rep["iterations"] == orep["iterations"]
        assert rel_err(_np(heat.temperature), T) < 1e-10
        assert abs(heat.l2_norm() - pb.l2_norm(T)) < 1e-10 * pb.l2_norm(T)
        Too, To, dto = To, T, dt


def test_l2_norm_and_predictor_vs_oracle():
    import ctypes
    import torch
    from meltpooldg_amd import HeatDiffuseOperation
    from meltpooldg_amd import _lib as L
    case = make_case((9, 4, 5), 4, faces=0)
    pb = oracle_problem(case)
    op = engine_operator(case)
    u = np.random.default_rng(4).uniform(-1.0, 1.0, pb.N)
    heat = HeatDiffuseOperation(op, _t(u))
    assert abs(heat.l2_norm() - pb.l2_norm(u)) < 1e-13 * pb.l2_norm(u)
    a, b = _t(u), _t(2.0 * u + 1.0)
    pred = torch.empty_like(a)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    L.check(op._lib.mpdg_heat_linear_predictor(op._h, vp(pred), vp(a), vp(b), 2.0, 1.0))
    assert rel_err(_np(pred), 3.0 * u - 2.0 * (2.0 * u + 1.0)) < 1e-15
    L.check(op._lib.mpdg_heat_linear_predictor(op._h, vp(pred), vp(b), vp(a), 1.0, 0.0))
    assert np.array_equal(_np(pred), _np(b))
