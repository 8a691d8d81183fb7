"""Paired CG iterations against single updates.

With an even convergence-check interval the FP64 device PCG runs its iterations in
pairs: the direction alternates between two vectors, the first iteration of a
pair leaves x alone and the second applies both steps,
x = (x + alpha_1 p_1) + alpha_2 p_2. With an odd interval every iteration
updates x. Both perform the same roundings in the same order, so iteration
counts, residuals and x must agree bitwise. The tolerance sweep ends solves on
both parities of the iteration count: an odd count stops on the first
iteration of a pair and takes the pending x update after the last chunk.
The iteration count of the paired path is also checked against the oracle's
(deal.II SolverCG restated, linear_solver.hpp:106-111).
"""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}


def _solve(op, b, x0, tol, check, mixed=False):
    import torch
    from meltpooldg_amd import LinearSolver, LinearSolverData
    x = torch.tensor(x0, dtype=torch.float64, device="cuda")
    data = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal", rel_tolerance=tol,
                            check_interval=check, mixed_precision=mixed)
    it = LinearSolver.solve(op, x, torch.tensor(b, dtype=torch.float64, device="cuda"), data)
    torch.cuda.synchronize()
    return it, dict(LinearSolver.last_report), x


def test_fp64_pairs_bitwise_equal_single_updates_both_parities():
    import torch
    case = make_case((10, 9, 8), 3, k_amplitude=0.5)
    pb = oracle_problem(case)
    b = pb.random(1)
    op = engine_operator(case)
    zero = np.zeros(op.n_dofs())
    parities = set()
    for tol in (1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10):
        it_p, rep_p, x_p = _solve(op, b, zero, tol, check=8)
        it_s, rep_s, x_s = _solve(op, b, zero, tol, check=7)
        print(f"tol {tol:g}: paired {it_p} single {it_s}", flush=True)
        assert it_p == it_s
        assert rep_p["final_residual"] == rep_s["final_residual"]
        assert torch.equal(x_p, x_s)
        parities.add(it_p % 2)
    assert parities == {0, 1}
    ref = pb.solve(b, rel_tol=1e-9, inv_diag=pb.inverse_diagonal())
    it, _, x = _solve(op, b, zero, 1e-9, check=8)
    assert it == ref["iterations"]
    assert rel_err(x.cpu().numpy(), ref["x"]) < 1e-8


def test_fp64_pairs_nonzero_initial_guess_and_two_phase():
    import torch
    from oracle.rng import random_vector
    n, p = 9, 2
    H = 0.5 * (random_vector(n, p, 3, dirichlet_faces=0) + 1.0)
    case = make_case(n, p, material=TWO, heaviside=H)
    pb = oracle_problem(case)
    b = pb.random(1)
    x0 = 0.1 * pb.random(2)
    op = engine_operator(case)
    for check in (2, 16):
        it_p, rep_p, x_p = _solve(op, b, x0, 1e-10, check=check)
        it_s, rep_s, x_s = _solve(op, b, x0, 1e-10, check=5)
        print(f"check {check}: paired {it_p} single {it_s}", flush=True)
        assert it_p == it_s
        assert rep_p["final_residual"] == rep_s["final_residual"]
        assert torch.equal(x_p, x_s)
    ref = pb.solve(b, rel_tol=1e-10, inv_diag=pb.inverse_diagonal(), x0=x0)
    assert it_p == ref["iterations"]


def test_mixed_solve_independent_of_check_interval():
    """The mixed solve's FP32 inner CG keeps single updates (paired ones
    measured slower there); its FP64 outer correction must not depend on the
    check interval either."""
    import torch
    case = make_case(12, 4, k_amplitude=0.5)
    pb = oracle_problem(case)
    b = pb.random(1)
    op = engine_operator(case)
    zero = np.zeros(op.n_dofs())
    for tol in (1e-6, 1e-8):
        it_p, rep_p, x_p = _solve(op, b, zero, tol, check=8, mixed=True)
        it_s, rep_s, x_s = _solve(op, b, zero, tol, check=7, mixed=True)
        print(f"mixed tol {tol:g}: check 8 {it_p}/{rep_p['outer_iterations']} "
              f"check 7 {it_s}/{rep_s['outer_iterations']}", flush=True)
        assert it_p == it_s
        assert rep_p["outer_iterations"] == rep_s["outer_iterations"]
        assert rep_p["final_residual"] == rep_s["final_residual"]
        assert torch.equal(x_p, x_s)
        assert rep_p["final_residual"] <= tol * rep_p["initial_residual"]
