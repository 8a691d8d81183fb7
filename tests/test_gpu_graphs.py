"""Krylov iteration chunks replayed as hipGraphs: the PCG chunks and the GMRES
restart cycles (device-side Givens rotations and back substitution) of a
handle give bitwise the same iterates, iteration counts and reports whether
their launches are replayed from captured graphs or issued eagerly.

Each test solves several times in a row with the same vectors, so the cached
graph is replayed many times, with the program's own eager work (vmults,
fills, a solve with graphs off) between the replays, and every replay is
checked against the eager result. Changing the time increment, the
temperature or the vectors recaptures. GMRES runs with a short restart so
that most of its cycles are replays; its iteration counts are also checked
against the oracle (the device Givens rotation restates the oracle's).
"""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.tensor(a, dtype=torch.float64, device="cuda")


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _solve(op, b, solver, graphs, restart=30, rel=1e-10, mixed=False):
    import torch
    from meltpooldg_amd import LinearSolver, LinearSolverData
    op.set_graphs(graphs)
    x = torch.zeros_like(b)
    data = LinearSolverData(solver_type=solver, preconditioner_type="Diagonal", rel_tolerance=rel,
                            gmres_restart=restart, mixed_precision=mixed, check_interval=4)
    its = LinearSolver.solve(op, x, b, data)
    return _np(x), its, dict(LinearSolver.last_report)


@pytest.mark.parametrize("solver,restart", [("CG", 30), ("GMRES", 5), ("GMRES", 30)])
def test_graph_replays_match_eager_launches(solver, restart):
    import torch
    vel = "constant" if solver == "GMRES" else False
    case = make_case((9, 7, 8), 2, dt=1e-2, k_amplitude=0.0, velocity=vel,
                     velocity_value=(3.0, -2.0, 1.0), faces=0b110011)
    op = engine_operator(case)
    pb = oracle_problem(case)
    rng = np.random.default_rng(5)
    rhs = [pb.random(1 + i) for i in range(4)]
    eager = [_solve(op, _t(b), solver, False, restart) for b in rhs]
    scratch = torch.empty(op.n_dofs(), dtype=torch.float64, device="cuda")
    for rep in range(2):
        for i, b in enumerate(rhs):
            x, its, r = _solve(op, _t(b), solver, True, restart)
            print(solver, restart, rep, i, its, flush=True)
            assert its == eager[i][1]
            assert np.array_equal(x, eager[i][0]), (solver, restart, rep, i)
            assert r["final_residual"] == eager[i][2]["final_residual"]
            # eager work between the replays
            op.vmult(scratch, _t(rng.uniform(-1.0, 1.0, op.n_dofs())))
    # iteration counts of the oracle (same operator, same stopping rule)
    inv = pb.inverse_diagonal()
    for i, b in enumerate(rhs[:2]):
        ref = pb.solve(b, solver=solver, rel_tol=1e-10, inv_diag=inv, restart=restart)
        assert ref["status"] == 0 and ref["iterations"] == eager[i][1], (i, ref["iterations"])
        assert np.abs(eager[i][0] - ref["x"]).max() <= 1e-8 * np.abs(ref["x"]).max()
    # a new time increment changes the captured kernel arguments: recapture
    op.reset_time_increment(2e-2)
    e2 = _solve(op, _t(rhs[0]), solver, False, restart)
    g2 = _solve(op, _t(rhs[0]), solver, True, restart)
    assert g2[1] == e2[1] and np.array_equal(g2[0], e2[0])
    assert e2[1] != eager[0][1] or not np.array_equal(e2[0], eager[0][0])


@pytest.mark.parametrize("solver", ["CG", "GMRES"])
def test_graphs_in_the_mixed_precision_solve(solver):
    vel = "constant" if solver == "GMRES" else False
    case = make_case((10, 9, 8), 3, dt=1e-3, k_amplitude=0.5 if solver == "CG" else 0.0,
                     velocity=vel, velocity_value=(1.0, 0.5, -0.5))
    op = engine_operator(case)
    pb = oracle_problem(case)
    b = _t(pb.random(1))
    e = _solve(op, b, solver, False, restart=8, rel=1e-9, mixed=True)
    for _ in range(3):
        g = _solve(op, b, solver, True, restart=8, rel=1e-9, mixed=True)
        assert g[1] == e[1] and g[2]["outer_iterations"] == e[2]["outer_iterations"]
        assert np.array_equal(g[0], e[0])


def test_graphs_in_newton_time_steps_with_robin_and_solidification():
    """The time step's GMRES solves replay graphs that read the temperature and
    the old temperature through the pointers the Newton loop registers; the
    Robin faces' launches are part of the captured vmult."""
    import torch
    from meltpooldg_amd import HeatDiffuseOperation, LinearSolverData, NonlinearSolverData
    steel = {"gas_rho": 3925.0, "gas_cp": 230.0, "gas_k": 27.7815, "liq_rho": 7850.0,
             "liq_cp": 460.0, "liq_k": 55.563, "sol_rho": 7850.0, "sol_cp": 700.0,
             "sol_k": 17.0, "solidus": 1966.0, "liquidus": 1974.0, "type": "liquid_solid"}
    robin = dict(convection_faces=0b000010, convection_coefficient=25.0, convection_t_inf=300.0,
                 radiation_faces=0b000010, emissivity=0.8, radiation_t_inf=300.0)
    case = make_case((12, 3, 3), 2, dt=0.05, material=steel, faces=0, robin=robin,
                     velocity="constant", velocity_value=(-0.005, 0.0, 0.0),
                     lo=(0.0, 0.0, 0.0), hi=(0.1, 0.025, 0.025))
    n1 = [c * 2 + 1 for c in case["n"]]
    x = np.linspace(0.0, 1.0, n1[0])
    T0 = np.broadcast_to((1960.0 + 20.0 * x)[None, None, :], n1[::-1]).reshape(-1).copy()
    out = {}
    for graphs in (False, True):
        op = engine_operator(case)
        op.set_graphs(graphs)
        heat = HeatDiffuseOperation(op, _t(T0), linear_solver=LinearSolverData(
            solver_type="GMRES", preconditioner_type="Diagonal", rel_tolerance=1e-12,
            gmres_restart=10), nlsolve=NonlinearSolverData(residual_tolerance=1e-6),
            predictor="linear_extrapolation")
        reps = [heat.advance(0.05) for _ in range(4)]
        out[graphs] = (_np(heat.temperature), [(r["iterations"], r["linear_iterations"])
                                               for r in reps])
    assert out[True][1] == out[False][1]
    assert np.array_equal(out[True][0], out[False][0])
    assert sum(li for _, li in out[True][1]) > 40  # several restart cycles replayed


def test_check_interval_limit():
    """Chunks of at most 64 iterations are captured; a longer interval is
    refused before anything is launched (MPDG_ERR_ARGUMENT), and 64 solves
    to the eager result."""
    import torch
    from meltpooldg_amd import LinearSolver, LinearSolverData
    from meltpooldg_amd import _lib as L
    case = make_case((6, 5, 4), 2, dt=1e-2, k_amplitude=0.0)
    op = engine_operator(case)
    pb = oracle_problem(case)
    b = _t(pb.random(3))
    for bad in (65, 128, -1):
        x = torch.zeros_like(b)
        data = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal",
                                rel_tolerance=1e-10, check_interval=bad)
        with pytest.raises(L.MpdgError) as e:
            LinearSolver.solve(op, x, b, data)
        assert "check_interval" in str(e.value)
        assert float(torch.abs(x).max()) == 0.0
    x_e, its_e, _ = _solve(op, b, "CG", False)
    op.set_graphs(True)
    for rep in range(3):
        x = torch.zeros_like(b)
        its = LinearSolver.solve(op, x, b, LinearSolverData(
            solver_type="CG", preconditioner_type="Diagonal", rel_tolerance=1e-10, check_interval=64))
        assert its == its_e
        assert np.array_equal(_np(x), x_e)
