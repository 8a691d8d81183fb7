"""Slab schedule of the line forms: the eight colour passes split into two
z-slabs of work items on two streams, the slab whose boundary chunk is odd
waiting (one event) for the other slab's colour 3 before its colour 4. Every
lattice node receives its cell contributions in the same colour order as with
the single-stream passes, so dst is bitwise the passes' result: for every
degree, precision, z-chunk length, workgroup count, both parities of the
boundary chunks, meshes with one chunk (no split) and with empty colours.
Solves on it (eager and replayed as graphs, whose capture then spans both
streams) keep the oracle's iteration counts; the fused p . Ap is grouped per
slab launch (deterministic, not bitwise the passes' grouping)."""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu


def _t(a, dtype="float64"):
    import torch
    return torch.tensor(a, dtype=getattr(torch, dtype), device="cuda")


def _vmult(op, x, **tuning):
    import torch
    op.set_tuning(**tuning)
    y = torch.full_like(x, np.nan)
    op.vmult(y, x)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _check_same(got, ref, n, p, zl, precision):
    """Bitwise the passes' result (every node summed in the same order)."""
    assert np.array_equal(got, ref)


# nz and the z-chunk lengths below give 1..9 chunks: boundary chunk ks - 1
# even and odd, slab 1 one chunk longer than slab 0, one chunk (no split)
SHAPES = [((3, 2, 5), 2), ((17, 3, 4), 1), ((9, 7, 8), 3), ((6, 5, 7), 4), ((4, 3, 9), 5),
          ((3, 4, 3), 6), ((40, 2, 3), 2), ((2, 2, 1), 3)]


@pytest.mark.parametrize("n,p", SHAPES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_slabs_match_passes(n, p, precision):
    case = make_case(n, p, dt=1e-2, k_amplitude=0.4, velocity=True, faces=0b011101)
    op = engine_operator(case, precision=precision)
    dtype = "float32" if precision == "fp32" else "float64"
    x = _t(np.random.default_rng(7).uniform(-1.0, 1.0, op.n_dofs()), dtype)
    print("slabs", n, p, precision, flush=True)
    for zl in (1, 2, 3):
        for bpc in (1, 0):
            ref = _vmult(op, x, z_layers=zl, blocks_per_cu=bpc, variant="line", schedule="passes",
                         precision=precision)
            first = _vmult(op, x, z_layers=zl, blocks_per_cu=bpc, variant="line",
                           schedule="slabs", precision=precision)
            _check_same(first, ref, n, p, zl, precision)
            for rep in range(2):
                got = _vmult(op, x, z_layers=zl, blocks_per_cu=bpc, variant="line",
                             schedule="slabs", precision=precision)
                assert np.array_equal(got, first), (zl, bpc, rep)
    if precision == "fp64":
        pb = oracle_problem(case)
        assert rel_err(got, pb.vmult(x.cpu().numpy())) <= 1e-12


@pytest.mark.parametrize("n,p", SHAPES[:6])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_slabs_match_passes_lean(n, p, precision):
    """Operators without a velocity or a heaviside take the two-slot and
    two-wavefront launch shapes (FP32 degrees 1-2, FP64 degree 3) and the
    80-VGPR FP32 degree-4 budget: the same bitwise agreement."""
    case = make_case(n, p, dt=1e-2, k_amplitude=0.4, faces=0b110110)
    op = engine_operator(case, precision=precision)
    dtype = "float32" if precision == "fp32" else "float64"
    x = _t(np.random.default_rng(17).uniform(-1.0, 1.0, op.n_dofs()), dtype)
    print("slabs lean", n, p, precision, flush=True)
    for zl in (1, 2, 3):
        ref = _vmult(op, x, z_layers=zl, blocks_per_cu=0, variant="line", schedule="passes",
                     precision=precision)
        got = _vmult(op, x, z_layers=zl, blocks_per_cu=0, variant="line", schedule="slabs",
                     precision=precision)
        _check_same(got, ref, n, p, zl, precision)
    if precision == "fp64":
        pb = oracle_problem(case)
        assert rel_err(got, pb.vmult(x.cpu().numpy())) <= 1e-12
    else:
        pb = oracle_problem(case)
        assert rel_err(got, pb.vmult(x.cpu().numpy().astype(np.float64))) < 2e-5


@pytest.mark.parametrize("p", [3, 4, 5])
def test_slabs_mfma_form(p):
    case = make_case((7, 5, 6), p, dt=1e-3, k_amplitude=0.5)
    op = engine_operator(case, precision="fp32")
    x = _t(np.random.default_rng(8).uniform(-1.0, 1.0, op.n_dofs()), "float32")
    for zl in (1, 2):
        ref = _vmult(op, x, z_layers=zl, variant="mfma", schedule="passes", precision="fp32")
        got = _vmult(op, x, z_layers=zl, variant="mfma", schedule="slabs", precision="fp32")
        _check_same(got, ref, (7, 5, 6), p, zl, "fp32")


def test_slabs_heaviside_and_robin():
    rng = np.random.default_rng(9)
    n, p = (6, 5, 8), 3
    N = (n[0] * p + 1) * (n[1] * p + 1) * (n[2] * p + 1)
    mat = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 2.0, "gas_k": 0.5,
           "liq_rho": 3.0, "liq_cp": 1.0, "liq_k": 4.0}
    robin = dict(convection_faces=0b100000, convection_coefficient=25.0, convection_t_inf=300.0,
                 radiation_faces=0b000001, emissivity=0.8, radiation_t_inf=300.0)
    case = make_case(n, p, dt=1e-2, material=mat, velocity=True, faces=0b000110,
                     heaviside=rng.uniform(0.0, 1.0, N), robin=robin)
    op = engine_operator(case)
    T = rng.uniform(500.0, 1500.0, N)
    op.set_temperature(_t(T))
    pb = oracle_problem(case)
    pb.set_temperature(T)
    x = _t(rng.uniform(-1.0, 1.0, N))
    for zl in (1, 2):
        ref = _vmult(op, x, z_layers=zl, schedule="passes", precision="fp64")
        got = _vmult(op, x, z_layers=zl, schedule="slabs", precision="fp64")
        _check_same(got, ref, n, p, zl, "fp64")
        assert rel_err(got, pb.vmult(x.cpu().numpy())) <= 1e-12


@pytest.mark.parametrize("solver,mixed", [("CG", False), ("CG", True), ("GMRES", False),
                                          ("GMRES", True)])
def test_slab_solves(solver, mixed):
    import torch
    from meltpooldg_amd import LinearSolver, LinearSolverData
    vel = "constant" if solver == "GMRES" else False
    case = make_case((10, 9, 8), 3, dt=1e-3, k_amplitude=0.5 if solver == "CG" else 0.0,
                     velocity=vel, velocity_value=(1.0, 0.5, -0.5))
    op = engine_operator(case)
    pb = oracle_problem(case)
    b = _t(pb.random(1))
    out = {}
    for sched in ("passes", "slabs"):
        for prec in ("fp64", "fp32"):
            op.set_tuning(schedule=sched, z_layers=2, precision=prec)
        for graphs in (False, True, True):
            op.set_graphs(graphs)
            x = torch.zeros_like(b)
            data = LinearSolverData(solver_type=solver, preconditioner_type="Diagonal",
                                    rel_tolerance=1e-9, mixed_precision=mixed, gmres_restart=8)
            its = LinearSolver.solve(op, x, b, data)
            torch.cuda.synchronize()
            out.setdefault(sched, []).append((its, x.cpu().numpy()))
    # graph replays (captured across both streams) repeat the eager solve bitwise
    for its, x in out["slabs"][1:]:
        assert its == out["slabs"][0][0] and np.array_equal(x, out["slabs"][0][1])
    its_p, x_p = out["passes"][0]
    its_s, x_s = out["slabs"][0]
    assert abs(its_s - its_p) <= 1, (its_s, its_p)
    assert np.abs(x_s - x_p).max() <= 1e-7 * np.abs(x_p).max()
    if not mixed:
        ref = pb.solve(b.cpu().numpy(), solver=solver, rel_tol=1e-9, inv_diag=pb.inverse_diagonal(),
                       restart=8)
        assert ref["iterations"] == its_s
