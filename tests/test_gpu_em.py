"""Element-matrix form of the vmult on the matrix cores (MPDG_VMULT_CELL_MFMA,
em_kernels.hip): operators whose cell matrix is the same in every cell (gas
material, no conductivity field, no or a constant velocity), degrees 1 and 2.
Checked against the oracle's sum-factorised operator on the same inputs: FP64
to 1e-12, FP32 to 2e-5 (relative max norm), over tile shapes, z-chunk lengths,
partial tiles, Dirichlet masks, anisotropic boxes and a constant velocity;
repeats are bitwise equal; the fused p . Ap gives the oracle's CG iteration
count; ineligible operators refuse the form."""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

MAT = {"type": "gas", "gas_rho": 2.0, "gas_cp": 1.5, "gas_k": 0.7}


def _em(op, precision, tile, z):
    op.set_tuning(z_layers=z, blocks_per_cu=0, variant="cell_mfma", tile_log2=tile,
                  schedule="passes", precision=precision)
    assert op.get_tuning(precision)["variant"] == "cell_mfma"


def _vmult(op, u, dtype):
    import torch
    src = torch.tensor(u, dtype=dtype, device="cuda")
    dst = torch.full_like(src, float("nan"))
    op.vmult(dst, src)
    torch.cuda.synchronize()
    return dst.double().cpu().numpy()


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("tile", [1, 2, 3, 4])
def test_em_vmult_vs_oracle_tiles(p, tile):
    import torch
    for n, z, faces in (((7, 5, 6), 2, 0x3F), ((9, 4, 3), 1, 0b010101), ((5, 6, 7), 3, 0)):
        case = make_case(n, p, material=MAT, faces=faces, dt=2e-3)
        print("em", p, tile, n, z, faces, flush=True)
        pb = oracle_problem(case)
        u = pb.random(0)
        ref = pb.vmult(u)
        op = engine_operator(case)
        _em(op, "fp64", tile, z)
        assert rel_err(_vmult(op, u, torch.float64), ref) < 1e-12
        op32 = engine_operator(case, precision="fp32")
        _em(op32, "fp32", tile, z)
        assert rel_err(_vmult(op32, u.astype(np.float32), torch.float32), ref) < 2e-5


@pytest.mark.parametrize("p", [1, 2])
def test_em_constant_velocity_and_box(p):
    import torch
    case = make_case((6, 5, 4), p, material=MAT, velocity="constant",
                     velocity_value=(0.8, -0.5, 0.3), lo=(0.0, -0.2, 0.1), hi=(0.6, 0.4, 0.3),
                     faces=0b110011)
    pb = oracle_problem(case)
    u = pb.random(3)
    ref = pb.vmult(u)
    for precision, dtype, tol in (("fp64", torch.float64, 1e-12), ("fp32", torch.float32, 2e-5)):
        op = engine_operator(case, precision=precision)
        _em(op, precision, 2, 2)
        uu = u.astype(np.float32) if precision == "fp32" else u
        assert rel_err(_vmult(op, uu, dtype), ref) < tol


def test_em_bitwise_repeats():
    import torch
    case = make_case((20, 18, 16), 2, material=MAT)
    op = engine_operator(case)
    _em(op, "fp64", 2, 2)
    src = torch.tensor(oracle_problem(case).random(0), dtype=torch.float64, device="cuda")
    a = torch.empty_like(src)
    b = torch.empty_like(src)
    op.vmult(a, src)
    for _ in range(3):
        op.vmult(b, src)
        torch.cuda.synchronize()
        assert torch.equal(a, b)


@pytest.mark.parametrize("p", [1, 2])
def test_em_cg_iterations_match_oracle(p):
    import torch
    from meltpooldg_amd import LinearSolver, LinearSolverData
    case = make_case(12, p, material=MAT)
    pb = oracle_problem(case)
    b = pb.random(1)
    ref = pb.solve(b, rel_tol=1e-10, inv_diag=pb.inverse_diagonal())
    op = engine_operator(case)
    _em(op, "fp64", 2, 2)
    _em(op, "fp32", 2, 2)
    x = torch.zeros(pb.N, dtype=torch.float64, device="cuda")
    it = LinearSolver.solve(op, x, torch.tensor(b, dtype=torch.float64, device="cuda"),
                            LinearSolverData(solver_type="CG", preconditioner_type="Diagonal",
                                             rel_tolerance=1e-10))
    torch.cuda.synchronize()
    print("em cg", p, it, ref["iterations"], flush=True)
    assert it == ref["iterations"]
    assert rel_err(x.cpu().numpy(), ref["x"]) < 1e-9
    xm = torch.zeros_like(x)
    LinearSolver.solve(op, xm, torch.tensor(b, dtype=torch.float64, device="cuda"),
                       LinearSolverData(solver_type="CG", preconditioner_type="Diagonal",
                                        rel_tolerance=1e-10, mixed_precision=True))
    torch.cuda.synchronize()
    assert rel_err(xm.cpu().numpy(), ref["x"]) < 1e-8


def test_em_refused_where_cells_differ():
    from meltpooldg_amd._lib import MpdgError
    case = make_case((4, 4, 4), 2, k_amplitude=0.5)
    op = engine_operator(case)
    with pytest.raises(MpdgError):
        _em(op, "fp64", 2, 2)
    op3 = engine_operator(make_case((4, 4, 4), 3, material=MAT))
    with pytest.raises(MpdgError):
        _em(op3, "fp64", 2, 2)
