"""GPU parity of the MFMA form of the line vmult (FP32 1D contractions on
v_mfma_f32_4x4x1_16b, degree >= 3) against the FP64 oracle, through the C ABI.

The FP32 bar of the FMA form applies: within 2e-5 relative (max-norm scaled by
the max entry) of the FP64 oracle; the mixed-precision solves with the MFMA
form reach the FP64 oracle solution to the stated tolerance.
"""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}


def _t(a, dtype=None):
    import torch
    return torch.tensor(a, dtype=dtype or torch.float64, device="cuda")


def _vmult32(op, src_np):
    import torch
    src = _t(src_np.astype(np.float32), torch.float32)
    dst = torch.full_like(src, float("nan"))
    op.vmult(dst, src)
    torch.cuda.synchronize()
    return dst.double().cpu().numpy()


@pytest.mark.parametrize("p", [3, 4, 5, 6])
def test_mfma_vmult_vs_oracle(p):
    cases = [make_case((7, 5, 4), p, k_amplitude=0.5),
             make_case((5, 3, 6), p, velocity=True, dt=5e-4, faces=0b011010),
             make_case((1, 1, 1), p)]
    for case in cases:
        print("mfma", p, case["n"], case["k_amplitude"], case["velocity"], flush=True)
        pb = oracle_problem(case)
        op = engine_operator(case, precision="fp32")
        op.set_tuning(variant="mfma", z_layers=2)
        assert op.get_tuning()["variant"] == "mfma"
        u = pb.random(0)
        assert rel_err(_vmult32(op, u), pb.vmult(u)) < 2e-5


@pytest.mark.parametrize("p", [3, 4])
def test_mfma_two_phase_vs_oracle(p):
    from oracle.rng import random_vector
    n = (6, 4, 3)
    H = 0.5 * (random_vector(n, p, 11, dirichlet_faces=0) + 1.0)
    for vel in (False, True):
        case = make_case(n, p, material=dict(TWO), heaviside=H, velocity=vel)
        print("mfma two-phase", p, vel, flush=True)
        pb = oracle_problem(case)
        op = engine_operator(case, precision="fp32")
        op.set_tuning(variant="mfma")
        u = pb.random(0)
        assert rel_err(_vmult32(op, u), pb.vmult(u)) < 2e-5


def test_mfma_mixed_cg_reaches_fp64_solution():
    """FP32-MFMA cell loop inside the FP64 defect-correction CG: the solution
    matches the FP64 oracle CG solution to the tolerance of the solve."""
    import torch
    from meltpooldg_amd import LinearSolver, LinearSolverData
    case = make_case((12, 10, 9), 4, k_amplitude=0.5)
    pb = oracle_problem(case)
    b = pb.random(1)
    ref = pb.solve(b, rel_tol=1e-12, inv_diag=pb.inverse_diagonal())
    op = engine_operator(case)
    op.set_tuning(variant="mfma", precision="fp32")
    x = torch.zeros(op.n_dofs(), dtype=torch.float64, device="cuda")
    data = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal", rel_tolerance=1e-9,
                            mixed_precision=True)
    LinearSolver.solve(op, x, _t(b), data)
    rep = LinearSolver.last_report
    print("mfma mixed", rep, flush=True)
    assert rep["final_residual"] <= 1e-9 * rep["initial_residual"]
    torch.cuda.synchronize()
    assert rel_err(x.cpu().numpy(), ref["x"]) < 1e-7


def test_mfma_rejected_for_fp64_and_low_degree():
    op = engine_operator(make_case((4, 4, 4), 4))
    with pytest.raises(Exception):
        op.set_tuning(variant="mfma", precision="fp64")
    op2 = engine_operator(make_case((4, 4, 4), 2))
    with pytest.raises(Exception):
        op2.set_tuning(variant="mfma", precision="fp32")
