"""GPU parity of the constant prescribed velocity (advection rho_cp u . grad T
with u a constant field) against the oracle (pinned in
tests/test_velocity_oracle.py), through the C ABI: tangent vmult in every
kernel form (line, cell for degree <= 2, matrix-core line form for FP32 degree
>= 3), inverse diagonal and create_rhs, with and without the two-phase
material. Bars: FP64 within 1e-12 relative (max-norm scaled by the largest
entry), FP32 within 2e-5.
"""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 2.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}
U = (-0.7, 0.4, 1.3)


def _t(a, dtype=None):
    import torch
    return torch.tensor(a, dtype=dtype or torch.float64, device="cuda")


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_constant_velocity_vs_oracle(p):
    import torch
    from oracle.rng import random_vector
    n = (6, 5, 4)
    H = 0.5 * (random_vector(n, p, 11, dirichlet_faces=0) + 1.0)
    cases = [make_case(n, p, dt=0.05, velocity="constant", velocity_value=U, faces=0b101101),
             make_case(n, p, dt=0.05, velocity="constant", velocity_value=U, material=dict(TWO),
                       heaviside=H, faces=0)]
    for case in cases:
        print("constant velocity", p, case["material"]["type"], flush=True)
        pb = oracle_problem(case)
        rng = np.random.default_rng(p)
        src = rng.uniform(-1.0, 1.0, pb.N)
        ref = pb.vmult(src)
        variants = ["line"] + (["cell"] if p <= 2 else [])
        for v in variants:
            op = engine_operator(case)
            op.set_tuning(variant=v, precision="fp64")
            dst = torch.empty(pb.N, dtype=torch.float64, device="cuda")
            op.vmult(dst, _t(src))
            assert rel_err(_np(dst), ref) < 1e-12, v
        inv = torch.empty(pb.N, dtype=torch.float64, device="cuda")
        op.compute_inverse_diagonal_from_matrixfree(inv)
        assert rel_err(_np(inv), pb.inverse_diagonal()) < 1e-12
        T, To, f, d0 = (rng.uniform(-1.0, 1.0, pb.N) for _ in range(4))
        dst = _t(d0)
        op.create_rhs(dst, _t(To), _t(T), _t(f))
        assert rel_err(_np(dst), pb.create_rhs(T, To, f, dst=d0)) < 1e-12
        op32 = engine_operator(case, precision="fp32")
        for v in ["line"] + (["mfma"] if p >= 3 else []) + (["cell"] if p <= 2 else []):
            op32.set_tuning(variant=v, precision="fp32")
            d32 = torch.empty(pb.N, dtype=torch.float32, device="cuda")
            op32.vmult(d32, _t(src, torch.float32))
            assert rel_err(_np(d32), ref) < 2e-5, v
