"""Heaviside of a level set on FE_Q_iso_Q1(p) (the reference's
do_level_set_temperature_gradient_interpolation path,
source/heat/heat_diffuse_operator.cpp:834-838, 982-993): the engine
interpolates the iso field to the heat element's support points on the device
(mpdg_heat_set_heaviside_iso_q1); the vmult and inverse diagonal then match
the oracle given the oracle's own interpolation (oracle/iso_q1.py) to 1e-12,
for every degree, both two-phase transition laws and with a velocity."""
import numpy as np
import pytest

from oracle import iso_q1
from oracle.rng import random_vector
from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 2.0, "liq_k": 50.0}


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("sharp", [False, True])
def test_iso_q1_heaviside_vs_oracle(p, sharp):
    import torch
    n = (4, 3, 3)
    m = dict(TWO, sharp=1 if sharp else 0)
    h_iso = 0.5 * (random_vector(n, p, 21, dirichlet_faces=0) + 1.0)
    h_gll = iso_q1.iso_q1_to_gll(n, p, h_iso)
    for vel in (False, True):
        print("iso-q1", p, sharp, vel, flush=True)
        pb = oracle_problem(make_case(n, p, material=m, heaviside=h_gll, velocity=vel))
        op = engine_operator(make_case(n, p, material=m, heaviside=None, velocity=vel))
        op.set_heaviside(torch.tensor(h_iso, dtype=torch.float64, device="cuda"), fe="FE_Q_iso_Q1")
        u = pb.random(0)
        src = torch.tensor(u, dtype=torch.float64, device="cuda")
        dst = torch.full_like(src, float("nan"))
        op.vmult(dst, src)
        d = torch.full_like(src, float("nan"))
        op.compute_inverse_diagonal_from_matrixfree(d)
        torch.cuda.synchronize()
        assert rel_err(dst.cpu().numpy(), pb.vmult(u)) < 1e-12
        assert rel_err(d.cpu().numpy(), pb.inverse_diagonal()) < 1e-12


def test_iso_q1_differs_from_plain_heaviside():
    # the interpolation is not the identity from degree 3 on
    import torch
    n, p = (3, 3, 3), 4
    h = 0.5 * (random_vector(n, p, 22, dirichlet_faces=0) + 1.0)
    ops = []
    for fe in ("FE_Q", "FE_Q_iso_Q1"):
        op = engine_operator(make_case(n, p, material=TWO, heaviside=None))
        op.set_heaviside(torch.tensor(h, dtype=torch.float64, device="cuda"), fe=fe)
        ops.append(op)
    u = torch.tensor(random_vector(n, p, 0), dtype=torch.float64, device="cuda")
    y = [torch.empty_like(u) for _ in ops]
    for op, yy in zip(ops, y):
        op.vmult(yy, u)
    torch.cuda.synchronize()
    assert float((y[0] - y[1]).abs().max()) > 1e-6 * float(y[0].abs().max())
