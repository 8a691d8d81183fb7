"""Pins of the oracle's Robin boundary faces (convection alpha (T - T_inf) and
radiation eps sigma (T^4 - T_inf^4) on box faces): tangent_boundary_loop /
tangent_local_boundary_operation, the boundary part of the diagonal and
rhs_boundary_loop (source/heat/heat_diffuse_operator.cpp:243-269, 352-356,
719-781, 1088-1129).

Independent checks:
* the dense brute-force face assembly (oracle/brute.py: explicit 2D face basis
  matrices from the Vandermonde construction) for the tangent, its diagonal
  and the residual, with Dirichlet faces meeting Robin faces so that the
  tangent's constraint-resolved temperature read is exercised;
* closed forms at a constant temperature c without Dirichlet faces:
  1^T A 1 = rho_cp |Omega| / dt + (alpha + 4 eps sigma c^3) |Gamma| and
  sum(rhs) = -(alpha (c - T_inf) + eps sigma (c^4 - T_inf^4)) |Gamma|;
* the tangent is the derivative of minus the residual (central differences),
  which fixes the factor 4 and the cube of the radiation linearisation;
* the reference's own regression output of the unidirectional_heat_transfer
  case with convection, radiation and both (5 Newton time steps each).

The reference case is 2D (32 x 32 Q1 cells on [0, 0.1]^2, T0 = 1000, the
Robin condition on the right face x = 0.1, every other face insulated), and its
solution does not depend on y. The 3D operator on the extruded mesh of
32 x 1 x 1 cells on [0, 0.1] x [0, 0.1] x [0, Lz] has the same discrete
solution for any depth Lz and an L2 norm sqrt(Lz) times the 2D one. Newton's
residual tolerance is absolute (5e-9), so the depth is chosen small enough
(Lz = 0.1) that the residual, which scales with the face and cell measures,
lies below the 2D run's: with Lz = 1 the round-off floor of the residual
(about 1e-8) is above the tolerance and Newton would not stop.
"""
import json
import os

import numpy as np
import pytest

from oracle import brute, cpu
from oracle.newton import NewtonData, linear_predictor, newton_time_step
from oracle.rng import constrained_mask

HERE = os.path.dirname(os.path.abspath(__file__))
SIGMA = cpu.STEFAN_BOLTZMANN


def _robin(conv, rad, alpha=3.0, eps=0.7, tc=1.5, tr=0.5):
    return dict(convection_faces=conv, convection_coefficient=alpha, convection_t_inf=tc,
                radiation_faces=rad, emissivity=eps, radiation_t_inf=tr)


CASES = [
    ((2, 3, 2), 1, 0b000011, 0b000010, 0b100000),
    ((2, 2, 2), 2, 0b010010, 0b001010, 0b000001),
    ((2, 2, 1), 3, 0b000000, 0b111111, 0b000000),
    ((1, 2, 2), 4, 0b000100, 0b100010, 0b011000),
]


@pytest.mark.parametrize("n,p,conv,rad,dirichlet", CASES)
def test_robin_faces_match_dense_assembly(n, p, conv, rad, dirichlet):
    rb = _robin(conv, rad)
    mat = {"type": "gas", "gas_rho": 2.0, "gas_cp": 1.5, "gas_k": 0.7}
    lo, hi = (0.0, -0.5, 0.25), (1.0, 0.5, 1.0)
    pb = cpu.Problem(n, p, dt=0.05, material=mat, lo=lo, hi=hi, dirichlet_faces=dirichlet,
                     robin=rb)
    rng = np.random.default_rng(17 * p + len(n))
    T = rng.uniform(0.5, 2.0, pb.N)
    T_old = rng.uniform(0.5, 2.0, pb.N)
    src = rng.uniform(-1.0, 1.0, pb.N)
    pb.set_temperature(T)
    A, Mrc, M1, mask = brute.assemble_parts(n, p, 0.05, mat, lo=lo, hi=hi,
                                            dirichlet_faces=dirichlet)
    Af, rf = brute.robin_face_parts(n, p, rb, T, T, mask, lo=lo, hi=hi)
    assert np.abs(Af).max() > 0.0
    # tangent (source constraint-resolved, constrained rows dropped)
    ref = brute.vmult(A + Af, mask, np.where(mask, 0.0, src))
    got = pb.vmult(src)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # diagonal
    dref = brute.diagonal(A + Af, mask)
    assert np.abs(pb.diagonal() - dref).max() <= 1e-12 * np.abs(dref).max()
    # residual: cell part from the dense assembly, face part with T plain
    cell = brute.create_rhs(A, Mrc, M1, mask, T, T_old)
    rref = cell + np.where(mask, 0.0, rf)
    got_r = pb.create_rhs(T, T_old)
    assert np.abs(got_r - rref).max() <= 1e-12 * np.abs(rref).max()


def test_tangent_reads_temperature_with_constraints_resolved():
    # a Dirichlet face next to a radiating face: the tangent uses T = 0 on the
    # constrained nodes (read_dof_values), so it differs from the tangent at
    # the plain temperature on the free rows near the shared edge
    n, p = (2, 2, 2), 2
    rb = _robin(0, 0b000010, eps=0.9)
    mat = {"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0}
    pb = cpu.Problem(n, p, dt=0.1, material=mat, dirichlet_faces=0b000100, robin=rb)
    T = np.full(pb.N, 300.0)  # 4 eps sigma T^3 ~ 5.5, comparable with the cell part
    pb.set_temperature(T)
    mask = constrained_mask(n, p, 0b000100)
    src = np.where(mask, 0.0, np.random.default_rng(3).uniform(-1, 1, pb.N))
    A, _, _, _ = brute.assemble_parts(n, p, 0.1, mat, dirichlet_faces=0b000100)
    Af_res, _ = brute.robin_face_parts(n, p, rb, T, T, mask)
    Af_plain, _ = brute.robin_face_parts(n, p, rb, T, T, np.zeros_like(mask))
    got = pb.vmult(src)
    ok = brute.vmult(A + Af_res, mask, src)
    wrong = brute.vmult(A + Af_plain, mask, src)
    assert np.abs(got - ok).max() <= 1e-12 * np.abs(ok).max()
    assert np.abs(got - wrong).max() > 1e-6 * np.abs(ok).max()


@pytest.mark.parametrize("p", [1, 3])
def test_robin_closed_forms_at_constant_temperature(p):
    n = (2, 3, 1)
    lo, hi = (0.0, 0.0, 0.0), (2.0, 0.5, 1.5)
    rb = _robin(0b100001, 0b010110, alpha=4.0, eps=0.8, tc=2.0, tr=1.0)
    rho, cp, dt, c = 3.0, 2.0, 0.25, 5.0
    mat = {"type": "gas", "gas_rho": rho, "gas_cp": cp, "gas_k": 1.3}
    pb = cpu.Problem(n, p, dt=dt, material=mat, lo=lo, hi=hi, dirichlet_faces=0, robin=rb)
    T = np.full(pb.N, c)
    pb.set_temperature(T)
    ext = [hi[d] - lo[d] for d in range(3)]
    area = lambda f: ext[[e for e in range(3) if e != f >> 1][0]] * ext[[e for e in range(3) if e != f >> 1][1]]
    conv_area = sum(area(f) for f in range(6) if (rb["convection_faces"] >> f) & 1)
    rad_area = sum(area(f) for f in range(6) if (rb["radiation_faces"] >> f) & 1)
    vol = ext[0] * ext[1] * ext[2]
    one = np.ones(pb.N)
    expect = rho * cp * vol / dt + 4.0 * conv_area + 4.0 * 0.8 * SIGMA * c ** 3 * rad_area
    assert abs(one @ pb.vmult(one) - expect) <= 1e-12 * expect
    r = pb.create_rhs(T, T)
    expect_r = -(4.0 * (c - 2.0) * conv_area + 0.8 * SIGMA * (c ** 4 - 1.0) * rad_area)
    assert abs(r.sum() - expect_r) <= 1e-12 * abs(expect_r)


def test_tangent_is_derivative_of_residual():
    n, p = (2, 2, 2), 2
    rb = _robin(0b000011, 0b001001, alpha=2.0, eps=0.9, tc=1.0, tr=0.3)
    mat = {"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 0.5}
    pb = cpu.Problem(n, p, dt=0.1, material=mat, dirichlet_faces=0, robin=rb)
    rng = np.random.default_rng(5)
    # large temperatures so that the radiation term dominates the face part
    T = rng.uniform(800.0, 1200.0, pb.N)
    T_old = rng.uniform(800.0, 1200.0, pb.N)
    v = rng.uniform(-1.0, 1.0, pb.N)
    pb.set_temperature(T)
    Jv = pb.vmult(v)
    eps = 1e-3
    fd = -(pb.create_rhs(T + eps * v, T_old) - pb.create_rhs(T - eps * v, T_old)) / (2 * eps)
    assert np.abs(fd - Jv).max() <= 1e-7 * np.abs(Jv).max()
    # the face part alone, to make sure the check is not dominated by the cells
    pb0 = cpu.Problem(n, p, dt=0.1, material=mat, dirichlet_faces=0)
    face = Jv - pb0.vmult(v)
    assert np.abs(face).max() > 1e-2 * np.abs(Jv).max()


CASES_REF = ["with_convection", "with_radiation", "with_radiation_and_convection"]


@pytest.mark.parametrize("case", CASES_REF)
def test_oracle_time_steps_match_reference_robin_output(case):
    g = json.load(open(os.path.join(HERE, "golden", "unidirectional_heat_transfer_robin.json")))
    s = g["setup"]
    c = g["cases"][case]
    n = 2 ** s["global_refinements"]
    rb = dict(convection_faces=c["convection_faces"],
              convection_coefficient=c["convection_coefficient"],
              convection_t_inf=c["convection_temperature_infinity"],
              radiation_faces=c["radiation_faces"], emissivity=c["emissivity"],
              radiation_t_inf=c["radiation_temperature_infinity"])
    mat = {"type": "gas", "gas_rho": s["density"], "gas_cp": s["specific_heat_capacity"],
           "gas_k": s["thermal_conductivity"]}
    dt = s["time_step"]
    depth = 0.1
    pb = cpu.Problem((n, 1, 1), s["degree"], dt=dt, material=mat, lo=(0.0, 0.0, 0.0),
                     hi=(0.1, 0.1, depth), dirichlet_faces=0, robin=rb)
    T_old = np.full(pb.N, s["initial_temperature"])
    T_old_old, dt_old = T_old, 0.0
    data = NewtonData(residual_tolerance=s["nlsolve"]["residual_tolerance"])
    for step, ref in enumerate(c["l2_norms"]):
        pred = linear_predictor(T_old, T_old_old, dt, dt_old)
        T, rep = newton_time_step(pb, pred, T_old, data=data, solver="GMRES",
                                  rel_tol=s["linear_solver"]["rel_tolerance"])
        assert rep["converged"], rep
        norm = pb.l2_norm(T) / np.sqrt(depth)
        assert float(f"{norm:.5e}") == ref, (case, step, norm, ref)
        T_old_old, T_old, dt_old = T_old, T, dt
