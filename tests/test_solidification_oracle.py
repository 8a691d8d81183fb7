"""Pins of the oracle's temperature-dependent (solidification) materials:
Material liquid_solid / gas_liquid_solid (include/meltpooldg/core/
material.templates.hpp:296-495, 540-756), the rho_cp combination of
get_material_parameters(_and_derivatives) (source/heat/heat_diffuse_operator.cpp:
1133-1210), the solidification terms of the tangent (:1003-1014, 1044-1056)
and the material evaluated at T_new in the residual (rhs_cell_loop :393-529).

Independent checks:
* the reference's own known answers: unit_tests/material.cc prints k, cp, rho
  and their temperature derivatives for liquid_solid at T = 20, 60 and
  gas_liquid_solid at (H, T) = (0.2, 20), (0.6, 60) (unit_tests/material.output,
  transcribed in tests/golden/material_known_answers.json); both the C++
  oracle and the numpy restatement in oracle/brute.py must print them;
* the sharp transitions and the pure phases by their definitions;
* the dense brute-force assembly (oracle/brute.py solidification_parts:
  explicit 3D quadrature, Vandermonde basis, own material code) for the
  tangent's vmult and diagonal and for the residual, on meshes where the
  temperature field crosses the whole mushy zone, with and without a
  constant velocity and a heaviside, with partial Dirichlet faces;
* the tangent is the derivative of minus the residual (central differences),
  which fixes every derivative term's sign, factor and argument;
* Newton converges quadratically with this tangent;
* the apparent capacity of a latent heat of fusion (Heat::ApparentCapacity,
  source/heat/apparent_capacity.cpp; poly4_bell, constant, qlq) against the
  properties the reference's own unit test asserts
  (unit_tests/heat/apparent_capacity.cpp: L = 338e3, solidus 271, liquidus
  275): it vanishes with its derivative at the solidus and liquidus, its
  midpoint-rule integral over the mushy zone (1000 steps) is L to 1e-8, and
  its temperature derivative matches central differences; the C++ oracle and
  the independent numpy statement in oracle/brute.py agree.
"""
import json
import os

import numpy as np
import pytest

from oracle import brute, cpu
from oracle.newton import NewtonData, newton_time_step
from oracle.rng import constrained_mask

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "material_known_answers.json")))


def _unit_test_material(kind):
    md = KNOWN["material_data"]
    g, l, s = md["gas"], md["liquid"], md["solid"]
    return {"type": kind, "gas_rho": g["density"], "gas_cp": g["specific_heat_capacity"],
            "gas_k": g["thermal_conductivity"], "liq_rho": l["density"],
            "liq_cp": l["specific_heat_capacity"], "liq_k": l["thermal_conductivity"],
            "sol_rho": s["density"], "sol_cp": s["specific_heat_capacity"],
            "sol_k": s["thermal_conductivity"], "solidus": md["solidus_temperature"],
            "liquidus": md["liquidus_temperature"], "solid_liquid_sharp": 0, "sharp": 0}


def test_known_answers_of_the_reference_unit_test():
    rows = [r for r in KNOWN["rows"] if "temperature" in r]
    assert len(rows) == 4
    for row in rows:
        mat = _unit_test_material(row["type"])
        out = cpu.material_state(row["heaviside"], row["temperature"], mat)
        rc, k, drc, dk = brute.material_t(row["heaviside"], row["temperature"], mat)
        # the reference prints 6 significant digits
        for key in ("k", "cp", "rho", "d_k", "d_cp", "d_rho"):
            assert out[key] == pytest.approx(row[key], rel=5e-6), (row, key)
        assert out["rho_cp"] == pytest.approx(row["rho"] * row["cp"], rel=1e-5)
        assert out["d_rho_cp"] == pytest.approx(row["d_cp"] * row["rho"] + row["d_rho"] * row["cp"],
                                                rel=1e-5)
        for a, b in ((rc, out["rho_cp"]), (k, out["k"]), (drc, out["d_rho_cp"]), (dk, out["d_k"])):
            assert float(a) == pytest.approx(b, rel=1e-14)


def test_pure_phases_and_sharp_transitions():
    mat = _unit_test_material("liquid_solid")
    # outside the mushy zone: the phase values, no derivative
    for T, ph in ((150.0, "liq"), (-5.0, "sol")):
        out = cpu.material_state(0.0, T, mat)
        assert out["k"] == mat[ph + "_k"] and out["d_k"] == 0.0 and out["d_rho_cp"] == 0.0
        assert out["rho_cp"] == mat[ph + "_rho"] * mat[ph + "_cp"]
    # mushy midpoint: smoothstep(1/2) = 1/2, slope -(s - l) 1.5 / (T_l - T_s)
    out = cpu.material_state(0.0, 50.0, mat)
    assert out["k"] == pytest.approx(35.0, rel=1e-15)
    assert out["d_k"] == pytest.approx(10.0 * 1.5 / 100.0, rel=1e-15)
    # sharp solid/liquid transition at the solidus: pure phases either side
    sharp = dict(mat, solid_liquid_sharp=1, solidus=40.0, liquidus=40.0)
    assert cpu.material_state(0.0, 39.999, sharp)["k"] == mat["sol_k"]
    assert cpu.material_state(0.0, 40.0, sharp)["k"] == mat["liq_k"]
    assert cpu.material_state(0.0, 39.0, sharp)["d_k"] == 0.0
    # sharp two-phase blend: H > 0.5 is the liquid/solid material, else gas;
    # volume-specific capacity blends rho cp of each phase
    gls = dict(_unit_test_material("gas_liquid_solid"), sharp=1, volume_specific=1)
    hi = cpu.material_state(0.51, 50.0, gls)
    lo = cpu.material_state(0.5, 50.0, gls)
    assert hi["k"] == pytest.approx(35.0, rel=1e-15) and lo["k"] == gls["gas_k"]
    assert lo["d_k"] == 0.0 and lo["rho_cp"] == gls["gas_rho"] * gls["gas_cp"]
    vc = 8000.0 * 1000.0 + (7000.0 * 900.0 - 8000.0 * 1000.0) * 0.5
    assert hi["rho_cp"] == pytest.approx(vc, rel=1e-15)


STEEL = {"gas_rho": 3925.0, "gas_cp": 230.0, "gas_k": 27.7815, "liq_rho": 7850.0,
         "liq_cp": 460.0, "liq_k": 55.563, "sol_rho": 7850.0, "sol_cp": 700.0, "sol_k": 17.0,
         "solidus": 1966.0, "liquidus": 1974.0, "solid_liquid_sharp": 0}

CASES = [
    # n, p, material, sharp two-phase, volume specific, velocity, dirichlet faces
    ((3, 2, 1), 1, "liquid_solid", 0, 0, None, 0),
    ((2, 2, 2), 2, "liquid_solid", 0, 0, (-0.3, 0.2, 0.1), 0b000011),
    ((2, 1, 2), 3, "gas_liquid_solid", 1, 0, (-0.3, 0.0, 0.0), 0b010000),
    ((2, 2, 1), 2, "gas_liquid_solid", 0, 1, None, 0b100100),
    ((1, 2, 2), 4, "gas_liquid_solid", 0, 0, (0.1, 0.4, -0.2), 0),
]


def _fields(n, p, seed, span=(1958.0, 1982.0)):
    """Temperature crossing the mushy zone [1966, 1974] (linear in x plus noise),
    an old temperature close to it and a heaviside in [0, 1]."""
    rng = np.random.default_rng(seed)
    n1 = [c * p + 1 for c in n]
    x = np.linspace(0.0, 1.0, n1[0])
    T = (span[0] + (span[1] - span[0]) * x)[None, None, :] + 0.0 * np.zeros(n1[::-1])
    T = T + rng.uniform(-1.0, 1.0, T.shape)
    T = T.reshape(-1)
    T_old = T + rng.uniform(-0.5, 0.5, T.size)
    H = rng.uniform(0.0, 1.0, T.size)
    return T, T_old, H


def _problem(n, p, mat, sharp, vs, vel, faces, T, T_old, H, dt=0.05, lo=(0.0, 0.0, 0.0),
             hi=(0.1, 0.1, 0.1)):
    m = dict(STEEL, type=mat, sharp=sharp, volume_specific=vs)
    return cpu.Problem(n, p, dt=dt, material=m, heaviside=H if mat == "gas_liquid_solid" else None,
                       lo=lo, hi=hi, dirichlet_faces=faces,
                       velocity="constant" if vel else None, velocity_value=vel,
                       temperature=T, temperature_old=T_old), m


@pytest.mark.parametrize("n,p,mat,sharp,vs,vel,faces", CASES)
def test_tangent_and_residual_match_dense_assembly(n, p, mat, sharp, vs, vel, faces):
    T, T_old, H = _fields(n, p, 11 + p)
    pb, m = _problem(n, p, mat, sharp, vs, vel, faces, T, T_old, H)
    mask = constrained_mask(n, p, faces)
    J, r = brute.solidification_parts(n, p, 0.05, m, T, T_old,
                                      heaviside=H if mat == "gas_liquid_solid" else None,
                                      velocity="constant" if vel else None,
                                      velocity_value=vel or (0.0, 0.0, 0.0),
                                      hi=(0.1, 0.1, 0.1))
    Jc = brute.constrained(J, mask)
    src = np.random.default_rng(3).uniform(-1.0, 1.0, pb.N)
    ref = Jc @ src
    out = pb.vmult(src)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    d_ref = np.where(mask, 0.0, np.diag(J))
    d = pb.diagonal()
    assert np.abs(d - d_ref).max() <= 1e-12 * np.abs(d_ref).max()
    rhs = pb.create_rhs(T, T_old)
    r_ref = np.where(mask, 0.0, r)
    assert np.abs(rhs - r_ref).max() <= 1e-11 * np.abs(r_ref).max()
    # the solidification terms are present: the plain (temperature-frozen)
    # tangent differs
    frozen = brute.solidification_parts(n, p, 0.05, dict(m, liquidus=1e9, solidus=-1e9), T, T_old,
                                        heaviside=H, velocity="constant" if vel else None,
                                        velocity_value=vel or (0.0, 0.0, 0.0),
                                        hi=(0.1, 0.1, 0.1))[0]
    assert np.abs(brute.constrained(frozen, mask) @ src - ref).max() > 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("n,p,mat,sharp,vs,vel,faces", CASES[:4])
def test_tangent_is_the_derivative_of_minus_the_residual(n, p, mat, sharp, vs, vel, faces):
    # temperatures strictly inside the mushy zone (no kink of the solid fraction
    # within the difference step)
    T, T_old, H = _fields(n, p, 5, span=(1967.5, 1972.5))
    T = np.clip(T, 1967.0, 1973.0)
    pb, _ = _problem(n, p, mat, sharp, vs, vel, faces, T, T_old, H)
    mask = pb.mask()
    delta = np.where(mask, 0.0, np.random.default_rng(9).uniform(-1.0, 1.0, pb.N))
    eps = 1e-4
    rp = pb.create_rhs(T + eps * delta, T_old)
    rm = pb.create_rhs(T - eps * delta, T_old)
    fd = -(rp - rm) / (2.0 * eps)
    Jd = pb.vmult(delta)
    assert np.abs(fd - Jd).max() <= 1e-6 * np.abs(Jd).max()


def test_newton_converges_quadratically_through_the_mushy_zone():
    n, p = (8, 1, 1), 2
    T, T_old, _ = _fields(n, p, 21)
    T_old = T.copy()
    pb, _ = _problem(n, p, "liquid_solid", 0, 0, (-0.01, 0.0, 0.0), 0, T, T_old, None, dt=0.5)
    data = NewtonData(residual_tolerance=1e-9, field_correction_tolerance=1e-10)
    Tn, rep = newton_time_step(pb, T, T_old, data=data, rel_tol=1e-13)
    assert rep["converged"] and rep["iterations"] <= 5
    # the converged state is a root of the residual
    assert np.abs(pb.create_rhs(Tn, T_old)).max() < 1e-9


AC_MATERIAL = {"type": "liquid_solid", "liq_rho": 1.0, "liq_cp": 0.0, "liq_k": 1.0,
               "sol_rho": 1.0, "sol_cp": 0.0, "sol_k": 1.0, "solidus": 271.0,
               "liquidus": 275.0, "latent_heat": 338e3}


@pytest.mark.parametrize("kind", ["poly4_bell", "constant", "qlq"])
def test_apparent_capacity_follows_the_reference_unit_test(kind):
    m = dict(AC_MATERIAL, apparent_capacity=kind)
    # with c_p = 0 in both phases the material's c_p is the apparent capacity
    ac = lambda T: cpu.material_state(0.0, T, m)["cp"]
    dac = lambda T: cpu.material_state(0.0, T, m)["d_cp"]
    Ts, Tl, L = m["solidus"], m["liquidus"], m["latent_heat"]
    if kind != "constant":
        for T in (Ts, Tl):
            assert abs(ac(T)) <= 1e-6 and abs(dac(T)) <= 1e-6
    n = 1000
    h = (Tl - Ts) / n
    integral = sum(ac(Ts + (i + 0.5) * h) for i in range(n)) * h
    assert abs(integral - L) <= 1e-8 * L, integral
    eps = 1e-5
    T = Ts + 0.1
    for _ in range(3):
        fd = (ac(T + eps) - ac(T - eps)) / (2.0 * eps)
        assert abs(dac(T) - fd) <= 1e-5 * max(1.0, abs(fd)), (kind, T)
        T += (Tl - Ts) / 5.0
    # outside the mushy zone nothing is added
    assert ac(Ts - 0.5) == 0.0 and ac(Tl + 0.5) == 0.0
    # the independent numpy statement
    grid = np.linspace(270.0, 276.0, 61)
    c, dc = brute.apparent_capacity_t(grid, m)
    ref = np.array([ac(t) for t in grid])
    dref = np.array([dac(t) for t in grid])
    assert np.abs(c - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1.0)
    assert np.abs(dc - dref).max() <= 1e-12 * max(np.abs(dref).max(), 1.0)


LATENT_CASES = [
    ((3, 2, 1), 1, "liquid_solid", 0, (-0.3, 0.2, 0.0), 0, "qlq"),
    ((2, 2, 2), 2, "gas_liquid_solid", 1, None, 0b000011, "poly4_bell"),
    ((2, 1, 2), 3, "gas_liquid_solid", 0, (0.1, 0.0, -0.2), 0, "constant"),
]


@pytest.mark.parametrize("n,p,mat,sharp,vel,faces,kind", LATENT_CASES)
def test_latent_heat_tangent_and_residual(n, p, mat, sharp, vel, faces, kind):
    T, T_old, H = _fields(n, p, 23 + p)
    pb, m = _problem(n, p, mat, sharp, 0, vel, faces, T, T_old, H)
    m.update(latent_heat=2.7e5, apparent_capacity=kind)
    pb = cpu.Problem(n, p, dt=0.05, material=m, heaviside=H if mat == "gas_liquid_solid" else None,
                     lo=(0.0, 0.0, 0.0), hi=(0.1, 0.1, 0.1), dirichlet_faces=faces,
                     velocity="constant" if vel else None, velocity_value=vel,
                     temperature=T, temperature_old=T_old)
    mask = constrained_mask(n, p, faces)
    J, r = brute.solidification_parts(n, p, 0.05, m, T, T_old,
                                      heaviside=H if mat == "gas_liquid_solid" else None,
                                      velocity="constant" if vel else None,
                                      velocity_value=vel or (0.0, 0.0, 0.0), hi=(0.1, 0.1, 0.1))
    src = np.random.default_rng(4).uniform(-1.0, 1.0, pb.N)
    ref = brute.constrained(J, mask) @ src
    assert np.abs(pb.vmult(src) - ref).max() <= 1e-12 * np.abs(ref).max()
    d_ref = np.where(mask, 0.0, np.diag(J))
    assert np.abs(pb.diagonal() - d_ref).max() <= 1e-12 * np.abs(d_ref).max()
    r_ref = np.where(mask, 0.0, r)
    assert np.abs(pb.create_rhs(T, T_old) - r_ref).max() <= 1e-11 * np.abs(r_ref).max()
    # derivative of minus the residual, inside the mushy zone away from the
    # qlq ramp's joints (central differences with a small step)
    Tm = np.clip(T, 1967.2, 1972.8)
    pb.set_temperature(Tm)
    delta = np.where(pb.mask(), 0.0, np.random.default_rng(8).uniform(-1.0, 1.0, pb.N))
    e = 1e-6
    fd = -(pb.create_rhs(Tm + e * delta, T_old) - pb.create_rhs(Tm - e * delta, T_old)) / (2 * e)
    Jd = pb.vmult(delta)
    assert np.abs(fd - Jd).max() <= 1e-4 * np.abs(Jd).max()
