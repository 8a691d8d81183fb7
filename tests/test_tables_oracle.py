"""Pins the CPU oracle (oracle/heat_oracle.cpp) against checks that do not
share its code: numpy's Gauss rule and Legendre roots, closed-form identities,
the Kronecker-product closed form, dense brute-force assembly, the golden
fixtures, and the reference's own material known answers."""
import json
import os

import numpy as np
import pytest

from oracle import brute, kron, rng
from oracle import tables as T
from tests.cases import GOLDEN, golden_cases, oracle_problem, rel_err

GAS = {"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0}
TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}


@pytest.mark.parametrize("p", range(1, 8))
def test_tables_against_numpy(oracle_lib, p):
    g, x, w, S, D = oracle_lib.tables(p)
    nodes, xq, wq, S2, D2 = T.tables(p)
    assert np.abs(g - nodes).max() < 1e-14
    assert np.abs(x - xq).max() < 1e-14
    assert np.abs(w - wq).max() < 1e-14
    assert np.abs(S - S2).max() < 1e-13
    assert np.abs(D - D2).max() < 1e-12 * max(1.0, np.abs(D).max())


@pytest.mark.parametrize("p", range(1, 8))
def test_table_identities(oracle_lib, p):
    g, x, w, S, D = oracle_lib.tables(p)
    # Gauss rule with p+1 points is exact to degree 2p+1 on [0,1]
    for k in range(2 * p + 2):
        assert abs((w * x**k).sum() - 1.0 / (k + 1)) < 1e-14
    # partition of unity and its derivative
    assert np.abs(S.sum(axis=1) - 1.0).max() < 1e-14
    assert np.abs(D.sum(axis=1)).max() < 1e-12
    # interior GLL points are the roots of P_p'
    if p > 1:
        dP = np.polynomial.legendre.Legendre.basis(p).deriv()
        assert np.abs(dP(2 * g[1:-1] - 1)).max() < 1e-12
    assert g[0] == 0.0 and g[-1] == 1.0
    # the even-odd symmetries the GPU kernels rely on
    n = p + 1
    assert np.abs(S[::-1, ::-1] - S).max() < 1e-14
    assert np.abs(D[::-1, ::-1] + D).max() < 1e-12
    # S interpolates: sum_i S[q,i] x_i^k = x_q^k for k <= p
    for k in range(n):
        assert np.abs(S @ g**k - x**k).max() < 1e-13


def test_rng_matches_c(oracle_lib):
    pb = oracle_lib.Problem((3, 4, 2), 3)
    assert np.array_equal(pb.random(0), rng.random_vector((3, 4, 2), 3, 0))
    assert np.array_equal(pb.random(12345), rng.random_vector((3, 4, 2), 3, 12345))
    v = pb.random(1)
    assert np.all(v[pb.mask()] == 0.0)
    assert np.all(np.abs(v) <= 1.0)


@pytest.mark.parametrize("n,p,amp,vel", [(3, 2, 0.0, False), (2, 4, 0.5, False), (3, 3, 0.0, True),
                                         (2, 5, 0.5, True), (1, 6, 0.0, False), ((4, 2, 3), 1, 0.5, True)])
def test_oracle_vmult_diag_vs_kronecker(oracle_lib, n, p, amp, vel):
    pb = oracle_lib.Problem(n, p, dt=1e-3, k_amplitude=amp, velocity=vel)
    K = kron.KronOperator(n, p, dt=1e-3, k_amplitude=amp, velocity=vel)
    u = pb.random(0)
    assert rel_err(pb.vmult(u), K.vmult(u)) < 1e-13
    assert rel_err(pb.diagonal(), K.diagonal()) < 1e-13
    assert rel_err(pb.inverse_diagonal(), K.inverse_diagonal()) < 1e-13


@pytest.mark.parametrize("mat", [TWO, dict(TWO, sharp=1), dict(TWO, volume_specific=1, liq_cp=2.0),
                                 dict(TWO, consistent_with_evaporation=1, liq_cp=3.0)])
@pytest.mark.parametrize("vel", [False, True])
def test_oracle_vs_brute_two_phase(oracle_lib, mat, vel):
    n, p = (2, 1, 2), 3
    H = 0.5 * (rng.random_vector(n, p, 5, dirichlet_faces=0) + 1.0)
    A, mask = brute.assemble(n, p, 1e-3, mat, heaviside=H, velocity=vel)
    pb = oracle_lib.Problem(n, p, dt=1e-3, material=mat, heaviside=H, velocity=vel)
    u = pb.random(0)
    assert rel_err(pb.vmult(u), brute.vmult(A, mask, u)) < 1e-13
    assert rel_err(pb.diagonal(), brute.diagonal(A, mask)) < 1e-13


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["name"])
def test_oracle_vs_golden(oracle_lib, case):
    pb = oracle_problem(case)
    assert np.array_equal(pb.mask(), case["mask"])
    assert rel_err(pb.vmult(case["src"]), case["dst"]) < 1e-13
    assert rel_err(pb.diagonal(), case["diag"]) < 1e-13


def test_oracle_symmetry_and_row_sums(oracle_lib):
    # without convection the constrained operator is symmetric
    n, p = (3, 2, 2), 3
    H = 0.5 * (rng.random_vector(n, p, 9, dirichlet_faces=0) + 1.0)
    pb = oracle_lib.Problem(n, p, dt=2e-3, material=TWO, heaviside=H)
    u, v = pb.random(1), pb.random(2)
    assert abs(u @ pb.vmult(v) - v @ pb.vmult(u)) < 1e-12 * abs(u @ pb.vmult(v))
    # no Dirichlet: K 1 = 0, so 1.(A 1) = rho_cp/dt * |Omega| for a constant material
    mat = {"type": "gas", "gas_rho": 2.0, "gas_cp": 3.0, "gas_k": 7.0}
    pb = oracle_lib.Problem((2, 3, 1), 4, dt=0.5, material=mat, dirichlet_faces=0,
                            lo=(0.0, -1.0, 0.5), hi=(2.0, 1.0, 1.0), k_amplitude=0.3)
    ones = np.ones(pb.N)
    Av = pb.vmult(ones)
    assert abs(Av.sum() - 6.0 / 0.5 * 2.0) < 1e-11
    # linearity
    u, v = pb.random(3), pb.random(4)
    assert rel_err(pb.vmult(2.0 * u - 3.0 * v), 2.0 * pb.vmult(u) - 3.0 * pb.vmult(v)) < 1e-13


def test_oracle_pcg_matches_textbook(oracle_lib):
    n, p = 3, 2
    A, mask = brute.assemble(n, p, 1e-3, GAS)
    Ac = brute.constrained(A, mask)
    pb = oracle_lib.Problem(n, p, dt=1e-3)
    b = pb.random(1)
    D = pb.inverse_diagonal()
    x_ref, it_ref = brute.pcg_dense(Ac, b, D, 1e-10)
    r = pb.solve(b, rel_tol=1e-10, inv_diag=D)
    assert r["status"] == 0
    assert r["iterations"] == it_ref
    free = ~mask
    x_exact = np.zeros_like(b)
    x_exact[free] = np.linalg.solve(Ac[np.ix_(free, free)], b[free])
    assert rel_err(r["x"], x_exact) < 1e-8
    assert r["final_res"] <= 1e-10 * r["initial_res"]


def test_oracle_no_convergence(oracle_lib):
    pb = oracle_lib.Problem(4, 2, dt=1.0)
    b = pb.random(1)
    r = pb.solve(b, rel_tol=1e-12, max_it=3, inv_diag=pb.inverse_diagonal())
    assert r["status"] == 1 and r["iterations"] == 3


def test_oracle_gmres_convection(oracle_lib):
    n, p = 2, 3
    mat = dict(GAS, gas_k=1e-2)
    A, mask = brute.assemble(n, p, 5e-4, mat, velocity=True)
    Ac = brute.constrained(A, mask)
    pb = oracle_lib.Problem(n, p, dt=5e-4, material=mat, velocity=True)
    b = pb.random(1)
    free = ~mask
    x_exact = np.zeros_like(b)
    x_exact[free] = np.linalg.solve(Ac[np.ix_(free, free)], b[free])
    for restart in (3, 30):
        r = pb.solve(b, solver="GMRES", restart=restart, rel_tol=1e-12,
                     inv_diag=pb.inverse_diagonal())
        assert r["status"] == 0
        assert rel_err(r["x"], x_exact) < 1e-9


def test_material_known_answers(oracle_lib):
    with open(os.path.join(GOLDEN, "material_known_answers.json")) as f:
        known = json.load(f)
    g = known["material_data"]["gas"]
    l = known["material_data"]["liquid"]
    for row in known["rows"]:
        t = row["type"]
        if "temperature" in row:
            continue  # temperature-dependent rows: tests/test_solidification_oracle.py
        mat = {"type": "gas" if t in ("gas", "liquid") else "gas_liquid",
               "gas_rho": g["density"], "gas_cp": g["specific_heat_capacity"],
               "gas_k": g["thermal_conductivity"], "liq_rho": l["density"],
               "liq_cp": l["specific_heat_capacity"], "liq_k": l["thermal_conductivity"],
               "consistent_with_evaporation": int(t.endswith("evaporation"))}
        if t == "liquid":
            mat.update(gas_rho=l["density"], gas_cp=l["specific_heat_capacity"],
                       gas_k=l["thermal_conductivity"])
        out = oracle_lib.material(row["heaviside"], mat)
        # the reference prints 6 significant digits
        for key, ref in (("k", row["k"]), ("cp", row["cp"]), ("rho", row["rho"])):
            assert out[key] == pytest.approx(ref, rel=5e-6, abs=1e-12), (t, key)
        assert out["rho_cp"] == pytest.approx(out["rho"] * out["cp"], rel=1e-15)
