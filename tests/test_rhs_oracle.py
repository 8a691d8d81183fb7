"""Pins of the oracle's create_rhs (the residual right-hand side of the heat
operator, source/heat/heat_diffuse_operator.cpp:393-529, 783-806) on the CPU.

Independent checks, each chosen so that a dropped term, a wrong sign or a
constraint mistake fails one of them:
* the dense brute-force assembly (oracle/brute.py: 3D basis built from a
  monomial Vandermonde solve, full element matrices) on tiny meshes;
* rhs(T, T_old = 0, f = 0) = -J T on the free rows when T is zero on the
  constrained nodes (J = the already pinned tangent vmult);
* a constant temperature field with T_old = T and f = 0 gives 0 on every free
  row (no gradient, no time change), also with convection and variable k;
* the heat-source term alone sums to the box volume for f = 1 without
  Dirichlet faces (sum_i int phi_i = |Omega|);
* create_rhs adds into dst (the reference's loop runs with zero_dst = false)
  and leaves the constrained rows of dst untouched.
"""
import numpy as np
import pytest

from oracle import brute, cpu
from oracle.rng import constrained_mask

GAS = {"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0}
TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 2.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}


def _fields(N, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1.0, 1.0, N) for _ in range(4)]


@pytest.mark.parametrize("n,p,kw", [
    ((2, 3, 2), 2, {}),
    ((2, 2, 2), 3, dict(k_amplitude=0.5, velocity=True)),
    ((3, 2, 2), 1, dict(dirichlet_faces=0b100101)),
    ((2, 2, 1), 4, dict(dirichlet_faces=0)),
])
def test_create_rhs_matches_dense_assembly(n, p, kw):
    pb = cpu.Problem(n, p, dt=2e-3, **kw)
    A, Mrc, M1, mask = brute.assemble_parts(n, p, 2e-3, GAS, **kw)
    T, To, f, d0 = _fields(pb.N, 1)
    ref = brute.create_rhs(A, Mrc, M1, mask, T, To, f, dst=d0)
    assert np.abs(pb.create_rhs(T, To, f, dst=d0) - ref).max() < 1e-13 * np.abs(ref).max()


def test_create_rhs_two_phase_matches_dense_assembly():
    n, p = (2, 2, 2), 2
    rng = np.random.default_rng(3)
    from oracle.rng import lattice_shape
    n1 = lattice_shape(n, p)
    H = rng.uniform(0.0, 1.0, n1[0] * n1[1] * n1[2])
    pb = cpu.Problem(n, p, dt=1e-3, material=TWO, heaviside=H, velocity=True)
    A, Mrc, M1, mask = brute.assemble_parts(n, p, 1e-3, TWO, heaviside=H, velocity=True)
    T, To, f, _ = _fields(pb.N, 2)
    ref = brute.create_rhs(A, Mrc, M1, mask, T, To, f)
    assert np.abs(pb.create_rhs(T, To, f) - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_rhs_of_homogeneous_field_is_minus_tangent(p):
    pb = cpu.Problem((4, 3, 3), p, dt=1e-3, k_amplitude=0.5, velocity=True)
    T = pb.random(5)  # zero on the constrained nodes
    out = pb.create_rhs(T, np.zeros(pb.N))
    assert np.abs(out + pb.vmult(T)).max() < 1e-12 * np.abs(out).max()


@pytest.mark.parametrize("p", [1, 3, 5])
def test_rhs_of_constant_steady_field_vanishes(p):
    pb = cpu.Problem((3, 4, 2), p, dt=1e-3, k_amplitude=0.5, velocity=True)
    T = np.full(pb.N, 2.5)
    out = pb.create_rhs(T, T)
    # scale: the diffusion part of one row for a unit nodal jump
    scale = np.abs(pb.create_rhs(pb.random(1), np.zeros(pb.N))).max()
    assert np.abs(out).max() < 1e-13 * scale


@pytest.mark.parametrize("p", [1, 2, 4])
def test_heat_source_integrates_to_volume(p):
    lo, hi = (0.0, -0.5, 0.25), (2.0, 1.0, 1.0)
    pb = cpu.Problem((3, 2, 2), p, dt=1e-3, lo=lo, hi=hi, dirichlet_faces=0)
    zero = np.zeros(pb.N)
    out = pb.create_rhs(zero, zero, np.ones(pb.N))
    vol = np.prod(np.subtract(hi, lo))
    assert abs(out.sum() - vol) < 1e-13 * vol
    # and the time term of a uniform change: rho_cp/dt * |Omega| * (T - T_old) with sign -1
    out = pb.create_rhs(np.full(pb.N, 3.0), np.full(pb.N, 1.0))
    assert abs(out.sum() + 2.0 / 1e-3 * vol) < 1e-12 * 2.0 / 1e-3 * vol


def test_create_rhs_accumulates_and_keeps_constrained_rows():
    n, p = (3, 3, 2), 2
    pb = cpu.Problem(n, p, dt=1e-3)
    T, To, f, d0 = _fields(pb.N, 4)
    mask = constrained_mask(n, p, 0x3F)
    fresh = pb.create_rhs(T, To, f)
    acc = pb.create_rhs(T, To, f, dst=d0)
    assert np.array_equal(acc[mask], d0[mask])
    assert np.abs(fresh[mask]).max() == 0.0
    assert np.abs(acc - (d0 + fresh)).max() < 1e-14 * np.abs(acc).max()
