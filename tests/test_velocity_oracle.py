"""Pins of the oracle's constant prescribed velocity (the advection term
rho_cp u . grad T of tangent_local_cell_operation and rhs_cell_loop,
source/heat/heat_diffuse_operator.cpp:1034-1042, 493-523, with u the nodal
interpolant of a constant prescribed_velocity function, which is the constant
itself at every point).

* dense brute-force assembly (oracle/brute.py) for the tangent, its diagonal
  and the residual, with Dirichlet faces and two-phase material;
* closed form without Dirichlet faces: for T = a . x (linear, reproduced by
  every FE_Q degree), 1^T A T = rho_cp/dt int T + rho_cp (u . a) |Omega| (the
  stiffness part vanishes against the constant), which fixes the sign and the
  component order of the advection term;
* on a field of z alone only u_z acts, so the constant field (0, 0, w0) and
  the vortex field with axial velocity w0 add the same advection term (both
  place u_z on d/dz).
"""
import numpy as np
import pytest

from oracle import brute, cpu
from oracle.rng import constrained_mask

TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 2.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}
GAS = {"type": "gas", "gas_rho": 2.0, "gas_cp": 1.5, "gas_k": 0.7}


@pytest.mark.parametrize("n,p,mat,faces", [
    ((2, 3, 2), 1, GAS, 0b000000),
    ((2, 2, 2), 2, GAS, 0b100101),
    ((2, 2, 1), 3, TWO, 0b000011),
    ((1, 2, 2), 4, GAS, 0b001000),
])
def test_constant_velocity_matches_dense_assembly(n, p, mat, faces):
    u = (-0.7, 0.4, 1.3)
    lo, hi = (0.0, -0.5, 0.25), (1.0, 0.5, 1.0)
    rng = np.random.default_rng(p)
    from oracle.rng import lattice_shape
    N = int(np.prod(lattice_shape(n, p)))
    H = rng.uniform(0.0, 1.0, N) if mat["type"] == "gas_liquid" else None
    pb = cpu.Problem(n, p, dt=0.05, material=mat, heaviside=H, lo=lo, hi=hi,
                     dirichlet_faces=faces, velocity="constant", velocity_value=u)
    A, Mrc, M1, mask = brute.assemble_parts(n, p, 0.05, mat, heaviside=H, velocity="constant",
                                            velocity_value=u, lo=lo, hi=hi, dirichlet_faces=faces)
    src = rng.uniform(-1.0, 1.0, N)
    ref = brute.vmult(A, mask, np.where(mask, 0.0, src))
    assert np.abs(pb.vmult(src) - ref).max() <= 1e-12 * np.abs(ref).max()
    d = brute.diagonal(A, mask)
    assert np.abs(pb.diagonal() - d).max() <= 1e-12 * np.abs(d).max()
    T, To = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    r = brute.create_rhs(A, Mrc, M1, mask, T, To)
    assert np.abs(pb.create_rhs(T, To) - r).max() <= 1e-12 * np.abs(r).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_constant_velocity_closed_form(p):
    n = (3, 2, 2)
    lo, hi = (0.0, 0.0, 0.0), (1.5, 1.0, 0.5)
    u = np.array([0.3, -1.1, 2.0])
    a = np.array([1.0, 2.0, -3.0])
    rho, cp, dt = 2.0, 3.0, 0.1
    mat = {"type": "gas", "gas_rho": rho, "gas_cp": cp, "gas_k": 0.9}
    pb = cpu.Problem(n, p, dt=dt, material=mat, lo=lo, hi=hi, dirichlet_faces=0,
                     velocity="constant", velocity_value=tuple(u))
    from oracle.rng import lattice_shape
    from oracle import tables
    n1 = lattice_shape(n, p)
    gll = tables.gll(p)
    coords = [np.concatenate([lo[d] + (c + gll[:-1]) * (hi[d] - lo[d]) / n[d]
                              for c in range(n[d])] + [[hi[d]]]) for d in range(3)]
    Z, Y, X = np.meshgrid(coords[2], coords[1], coords[0], indexing="ij")
    T = (a[0] * X + a[1] * Y + a[2] * Z).reshape(-1)
    vol = np.prod(np.subtract(hi, lo))
    centre = 0.5 * (np.array(lo) + np.array(hi))
    int_T = vol * float(a @ centre)
    expect = rho * cp / dt * int_T + rho * cp * float(u @ a) * vol
    got = np.ones(pb.N) @ pb.vmult(T)
    assert abs(got - expect) <= 1e-12 * abs(expect)


def test_constant_z_velocity_equals_vortex_on_fields_of_z():
    n, p = (2, 2, 3), 2
    mat = {"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0}
    pc = cpu.Problem(n, p, dt=0.1, material=mat, dirichlet_faces=0, velocity="constant",
                     velocity_value=(0.0, 0.0, 0.37))
    pv = cpu.Problem(n, p, dt=0.1, material=mat, dirichlet_faces=0, velocity=True, w0=0.37)
    from oracle.rng import lattice_shape
    n1 = lattice_shape(n, p)
    z = np.linspace(0.0, 1.0, n1[2])
    T = np.repeat(np.sin(2.0 * z), n1[0] * n1[1])  # a function of z only
    # subtract the velocity-free operator: the advection parts must agree
    p0 = cpu.Problem(n, p, dt=0.1, material=mat, dirichlet_faces=0)
    dc = pc.vmult(T) - p0.vmult(T)
    dv = pv.vmult(T) - p0.vmult(T)
    assert np.abs(dc - dv).max() <= 1e-13 * np.abs(dc).max()
    assert np.abs(dc).max() > 1e-3


@pytest.mark.parametrize("n,p,mat,faces", [
    ((2, 3, 2), 1, GAS, 0b000000),
    ((2, 2, 2), 2, TWO, 0b100101),
    ((2, 1, 2), 3, GAS, 0b000011),
    ((1, 2, 1), 4, TWO, 0b001000),
])
def test_nodal_velocity_matches_dense_assembly(n, p, mat, faces):
    """A nodal (DoF-vector) velocity read plain and interpolated at q
    (vel_eval, heat_diffuse_operator.cpp:975-980): the oracle against dense
    assembly with the brute-force module's own basis, for a random field whose
    three components differ, for the tangent, its diagonal and the residual."""
    lo, hi = (0.0, -0.5, 0.25), (1.0, 0.5, 1.0)
    rng = np.random.default_rng(10 + p)
    from oracle.rng import lattice_shape
    N = int(np.prod(lattice_shape(n, p)))
    H = rng.uniform(0.0, 1.0, N) if mat["type"] == "gas_liquid" else None
    v = np.concatenate([rng.uniform(-1.0, 1.0, N), 0.5 * rng.uniform(-1.0, 1.0, N),
                        2.0 + rng.uniform(-1.0, 1.0, N)])
    pb = cpu.Problem(n, p, dt=0.05, material=mat, heaviside=H, lo=lo, hi=hi,
                     dirichlet_faces=faces, velocity="nodal", velocity_field=v)
    A, Mrc, M1, mask = brute.assemble_parts(n, p, 0.05, mat, heaviside=H, velocity="nodal",
                                            velocity_field=v, lo=lo, hi=hi, dirichlet_faces=faces)
    src = rng.uniform(-1.0, 1.0, N)
    ref = brute.vmult(A, mask, np.where(mask, 0.0, src))
    assert np.abs(pb.vmult(src) - ref).max() <= 1e-12 * np.abs(ref).max()
    d = brute.diagonal(A, mask)
    assert np.abs(pb.diagonal() - d).max() <= 1e-12 * np.abs(d).max()
    T, To = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    r = brute.create_rhs(A, Mrc, M1, mask, T, To)
    assert np.abs(pb.create_rhs(T, To) - r).max() <= 1e-12 * np.abs(r).max()
    # swapping two components changes the operator (the component order is pinned)
    vs = np.concatenate([v[N:2 * N], v[:N], v[2 * N:]])
    ps = cpu.Problem(n, p, dt=0.05, material=mat, heaviside=H, lo=lo, hi=hi,
                     dirichlet_faces=faces, velocity="nodal", velocity_field=vs)
    assert np.abs(ps.vmult(src) - ref).max() > 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("p", [1, 3])
def test_nodal_constant_velocity_equals_constant_mode(p):
    """The nodal interpolant of a constant field is the constant: the nodal
    mode with constant nodal values is the constant mode to rounding."""
    n = (2, 3, 2)
    u = (0.3, -1.1, 2.0)
    pc = cpu.Problem(n, p, dt=0.1, material=GAS, velocity="constant", velocity_value=u)
    N = pc.N
    v = np.concatenate([np.full(N, c) for c in u])
    pn = cpu.Problem(n, p, dt=0.1, material=GAS, velocity="nodal", velocity_field=v)
    x = pc.random(3)
    yc = pc.vmult(x)
    assert np.abs(pn.vmult(x) - yc).max() <= 1e-13 * np.abs(yc).max()
