"""Pins of the oracle's Neumann faces: rhs_boundary_loop's Neumann term
-(q . n) integrated against the face shape functions
(source/heat/heat_diffuse_operator.cpp:735-767), for a constant flux vector q
per face of the box.

* closed form, with T = T_old = 0 so that only the faces act: on a face node
  (i, j) the residual is -(q . n) int phi_i phi_j over the face, the product of
  the two assembled 1D integrals of the Lagrange basis (computed here with
  numpy's Gauss rule and this file's own Lagrange polynomials), which fixes
  the sign, the outward normal of each face and the face metric; nodes off the
  Neumann faces get nothing, and the sum is -sum_f (q_f . n_f) |face f|;
* linearity: the residual with Neumann faces minus the one without them is the
  Neumann part alone, for any T, T_old, heat source and material;
* a face that is both Neumann and Robin is refused (heat_diffuse_operator.cpp:
  744-747)."""
import numpy as np
import pytest

from oracle import cpu


def _lagrange_integrals(n, p, lo, hi):
    """int phi_I over [lo, hi] of the assembled FE_Q(p) basis on n cells."""
    from numpy.polynomial import legendre as Lg
    from meltpooldg_amd.inputs import gll_points
    nodes = gll_points(p)
    xg, wg = Lg.leggauss(p + 1)
    xg, wg = 0.5 * (xg + 1.0), 0.5 * wg
    loc = np.array([np.sum(wg * np.prod([(xg - nodes[j]) / (nodes[i] - nodes[j])
                                         for j in range(p + 1) if j != i], axis=0))
                    for i in range(p + 1)])
    h = (hi - lo) / n
    out = np.zeros(n * p + 1)
    for c in range(n):
        out[c * p: c * p + p + 1] += loc * h
    return out


@pytest.mark.parametrize("n,p", [((2, 3, 2), 1), ((3, 2, 2), 2), ((2, 2, 3), 3), ((1, 2, 1), 4)])
def test_neumann_closed_form(n, p):
    lo, hi = (0.0, -0.5, 0.25), (1.5, 0.5, 1.0)
    q = {0: (2.0, 0.3, -0.1), 3: (0.5, -1.5, 0.7), 5: (0.2, 0.4, 3.0)}
    pb = cpu.Problem(n, p, dt=0.1, lo=lo, hi=hi, dirichlet_faces=0, neumann=q)
    N = pb.N
    r = pb.create_rhs(np.zeros(N), np.zeros(N))
    n1 = [c * p + 1 for c in n]
    R = r.reshape(n1[2], n1[1], n1[0])
    W = [_lagrange_integrals(n[d], p, lo[d], hi[d]) for d in range(3)]
    expect = np.zeros_like(R)
    total = 0.0
    for f, qf in q.items():
        d, side = f // 2, f % 2
        qn = (1.0 if side else -1.0) * qf[d]
        e = [k for k in range(3) if k != d]
        area = (hi[e[0]] - lo[e[0]]) * (hi[e[1]] - lo[e[1]])
        total += -qn * area
        face = -qn * np.einsum("i,j->ij", W[e[1]], W[e[0]])  # [e2, e1] = (slow, fast)
        sl = [slice(None)] * 3
        sl[2 - d] = n1[d] - 1 if side else 0
        expect[tuple(sl)] += face
    assert np.abs(R - expect).max() <= 1e-13 * np.abs(expect).max()
    assert abs(r.sum() - total) <= 1e-12 * abs(total)


def test_neumann_part_is_additive():
    n, p = (2, 3, 2), 2
    mat = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 2.0, "gas_k": 1.0,
           "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}
    rng = np.random.default_rng(3)
    base = cpu.Problem(n, p, dt=0.05, dirichlet_faces=0b000011)
    N = base.N
    H = rng.uniform(0, 1, N)
    kw = dict(dt=0.05, material=mat, heaviside=H, dirichlet_faces=0b000011, velocity="constant",
              velocity_value=(0.3, -0.2, 0.5))
    q = {2: (0.0, 1.25, 0.0), 4: (0.7, 0.0, -2.0)}
    with_n = cpu.Problem(n, p, neumann=q, **kw)
    without = cpu.Problem(n, p, **kw)
    only = cpu.Problem(n, p, dt=0.05, dirichlet_faces=0b000011, neumann=q)
    T, To, f = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    a = with_n.create_rhs(T, To, f)
    b = without.create_rhs(T, To, f) + only.create_rhs(np.zeros(N), np.zeros(N))
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    # the tangent does not see the Neumann faces
    x = rng.uniform(-1, 1, N)
    assert np.array_equal(with_n.vmult(x), without.vmult(x))


def test_neumann_and_robin_on_one_face_refused():
    with pytest.raises(AssertionError, match="invalid oracle problem"):
        cpu.Problem((2, 2, 2), 1, dirichlet_faces=0, neumann={1: (1.0, 0.0, 0.0)},
                    robin=dict(convection_faces=0b10, convection_coefficient=2.0))
    # on different faces both act
    cpu.Problem((2, 2, 2), 1, dirichlet_faces=0, neumann={0: (1.0, 0.0, 0.0)},
                robin=dict(convection_faces=0b10, convection_coefficient=2.0))
