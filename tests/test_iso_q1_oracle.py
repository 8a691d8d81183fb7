"""Pins of the FE_Q_iso_Q1 -> FE_Q(p) heaviside interpolation of the oracle
(oracle/iso_q1.py, restating source/utilities/dof_tools.cpp:20-49, 79-103 as
used by source/heat/heat_diffuse_operator.cpp:834-838, 982-993).

* A trilinear field is reproduced exactly: FE_Q_iso_Q1(p) holds every
  function that is trilinear on each sub-cell, so its interpolant at the
  Gauss-Lobatto points is the field's value there.
* Degrees 1 and 2: the equidistant and the Gauss-Lobatto points coincide, so
  the interpolation is the identity.
* A random field against the global sub-grid hat functions evaluated at the
  global support-point coordinates (no cells involved).
"""
import numpy as np
import pytest

from oracle import iso_q1
from oracle.rng import random_vector
from tests.cases import support_points


def _equidistant_points(n, p):
    axes = [np.linspace(0.0, 1.0, n[d] * p + 1) for d in range(3)]
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], axis=1)


def _trilinear(P):
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    return 0.3 + 1.1 * x - 0.7 * y + 0.4 * z + 0.9 * x * y - 0.5 * y * z + 0.6 * x * z + 1.3 * x * y * z


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_trilinear_field_is_reproduced(p):
    n = (3, 2, 4)
    h_iso = _trilinear(_equidistant_points(n, p))
    got = iso_q1.iso_q1_to_gll(n, p, h_iso)
    want = _trilinear(support_points(n, p, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
    assert np.abs(got - want).max() < 1e-14


def test_quadratic_field_is_not_reproduced():
    # the iso space is piecewise linear: x^2 differs from its interpolant
    n, p = (2, 2, 2), 4
    f = lambda P: P[:, 0] ** 2
    got = iso_q1.iso_q1_to_gll(n, p, f(_equidistant_points(n, p)))
    want = f(support_points(n, p, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
    assert np.abs(got - want).max() > 1e-4


@pytest.mark.parametrize("p", [1, 2])
def test_identity_for_degrees_one_and_two(p):
    n = (3, 2, 4)
    h = random_vector(n, p, 5, dirichlet_faces=0)
    assert np.array_equal(iso_q1.iso_q1_to_gll(n, p, h), h)


@pytest.mark.parametrize("p", [3, 4, 5, 6])
def test_against_global_hat_functions(p):
    n = (3, 2, 2)
    h = random_vector(n, p, 7, dirichlet_faces=0)
    S = support_points(n, p, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    phi = []
    for d in range(3):
        nodes = np.arange(n[d] * p + 1)
        hs = 1.0 / (n[d] * p)  # sub-grid spacing
        phi.append(np.maximum(0.0, 1.0 - np.abs(S[:, d][:, None] / hs - nodes[None, :])))
    H = h.reshape(n[2] * p + 1, n[1] * p + 1, n[0] * p + 1)
    want = np.einsum("ni,nj,nk,kji->n", phi[0], phi[1], phi[2], H)
    assert np.abs(iso_q1.iso_q1_to_gll(n, p, h) - want).max() < 1e-14
