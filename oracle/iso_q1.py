"""FE_Q_iso_Q1 level set -> FE_Q(p) heaviside interpolation (test infrastructure).

Restates the reference's level-set-to-heat interpolation for a level set on
FE_Q_iso_Q1(p): DoFTools::create_dof_interpolation_matrix
(source/utilities/dof_tools.cpp:20-49) evaluates every iso shape function at
every support point of the heat element, and interpolate_dof_values
(dof_tools.cpp:79-103) applies that matrix to a cell's level-set values before
the heaviside is evaluated at the quadrature points
(source/heat/heat_diffuse_operator.cpp:834-838, 982-993).

FE_Q_iso_Q1(p) is a tensor product of 1D piecewise-linear hats on the p
equidistant sub-intervals of a cell; FE_Q(p)'s support points are the
Gauss-Lobatto points. Hence the cell matrix is M1 x M1 x M1 with
M1[a][b] = hat_b(x_a) = max(0, 1 - |p x_a - b|), applied here per lattice axis
(rows a = 0 and a = p are unit vectors, so cells sharing a face agree on it).

Only ``tests/`` import this module.
"""
import numpy as np

from . import tables as T


def hat_matrix(p):
    """M1[a][b]: iso-Q1 hat b (node b/p) at the Gauss-Lobatto point a of [0, 1]."""
    x = T.gll(p)
    b = np.arange(p + 1)
    return np.maximum(0.0, 1.0 - np.abs(p * x[:, None] - b[None, :]))


def _axis_matrix(n, p):
    """Lattice-wide 1D map: n cells of degree p, (n p + 1) x (n p + 1)."""
    m1 = hat_matrix(p)
    A = np.zeros((n * p + 1, n * p + 1))
    for c in range(n):
        A[c * p:c * p + p + 1, c * p:c * p + p + 1] = m1
    return A


def iso_q1_to_gll(n_cells, p, h_iso):
    """GLL nodal values of the FE_Q(p) interpolant of an FE_Q_iso_Q1(p) field
    given on the equidistant lattice (x fastest)."""
    n = (n_cells,) * 3 if np.isscalar(n_cells) else tuple(n_cells)
    shape = tuple(n[k] * p + 1 for k in (2, 1, 0))  # (z, y, x)
    v = np.asarray(h_iso, dtype=np.float64).reshape(shape)
    v = np.einsum("ia,zya->zyi", _axis_matrix(n[0], p), v)
    v = np.einsum("jb,zbx->zjx", _axis_matrix(n[1], p), v)
    v = np.einsum("kc,cyx->kyx", _axis_matrix(n[2], p), v)
    return v.reshape(-1)
