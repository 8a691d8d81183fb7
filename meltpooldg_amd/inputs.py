"""Synthetic inputs of the benchmark workloads (host numpy arrays in lattice order).

* ``random_vector``: uniform[-1, 1] on the free DoFs and 0 on the homogeneous
  Dirichlet ones, value(seed, I) = 2 (splitmix64(splitmix64(seed) ^ I) >> 11)
  2^-53 - 1 for the lexicographic lattice index I = ix + N1x (iy + N1y iz)
  (SplitMix64: Steele, Lea, Flood 2014), so a vector does not depend on any
  DoF numbering or thread order.
* ``batch_members``: the heterogeneous batch workload, cubes of n^3 cells
  with n ~ U{8..48}, degree ~ U{1..4}, conductivity 10^U(-1, 1) (numpy
  default_rng of the master seed), RHS seed 1000 + member.
* ``sphere_heaviside``: the nodal smoothed heaviside of a spherical level set,
  CharacteristicFunctions::smoothed_heaviside
  (include/meltpooldg/utilities/characteristic_functions.hpp:17-28) of
  radius - |x - centre| with half-width eps, at the FE_Q(p) support points.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix(z):
    with np.errstate(over="ignore"):
        z = z + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def lattice_shape(n, degree):
    """Lattice nodes per direction (x, y, z) of n cells (int or triple) of FE_Q(degree)."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    return tuple(int(c) * degree + 1 for c in n)


def dirichlet_mask(n, degree, dirichlet_faces=0x3F):
    """Flat boolean mask (x fastest) of the lattice nodes on the marked faces
    (bit 2d: face x_d = lo, bit 2d + 1: x_d = hi)."""
    n1 = lattice_shape(n, degree)
    m = np.zeros(n1[::-1], dtype=bool)
    for d in range(3):
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[2 - d] = 0
        sl_hi[2 - d] = n1[d] - 1
        if (dirichlet_faces >> (2 * d)) & 1:
            m[tuple(sl_lo)] = True
        if (dirichlet_faces >> (2 * d + 1)) & 1:
            m[tuple(sl_hi)] = True
    return m.reshape(-1)


def random_vector(n, degree, seed, dirichlet_faces=0x3F, chunk=1 << 24):
    n1 = lattice_shape(n, degree)
    N = n1[0] * n1[1] * n1[2]
    key = _splitmix(np.uint64(seed))
    out = np.empty(N, dtype=np.float64)
    for s in range(0, N, chunk):
        idx = np.arange(s, min(N, s + chunk), dtype=np.uint64)
        r = _splitmix(key ^ idx)
        out[s: s + idx.size] = 2.0 * ((r >> np.uint64(11)).astype(np.float64) * 2.0**-53) - 1.0
    out[dirichlet_mask(n, degree, dirichlet_faces)] = 0.0
    return out


def gll_points(degree):
    """Gauss-Lobatto support points of FE_Q(degree) on [0, 1]: 0, the roots of
    P_degree' (Newton-polished), 1."""
    from numpy.polynomial import legendre as Lg
    dP = Lg.Legendre.basis(degree).deriv()
    ddP = dP.deriv()
    z = np.sort(dP.roots().real) if degree > 1 else np.array([])
    for _ in range(3):
        z = z - dP(z) / ddP(z) if z.size else z
    return np.concatenate([[0.0], 0.5 * (z + 1.0), [1.0]])


def sphere_heaviside(n, degree, radius=0.3, centre=0.5, width_cells=2.0):
    """Smoothed heaviside of radius - |x - centre| on the unit cube with n^3
    cells, eps = width_cells / 2 cells: 1 for d > eps, 0 for d <= -eps,
    (d + eps) / (2 eps) + sin(pi d / eps) / (2 pi) in between."""
    gll = gll_points(degree)
    coords = np.concatenate([(c + gll[:-1]) / n for c in range(n)] + [[1.0]])
    eps = 0.5 * width_cells / n
    Z, Y, X = np.meshgrid(coords, coords, coords, indexing="ij")
    d = radius - np.sqrt((X - centre) ** 2 + (Y - centre) ** 2 + (Z - centre) ** 2)
    H = np.where(d > eps, 1.0, np.where(d <= -eps, 0.0,
                 (d + eps) / (2 * eps) + np.sin(np.pi * d / eps) / (2 * np.pi)))
    return H.reshape(-1)


def batch_members(count=64, seed=0, n_range=(8, 48), degrees=(1, 4)):
    """[{"n_cells": n, "degree": p, "k": k}, ...] drawn in member order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        p = int(rng.integers(degrees[0], degrees[1] + 1))
        k = float(10.0 ** rng.uniform(-1.0, 1.0))
        out.append({"n_cells": n, "degree": p, "k": k})
    return out
