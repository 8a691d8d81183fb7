"""SplitMix64 seeded input vectors (test infrastructure).

value(seed, I) = 2 * (splitmix64(splitmix64(seed) ^ I) >> 11) * 2^-53 - 1 on free
lattice DoFs, 0 on constrained ones. Identical to ``oracle_random_vector`` in
heat_oracle.cpp (Steele, Lea, Flood 2014 mixer).
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    with np.errstate(over="ignore"):
        z = z + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def splitmix64(x):
    return _mix(np.asarray(x, dtype=np.uint64))


def lattice_shape(n, degree):
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    return tuple(int(c) * degree + 1 for c in n)


def constrained_mask(n, degree, dirichlet_faces=0x3F):
    """Boolean lattice mask (flat, x fastest) of homogeneous Dirichlet DoFs."""
    n1 = lattice_shape(n, degree)
    iz, iy, ix = np.meshgrid(np.arange(n1[2]), np.arange(n1[1]), np.arange(n1[0]), indexing="ij")
    m = np.zeros(iz.shape, dtype=bool)
    for d, c in enumerate((ix, iy, iz)):
        if (dirichlet_faces >> (2 * d)) & 1:
            m |= c == 0
        if (dirichlet_faces >> (2 * d + 1)) & 1:
            m |= c == n1[d] - 1
    return m.reshape(-1)


def random_vector(n, degree, seed, dirichlet_faces=0x3F, chunk=1 << 24):
    n1 = lattice_shape(n, degree)
    N = n1[0] * n1[1] * n1[2]
    key = splitmix64(np.uint64(seed))
    out = np.empty(N, dtype=np.float64)
    for s in range(0, N, chunk):
        idx = np.arange(s, min(N, s + chunk), dtype=np.uint64)
        r = _mix(key ^ idx)
        out[s : s + idx.size] = 2.0 * ((r >> np.uint64(11)).astype(np.float64) * 2.0**-53) - 1.0
    out[constrained_mask(n, degree, dirichlet_faces)] = 0.0
    return out
