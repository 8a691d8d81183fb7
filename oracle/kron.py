"""Kronecker-product closed form of the discrete heat operator (test infrastructure).

On a box mesh of affine hex cells with FE_Q(p) x QGauss(p+1), every term of
the tangent (source/heat/heat_diffuse_operator.cpp:1029-1042)

    J = (rho_cp/dt) M + K_k + rho_cp C_u

is a sum of Kronecker products of assembled 1D matrices whenever the
coefficient fields are separable at the tensor quadrature points:
  * constant rho_cp, k (material ``gas``);
  * k(x) = k0 + a s(x) s(y) s(z), s = sin(2 pi t) (separable-sine conductivity);
  * the vortex velocity u = (-sin^2(pi x) sin(2 pi y), sin(2 pi x) sin^2(pi y), w0).
The 1D matrices use the same Gauss rule as the operator, so the closed form is
exact for the discrete operator (not an approximation). Independent of the
C++ oracle: own 1D tables (oracle/tables.py), own assembly, own application
by mode products.
"""
import numpy as np

from . import tables as T
from .rng import constrained_mask, lattice_shape


def assemble_1d(n, p, lo, hi, f=None, kind="mass"):
    """Assembled 1D matrix on n cells of [lo, hi] with weight f(x) at Gauss points.

    kind: mass  -> int f phi_i phi_j
          stiff -> int f phi_i' phi_j'
          conv  -> int f phi_i phi_j'      (row = test function i)
    """
    nodes, xq, wq, S, D = T.tables(p)
    h = (hi - lo) / n
    N1 = n * p + 1
    A = np.zeros((N1, N1))
    for c in range(n):
        xg = lo + (c + xq) * h
        fw = wq * h * (np.ones_like(xg) if f is None else f(xg))
        if kind == "mass":
            loc = S.T @ (fw[:, None] * S)
        elif kind == "stiff":
            loc = (D / h).T @ (fw[:, None] * (D / h))
        elif kind == "conv":
            loc = S.T @ (fw[:, None] * (D / h))
        else:
            raise ValueError(kind)
        sl = slice(c * p, c * p + p + 1)
        A[sl, sl] += loc
    return A


def _apply3(Az, Ay, Ax, u3):
    # (Az (x) Ay (x) Ax) u with u3 indexed [z, y, x]
    return np.einsum("ai,bj,ck,ijk->abc", Az, Ay, Ax, u3, optimize=True)


class KronOperator:
    def __init__(self, n, p, dt=1e-3, rho_cp=1.0, k=1.0, k_amplitude=0.0, velocity=False,
                 w0=0.2, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), dirichlet_faces=0x3F):
        self.n = (n, n, n) if np.isscalar(n) else tuple(n)
        self.p = p
        self.shape = lattice_shape(self.n, p)[::-1]  # (z, y, x)
        self.mask = constrained_mask(self.n, p, dirichlet_faces)
        c = rho_cp / dt
        s = lambda x: np.sin(2 * np.pi * x)
        a = lambda x: np.sin(np.pi * x) ** 2
        M = [assemble_1d(self.n[d], p, lo[d], hi[d], None, "mass") for d in range(3)]
        K = [assemble_1d(self.n[d], p, lo[d], hi[d], None, "stiff") for d in range(3)]
        self.terms = []  # list of (coef, Az, Ay, Ax)
        self.terms.append((c, M[2], M[1], M[0]))
        self.terms += [(k, M[2], M[1], K[0]), (k, M[2], K[1], M[0]), (k, K[2], M[1], M[0])]
        if k_amplitude:
            Ms = [assemble_1d(self.n[d], p, lo[d], hi[d], s, "mass") for d in range(3)]
            Ks = [assemble_1d(self.n[d], p, lo[d], hi[d], s, "stiff") for d in range(3)]
            self.terms += [(k_amplitude, Ms[2], Ms[1], Ks[0]), (k_amplitude, Ms[2], Ks[1], Ms[0]),
                           (k_amplitude, Ks[2], Ms[1], Ms[0])]
        if velocity:
            # vx = -a(x) s(y), vy = s(x) a(y), vz = w0
            Ca_x = assemble_1d(self.n[0], p, lo[0], hi[0], a, "conv")
            Ms_y = assemble_1d(self.n[1], p, lo[1], hi[1], s, "mass")
            Ca_y = assemble_1d(self.n[1], p, lo[1], hi[1], a, "conv")
            Ms_x = assemble_1d(self.n[0], p, lo[0], hi[0], s, "mass")
            C_z = assemble_1d(self.n[2], p, lo[2], hi[2], None, "conv")
            self.terms += [(-rho_cp, M[2], Ms_y, Ca_x), (rho_cp, M[2], Ca_y, Ms_x),
                           (rho_cp * w0, C_z, M[1], M[0])]

    def vmult(self, src):
        u = np.where(self.mask, 0.0, src).reshape(self.shape)
        out = np.zeros(self.shape)
        for coef, Az, Ay, Ax in self.terms:
            out += coef * _apply3(Az, Ay, Ax, u)
        out = out.reshape(-1)
        out[self.mask] = 0.0
        return out

    def diagonal(self):
        d = np.zeros(self.shape)
        for coef, Az, Ay, Ax in self.terms:
            d += coef * np.einsum("a,b,c->abc", np.diag(Az), np.diag(Ay), np.diag(Ax))
        d = d.reshape(-1)
        d[self.mask] = 0.0
        return d

    def inverse_diagonal(self):
        d = self.diagonal()
        return np.where(np.abs(d) > 1e-10, 1.0 / np.where(d == 0, 1.0, d), 1.0)
