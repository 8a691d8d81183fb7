"""Independent numpy construction of the 1D FE_Q(p) / QGauss(p+1) tables.

Test infrastructure. Uses numpy's own Gauss-Legendre rule and Legendre
polynomial roots, and the barycentric Lagrange formula, so that it shares no
code with the C++ oracle or the product's table builder.

FE_Q(p) places its support points at the Gauss-Lobatto points
(source/utilities/fe_util.cpp:16-58 selects FE_Q(p) + QGauss(p+1);
source/core/finite_element_data.cpp:61-65 sets n_q = p + 1).
"""
import numpy as np
from numpy.polynomial import legendre as L


def gauss(n):
    x, w = L.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def gll(p):
    inner = np.sort(L.Legendre.basis(p).deriv().roots().real) if p > 1 else np.array([])
    return np.concatenate([[0.0], 0.5 * (inner + 1.0), [1.0]])


def _bary_weights(nodes):
    n = len(nodes)
    w = np.ones(n)
    for j in range(n):
        for k in range(n):
            if k != j:
                w[j] /= nodes[j] - nodes[k]
    return w


def lagrange_values(nodes, x):
    """V[q, i] = l_i(x_q) by the barycentric formula."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    wb = _bary_weights(nodes)
    V = np.zeros((len(x), len(nodes)))
    for q, xq in enumerate(x):
        diff = xq - nodes
        hit = np.where(np.abs(diff) < 1e-15)[0]
        if hit.size:
            V[q, hit[0]] = 1.0
            continue
        t = wb / diff
        V[q] = t / t.sum()
    return V


def lagrange_derivs(nodes, x):
    """Dv[q, i] = l_i'(x_q), via the differentiation matrix on the nodes
    (D_jk = (w_k/w_j)/(x_j-x_k)) pushed through interpolation."""
    n = len(nodes)
    wb = _bary_weights(nodes)
    Dn = np.zeros((n, n))
    for j in range(n):
        for k in range(n):
            if j != k:
                Dn[j, k] = (wb[k] / wb[j]) / (nodes[j] - nodes[k])
        Dn[j, j] = -Dn[j].sum()
    # l_i' is a polynomial of degree p-1 <= p: interpolate its nodal values
    return lagrange_values(nodes, x) @ Dn


def tables(p):
    nodes = gll(p)
    xq, wq = gauss(p + 1)
    S = lagrange_values(nodes, xq)
    D = lagrange_derivs(nodes, xq)
    return nodes, xq, wq, S, D
