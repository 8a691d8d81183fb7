"""Dense brute-force assembly of the heat tangent on tiny meshes (test infrastructure).

Builds every cell matrix explicitly from 3D basis values/gradients at every
quadrature point (no sum factorisation), assembles the global dense matrix and
restricts it to the free DoFs. Implements the quadrature-point physics of
tangent_local_cell_operation (source/heat/heat_diffuse_operator.cpp:1029-1042)
with Material gas / gas_liquid (include/meltpooldg/core/material.templates.hpp:
223-295, 515-524) and the nodal heaviside read plain and interpolated at q
(:982-1002). The 1D basis comes from a monomial Vandermonde solve, a third
independent construction besides oracle/tables.py and the C++ oracle.
"""
import numpy as np
from numpy.polynomial import legendre as L

from .rng import constrained_mask, lattice_shape


def _basis_1d(p):
    inner = np.sort(L.Legendre.basis(p).deriv().roots().real) if p > 1 else np.array([])
    nodes = np.concatenate([[0.0], 0.5 * (inner + 1.0), [1.0]])
    V = np.vander(nodes, p + 1, increasing=True)  # V[i, m] = x_i^m
    C = np.linalg.solve(V, np.eye(p + 1))  # column i: monomial coefficients of l_i
    gx, gw = L.leggauss(p + 1)
    xq, wq = 0.5 * (gx + 1.0), 0.5 * gw
    P = np.vander(xq, p + 1, increasing=True)
    dP = np.zeros_like(P)
    for m in range(1, p + 1):
        dP[:, m] = m * xq ** (m - 1)
    return xq, wq, P @ C, dP @ C


def material(H, mat):
    """(rho_cp, k) at a quadrature point for material dict ``mat``."""
    if mat.get("type", "gas") == "gas":
        return mat["gas_rho"] * mat["gas_cp"], mat["gas_k"] + 0.0 * H
    w = (H > 0.5).astype(float) if mat.get("sharp") else H
    k = (1.0 - w) * mat["gas_k"] + w * mat["liq_k"]
    if mat.get("volume_specific"):
        rc = (1.0 - w) * (mat["gas_rho"] * mat["gas_cp"]) + w * (mat["liq_rho"] * mat["liq_cp"])
    else:
        cp = (1.0 - w) * mat["gas_cp"] + w * mat["liq_cp"]
        if mat.get("consistent_with_evaporation"):
            rho = 1.0 / ((1.0 - H) / mat["gas_rho"] + H / mat["liq_rho"])
        else:
            rho = (1.0 - w) * mat["gas_rho"] + w * mat["liq_rho"]
        rc = rho * cp
    return rc, k


def assemble(n, p, dt, mat, heaviside=None, k_amplitude=0.0, velocity=False, w0=0.2,
             lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), dirichlet_faces=0x3F, velocity_value=None,
             velocity_field=None):
    """Return (A_full, mask): dense lattice matrix without constraints and the
    constrained-DoF mask. The constrained operator is A_full restricted to free
    rows/columns, zero elsewhere."""
    A, _, _, mask = assemble_parts(n, p, dt, mat, heaviside, k_amplitude, velocity, w0, lo, hi,
                                   dirichlet_faces, velocity_value, velocity_field)
    return A, mask


def assemble_parts(n, p, dt, mat, heaviside=None, k_amplitude=0.0, velocity=False, w0=0.2,
                   lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), dirichlet_faces=0x3F,
                   velocity_value=None, velocity_field=None):
    """(A_full, M_rc, M_1, mask): the full tangent, its mass part rho_cp/dt M and
    the unit mass matrix M_1 (integral of phi_i phi_j), all without constraints.
    velocity: False, True / "vortex", "constant" with velocity_value, or
    "nodal" with velocity_field (3 N nodal values, component-blocked, evaluated
    at q with this module's own basis)."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    n1 = lattice_shape(n, p)
    N = n1[0] * n1[1] * n1[2]
    xq, wq, S, D = _basis_1d(p)
    nq = p + 1
    h = [(hi[d] - lo[d]) / n[d] for d in range(3)]
    # 3D basis at q: index q = (qz, qy, qx), i = (c, b, a)
    Phi = np.einsum("zc,yb,xa->zyxcba", S, S, S).reshape(nq**3, nq**3)
    Gx = np.einsum("zc,yb,xa->zyxcba", S, S, D).reshape(nq**3, nq**3) / h[0]
    Gy = np.einsum("zc,yb,xa->zyxcba", S, D, S).reshape(nq**3, nq**3) / h[1]
    Gz = np.einsum("zc,yb,xa->zyxcba", D, S, S).reshape(nq**3, nq**3) / h[2]
    W = np.einsum("z,y,x->zyx", wq, wq, wq).reshape(-1) * h[0] * h[1] * h[2]
    A = np.zeros((N, N))
    Mrc = np.zeros((N, N))
    M1 = np.zeros((N, N))
    loc = np.arange(nq)
    for cz in range(n[2]):
        for cy in range(n[1]):
            for cx in range(n[0]):
                ix = cx * p + loc
                iy = cy * p + loc
                iz = cz * p + loc
                dofs = (ix[None, None, :] + n1[0] * (iy[None, :, None] + n1[1] * iz[:, None, None])).reshape(-1)
                X = lo[0] + (cx + xq) * h[0]
                Y = lo[1] + (cy + xq) * h[1]
                Z = lo[2] + (cz + xq) * h[2]
                Zq, Yq, Xq = [a.reshape(-1) for a in np.meshgrid(Z, Y, X, indexing="ij")]
                Hq = Phi @ heaviside[dofs] if heaviside is not None else np.zeros(nq**3)
                rc, k = material(Hq, mat)
                rc = rc * np.ones(nq**3)
                k = k * np.ones(nq**3)
                if k_amplitude:
                    k = k + k_amplitude * np.sin(2 * np.pi * Xq) * np.sin(2 * np.pi * Yq) * np.sin(2 * np.pi * Zq)
                Mloc = Phi.T @ ((W * rc / dt)[:, None] * Phi)
                Mrc[np.ix_(dofs, dofs)] += Mloc
                M1[np.ix_(dofs, dofs)] += Phi.T @ (W[:, None] * Phi)
                Aloc = Mloc.copy()
                for G in (Gx, Gy, Gz):
                    Aloc += G.T @ ((W * k)[:, None] * G)
                if velocity == "constant":
                    vx, vy, vz = (np.full_like(Xq, c) for c in velocity_value)
                elif velocity == "nodal":
                    vx, vy, vz = (Phi @ velocity_field[c * N + dofs] for c in range(3))
                elif velocity:
                    vx = -np.sin(np.pi * Xq) ** 2 * np.sin(2 * np.pi * Yq)
                    vy = np.sin(2 * np.pi * Xq) * np.sin(np.pi * Yq) ** 2
                    vz = w0 * np.ones_like(Xq)
                if velocity:
                    adv = vx[:, None] * Gx + vy[:, None] * Gy + vz[:, None] * Gz
                    Aloc += Phi.T @ ((W * rc)[:, None] * adv)
                A[np.ix_(dofs, dofs)] += Aloc
    mask = constrained_mask(n, p, dirichlet_faces)
    return A, Mrc, M1, mask


def constrained(A, mask):
    Ac = A.copy()
    Ac[mask, :] = 0.0
    Ac[:, mask] = 0.0
    return Ac


def vmult(A, mask, src):
    return constrained(A, mask) @ src


def create_rhs(A, Mrc, M1, mask, T, T_old, f=None, dst=None):
    """dst + (-(A_full T - M_rc T_old - M_1 f)) on the free rows; T, T_old, f are
    used with their constrained entries (read plain), constrained rows of dst
    are left as they are."""
    r = -(A @ T - Mrc @ T_old - (M1 @ f if f is not None else 0.0))
    out = np.zeros(A.shape[0]) if dst is None else np.array(dst, dtype=float)
    out[~mask] += r[~mask]
    return out


def diagonal(A, mask):
    d = np.diag(A).copy()
    d[mask] = 0.0
    return d


def pcg_dense(Ac, b, inv_diag, rel_tol, abs_tol=1e-20, max_it=10000):
    """Textbook Jacobi-PCG (Saad, Iterative Methods, Alg. 9.1) with the
    ReductionControl rule ||r_k|| <= max(abs_tol, rel_tol ||r_0||), x0 = 0."""
    x = np.zeros_like(b)
    r = b.copy()
    r0 = np.linalg.norm(r)
    tol = max(abs_tol, rel_tol * r0)
    if r0 <= tol:
        return x, 0
    z = inv_diag * r
    p = z.copy()
    rz = r @ z
    for it in range(1, max_it + 1):
        q = Ac @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if np.linalg.norm(r) <= tol:
            return x, it
        z = inv_diag * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_it


STEFAN_BOLTZMANN = 5.670374419e-8


def _face_cells(n, p, f):
    """Yield (dofs, (e1, e2)) of every face cell on box face f = 2d + side; dofs
    in the face-local order a1 + (p+1) a2 over the tangential directions e1 < e2."""
    n1 = lattice_shape(n, p)
    d, side = f >> 1, f & 1
    e1, e2 = [e for e in range(3) if e != d]
    loc = np.arange(p + 1)
    for c2 in range(n[e2]):
        for c1 in range(n[e1]):
            I = [None, None, None]
            I[d] = np.full((p + 1, p + 1), (n1[d] - 1) if side else 0)
            I[e1] = np.broadcast_to(c1 * p + loc[None, :], (p + 1, p + 1))
            I[e2] = np.broadcast_to(c2 * p + loc[:, None], (p + 1, p + 1))
            yield (I[0] + n1[0] * (I[1] + n1[1] * I[2])).reshape(-1), (e1, e2)


def robin_face_parts(n, p, robin, T_tangent, T_rhs, mask, lo=(0.0, 0.0, 0.0),
                     hi=(1.0, 1.0, 1.0)):
    """(A_face, r_face): the dense face tangent linearised at T_tangent (read with
    the constrained entries as 0) and the face residual at T_rhs (read plain),
    both without constraints, from explicit 2D face basis matrices on each face
    cell (convection alpha (T - T_inf), radiation eps sigma (T^4 - T_inf^4))."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    n1 = lattice_shape(n, p)
    N = n1[0] * n1[1] * n1[2]
    h = [(hi[d] - lo[d]) / n[d] for d in range(3)]
    _, wq, S, _ = _basis_1d(p)
    Phi = np.einsum("yb,xa->yxba", S, S).reshape((p + 1) ** 2, (p + 1) ** 2)
    W = np.einsum("y,x->yx", wq, wq).reshape(-1)
    Tt = np.where(mask, 0.0, T_tangent) if T_tangent is not None else np.zeros(N)
    A = np.zeros((N, N))
    r = np.zeros(N)
    alpha = robin.get("convection_coefficient", 0.0)
    es = robin.get("emissivity", 0.0) * STEFAN_BOLTZMANN
    for f in range(6):
        conv = (robin.get("convection_faces", 0) >> f) & 1
        rad = (robin.get("radiation_faces", 0) >> f) & 1
        if not (conv or rad):
            continue
        for dofs, (e1, e2) in _face_cells(n, p, f):
            JxW = W * h[e1] * h[e2]
            c = np.zeros_like(W)
            val = np.zeros_like(W)
            if conv:
                c += alpha
                val -= alpha * (Phi @ T_rhs[dofs] - robin["convection_t_inf"])
            if rad:
                c += 4.0 * es * (Phi @ Tt[dofs]) ** 3
                val -= es * ((Phi @ T_rhs[dofs]) ** 4 - robin["radiation_t_inf"] ** 4)
            A[np.ix_(dofs, dofs)] += Phi.T @ ((JxW * c)[:, None] * Phi)
            r[dofs] += Phi.T @ (JxW * val)
    return A, r


def apparent_capacity_t(T, mat):
    """(c, dc/dT) of the apparent capacity at temperatures T (own numpy statement
    of source/heat/apparent_capacity.cpp): the latent heat L over the mushy zone
    (T_s, T_l), zero outside it, as a function of the temperature directly:
    poly4_bell  L (2/dT) 15/16 (1 - xi^2)^2 with xi = (2 T - T_s - T_l) / dT,
    constant    L / dT,
    qlq         two zero-slope quadratic-linear-quadratic ramps, 0 -> 2 L / dT
                on (T_s, T_m) and back on (T_m, T_l), transition width dT / 6,
                built here from the ramp's integral form."""
    T = np.asarray(T, dtype=float)
    L = float(mat.get("latent_heat", 0.0))
    zero = np.zeros_like(T)
    if L <= 0.0:
        return zero, zero
    Ts, Tl = mat["solidus"], mat["liquidus"]
    dT = Tl - Ts
    inside = (T > Ts) & (T < Tl)
    kind = mat.get("apparent_capacity", "qlq")
    if kind == "constant":
        return np.where(inside, L / dT, 0.0), zero
    if kind == "poly4_bell":
        xi = (2.0 * T - Ts - Tl) / dT
        c = L * (2.0 / dT) * 15.0 / 16.0 * (1.0 - xi ** 2) ** 2
        dc = L * (2.0 / dT) ** 2 * 15.0 / 4.0 * (xi ** 3 - xi)
        return np.where(inside, c, 0.0), np.where(inside, dc, 0.0)
    # qlq: slope profile of a ramp rising by `rise` over (a, b): linear up over
    # width d, flat, linear down over d; the ramp is its integral
    Tm, peak, d = 0.5 * (Ts + Tl), 2.0 * L / dT, dT / 6.0

    def ramp(x, a, b, y0, rise):
        s = rise / (b - a - d)
        u = np.clip(x - a, 0.0, d)                 # rising part of the slope
        w = np.clip(x - (b - d), 0.0, d)           # falling part
        mid = np.clip(x - a - d, 0.0, b - a - 2 * d)
        val = y0 + s * (u * u / (2 * d) + mid + (w - w * w / (2 * d)))
        slope = np.where(x < a + d, s * (x - a) / d, np.where(x > b - d, s * (b - x) / d, s))
        return val, slope

    up_v, up_s = ramp(T, Ts, Tm, 0.0, peak)
    dn_v, dn_s = ramp(T, Tm, Tl, peak, -peak)
    c = np.where(T <= Tm, up_v, dn_v)
    dc = np.where(T <= Tm, up_s, dn_s)
    return np.where(inside, c, 0.0), np.where(inside, dc, 0.0)


def material_t(H, T, mat):
    """(rho_cp, k, d rho_cp/dT, dk/dT) at quadrature points for the
    temperature-dependent materials liquid_solid and gas_liquid_solid (own numpy
    statement of material.templates.hpp:296-495, 540-756): solid fraction
    fs = clip((T_l - T) / (T_l - T_s), 0, 1) (mushy zone) or T < T_s (sharp);
    liquid/solid blend l + (s - l) fs^2 (3 - 2 fs), derivative
    -(s - l) 6 fs (1 - fs) / (T_l - T_s), both exact phase values (derivative 0)
    at fs in {0, 1}; gas_liquid_solid blends the gas value with weight 1 - w
    and scales the derivative by w (w = H, or H > 0.5 when sharp)."""
    T = np.asarray(T, dtype=float)
    span = mat["liquidus"] - mat["solidus"]
    if mat.get("solid_liquid_sharp"):
        fs = (T < mat["solidus"]).astype(float)
    else:
        fs = np.clip((mat["liquidus"] - T) / span, 0.0, 1.0)
    inner = (fs > 0.0) & (fs < 1.0)
    smooth = fs * fs * (3.0 - 2.0 * fs)
    dsmooth = np.where(inner, -6.0 * fs * (1.0 - fs) / (span if span else 1.0), 0.0)

    def blend(liq, sol):
        return liq + (sol - liq) * smooth, (sol - liq) * dsmooth

    k, dk = blend(mat["liq_k"], mat["sol_k"])
    cp, dcp = blend(mat["liq_cp"], mat["sol_cp"])
    ac, dac = apparent_capacity_t(T, mat)
    cp, dcp = cp + ac, dcp + dac
    rho, drho = blend(mat["liq_rho"], mat["sol_rho"])
    vc, dvc = blend(mat["liq_rho"] * mat["liq_cp"], mat["sol_rho"] * mat["sol_cp"])
    if mat["type"] == "gas_liquid_solid":
        H = np.asarray(H, dtype=float)
        w = (H > 0.5).astype(float) if mat.get("sharp") else H

        def gas(g, v, dv):
            return (1.0 - w) * g + w * v, w * dv

        k, dk = gas(mat["gas_k"], k, dk)
        cp, dcp = gas(mat["gas_cp"], cp, dcp)
        rho, drho = gas(mat["gas_rho"], rho, drho)
        vc, dvc = gas(mat["gas_rho"] * mat["gas_cp"], vc, dvc)
    if mat.get("volume_specific"):
        return vc, k, dvc, dk
    return rho * cp, k, dcp * rho + drho * cp, dk


def solidification_parts(n, p, dt, mat, T, T_old, heaviside=None, velocity=None,
                         velocity_value=(0.0, 0.0, 0.0), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                         heat_source=None):
    """(J_full, r_full): the dense tangent linearised at T (T and T_old read
    plain) and the residual right-hand side at T, both without constraints,
    for a temperature-dependent material, by explicit 3D quadrature:
      r_i = -sum_q JxW [(rc(T)/dt (T - T_old) - f + rc(T) u.grad T) phi_i
                        + k(T) grad T . grad phi_i]
      J_ij = sum_q JxW [(rc/dt + drc/dt (T - T_old) + drc u.grad T) phi_j phi_i
                        + rc (u.grad phi_j) phi_i + (k grad phi_j + dk grad T phi_j) . grad phi_i]
    velocity: None or "constant" (velocity_value)."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    n1 = lattice_shape(n, p)
    N = n1[0] * n1[1] * n1[2]
    xq, wq, S, D = _basis_1d(p)
    nq = p + 1
    h = [(hi[d] - lo[d]) / n[d] for d in range(3)]
    Phi = np.einsum("zc,yb,xa->zyxcba", S, S, S).reshape(nq**3, nq**3)
    Gs = [np.einsum("zc,yb,xa->zyxcba", S, S, D).reshape(nq**3, nq**3) / h[0],
          np.einsum("zc,yb,xa->zyxcba", S, D, S).reshape(nq**3, nq**3) / h[1],
          np.einsum("zc,yb,xa->zyxcba", D, S, S).reshape(nq**3, nq**3) / h[2]]
    W = np.einsum("z,y,x->zyx", wq, wq, wq).reshape(-1) * h[0] * h[1] * h[2]
    J = np.zeros((N, N))
    r = np.zeros(N)
    u = np.asarray(velocity_value, dtype=float) if velocity == "constant" else np.zeros(3)
    loc = np.arange(nq)
    for cz in range(n[2]):
        for cy in range(n[1]):
            for cx in range(n[0]):
                ix, iy, iz = cx * p + loc, cy * p + loc, cz * p + loc
                dofs = (ix[None, None, :] + n1[0] * (iy[None, :, None] + n1[1] * iz[:, None, None])).reshape(-1)
                Tq = Phi @ T[dofs]
                Toq = Phi @ T_old[dofs]
                fq = Phi @ heat_source[dofs] if heat_source is not None else np.zeros(nq**3)
                gT = [G @ T[dofs] for G in Gs]
                Hq = Phi @ heaviside[dofs] if heaviside is not None else np.zeros(nq**3)
                rc, k, drc, dk = material_t(Hq, Tq, mat)
                ugT = sum(u[d] * gT[d] for d in range(3))
                val = rc / dt * (Tq - Toq) - fq + rc * ugT
                r[dofs] -= Phi.T @ (W * val) + sum(Gs[d].T @ (W * k * gT[d]) for d in range(3))
                mass = rc / dt + drc / dt * (Tq - Toq) + drc * ugT
                A = Phi.T @ ((W * mass)[:, None] * Phi)
                adv = sum(u[d] * Gs[d] for d in range(3))
                A += Phi.T @ ((W * rc)[:, None] * adv)
                for d in range(3):
                    A += Gs[d].T @ ((W * k)[:, None] * Gs[d])
                    A += Gs[d].T @ ((W * dk * gT[d])[:, None] * Phi)
                J[np.ix_(dofs, dofs)] += A
    return J, r
