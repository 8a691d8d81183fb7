"""ctypes front end of the C++ oracle ``oracle/build/libheat_oracle.so`` (test infrastructure)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libheat_oracle.so")
_lib = None


class OracleProblem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("degree", ctypes.c_int),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("material", ctypes.c_int),
        ("gas_rho", ctypes.c_double),
        ("gas_cp", ctypes.c_double),
        ("gas_k", ctypes.c_double),
        ("liq_rho", ctypes.c_double),
        ("liq_cp", ctypes.c_double),
        ("liq_k", ctypes.c_double),
        ("sharp", ctypes.c_int),
        ("volume_specific", ctypes.c_int),
        ("consistent_with_evaporation", ctypes.c_int),
        ("conductivity_field", ctypes.c_int),
        ("k_amplitude", ctypes.c_double),
        ("velocity", ctypes.c_int),
        ("w0", ctypes.c_double),
        ("dirichlet_faces", ctypes.c_uint),
        ("heaviside", ctypes.POINTER(ctypes.c_double)),
        ("convection_faces", ctypes.c_uint),
        ("convection_coefficient", ctypes.c_double),
        ("convection_t_inf", ctypes.c_double),
        ("radiation_faces", ctypes.c_uint),
        ("emissivity", ctypes.c_double),
        ("radiation_t_inf", ctypes.c_double),
        ("temperature", ctypes.POINTER(ctypes.c_double)),
        ("velocity_value", ctypes.c_double * 3),
        ("sol_rho", ctypes.c_double),
        ("sol_cp", ctypes.c_double),
        ("sol_k", ctypes.c_double),
        ("solidus", ctypes.c_double),
        ("liquidus", ctypes.c_double),
        ("solid_liquid_sharp", ctypes.c_int),
        ("latent_heat", ctypes.c_double),
        ("apparent_capacity", ctypes.c_int),
        ("temperature_old", ctypes.POINTER(ctypes.c_double)),
        ("velocity_field", ctypes.POINTER(ctypes.c_double)),
        ("neumann_faces", ctypes.c_uint),
        ("neumann_flux", (ctypes.c_double * 3) * 6),
    ]

APPARENT_CAPACITY_TYPES = {"poly4_bell": 0, "constant": 1, "qlq": 2}

MATERIAL_TYPES = {"gas": 0, "gas_liquid": 1, "liquid_solid": 2, "gas_liquid_solid": 3}

STEFAN_BOLTZMANN = 5.670374419e-8


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        P = ctypes.POINTER(OracleProblem)
        L.oracle_tables.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
        L.oracle_n_dofs.argtypes = [P]
        L.oracle_n_dofs.restype = ctypes.c_int64
        L.oracle_constrained_mask.argtypes = [P, ctypes.POINTER(ctypes.c_ubyte)]
        L.oracle_random_vector.argtypes = [P, ctypes.c_uint64, dp]
        L.oracle_material.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 7 + [dp] * 4
        L.oracle_material_state.argtypes = [P, ctypes.c_double, ctypes.c_double, dp]
        L.oracle_vmult.argtypes = [P, dp, dp]
        L.oracle_create_rhs.argtypes = [P, dp, dp, dp, dp]
        L.oracle_l2_norm.argtypes = [P, dp, dp]
        L.oracle_diagonal.argtypes = [P, dp]
        L.oracle_inverse_diagonal.argtypes = [P, dp]
        L.oracle_solve.argtypes = [P, ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp,
                                   dp, dp, ctypes.c_int]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().oracle_set_threads(int(n))


def tables(p):
    g = np.zeros(p + 1)
    x = np.zeros(p + 1)
    w = np.zeros(p + 1)
    S = np.zeros((p + 1, p + 1))
    D = np.zeros((p + 1, p + 1))
    rc = lib().oracle_tables(p, _dp(g), _dp(x), _dp(w), _dp(S), _dp(D))
    assert rc == 0
    return g, x, w, S, D


def material(heaviside, mat):
    out = [ctypes.c_double() for _ in range(4)]
    kind = 0 if mat.get("type", "gas") == "gas" else 1
    lib().oracle_material(kind, int(mat.get("sharp", 0)), int(mat.get("volume_specific", 0)),
                          int(mat.get("consistent_with_evaporation", 0)), mat["gas_rho"],
                          mat["gas_cp"], mat["gas_k"], mat.get("liq_rho", 0.0),
                          mat.get("liq_cp", 0.0), mat.get("liq_k", 0.0), heaviside,
                          *[ctypes.byref(o) for o in out])
    rho_cp, k, rho, cp = (o.value for o in out)
    return {"rho_cp": rho_cp, "k": k, "rho": rho, "cp": cp}


def _fill_material(s, mat):
    s.material = MATERIAL_TYPES[mat.get("type", "gas")]
    s.gas_rho, s.gas_cp, s.gas_k = mat.get("gas_rho", 0.0), mat.get("gas_cp", 0.0), mat.get("gas_k", 0.0)
    s.liq_rho = mat.get("liq_rho", 0.0)
    s.liq_cp = mat.get("liq_cp", 0.0)
    s.liq_k = mat.get("liq_k", 0.0)
    s.sol_rho = mat.get("sol_rho", 0.0)
    s.sol_cp = mat.get("sol_cp", 0.0)
    s.sol_k = mat.get("sol_k", 0.0)
    s.solidus = mat.get("solidus", 0.0)
    s.liquidus = mat.get("liquidus", 0.0)
    s.solid_liquid_sharp = int(mat.get("solid_liquid_sharp", 0))
    s.latent_heat = float(mat.get("latent_heat", 0.0))
    s.apparent_capacity = APPARENT_CAPACITY_TYPES[mat.get("apparent_capacity", "qlq")]
    s.sharp = int(mat.get("sharp", 0))
    s.volume_specific = int(mat.get("volume_specific", 0))
    s.consistent_with_evaporation = int(mat.get("consistent_with_evaporation", 0))


def material_state(heaviside, temperature, mat):
    """Material state at (H, T): rho_cp, k, their temperature derivatives and
    rho, cp, d_rho, d_cp."""
    s = OracleProblem()
    _fill_material(s, mat)
    out = np.zeros(8)
    assert lib().oracle_material_state(ctypes.byref(s), heaviside, temperature, _dp(out)) == 0
    return dict(zip(("rho_cp", "k", "d_rho_cp", "d_k", "rho", "cp", "d_rho", "d_cp"),
                    (float(v) for v in out)))


class Problem:
    """Python mirror of ``oracle_problem``; keeps the heaviside and temperature
    arrays alive. ``robin`` (optional dict): convection_faces,
    convection_coefficient, convection_t_inf, radiation_faces, emissivity,
    radiation_t_inf; faces are bit masks like ``dirichlet_faces``."""

    def __init__(self, n, degree, dt=1e-3, material=None, k_amplitude=0.0, velocity=False,
                 w0=0.2, heaviside=None, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                 dirichlet_faces=0x3F, robin=None, temperature=None, velocity_value=None,
                 temperature_old=None, velocity_field=None, neumann=None):
        mat = material or {"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0}
        n = (n, n, n) if np.isscalar(n) else tuple(n)
        s = OracleProblem()
        s.n[:] = [int(v) for v in n]
        s.degree = degree
        s.lo[:] = list(lo)
        s.hi[:] = list(hi)
        s.dt = dt
        _fill_material(s, mat)
        s.conductivity_field = 1 if k_amplitude else 0
        s.k_amplitude = k_amplitude
        # velocity: False/None, True or "vortex", "constant" (velocity_value),
        # "nodal" (velocity_field: 3 n_dofs values, component-blocked)
        s.velocity = {"constant": 2, "nodal": 3}.get(velocity, 1 if velocity else 0)
        s.w0 = w0
        if velocity == "constant":
            s.velocity_value[:] = [float(c) for c in velocity_value]
        s.dirichlet_faces = dirichlet_faces
        rb = robin or {}
        s.convection_faces = int(rb.get("convection_faces", 0))
        s.convection_coefficient = float(rb.get("convection_coefficient", 0.0))
        s.convection_t_inf = float(rb.get("convection_t_inf", 0.0))
        s.radiation_faces = int(rb.get("radiation_faces", 0))
        s.emissivity = float(rb.get("emissivity", 0.0))
        s.radiation_t_inf = float(rb.get("radiation_t_inf", 0.0))
        # neumann: {face f: (qx, qy, qz)}, a constant flux vector per face
        for f, q in (neumann or {}).items():
            s.neumann_faces |= 1 << int(f)
            s.neumann_flux[int(f)][:] = [float(c) for c in q]
        self._T = None
        self._To = None
        self._h = None
        self._v = None
        if velocity == "nodal":
            self._v = np.ascontiguousarray(velocity_field, dtype=np.float64)
            s.velocity_field = _dp(self._v)
        if heaviside is not None:
            self._h = np.ascontiguousarray(heaviside, dtype=np.float64)
            s.heaviside = _dp(self._h)
        self.s = s
        self.n = n
        self.degree = degree
        self.N = int(lib().oracle_n_dofs(ctypes.byref(s)))
        assert self.N > 0, "invalid oracle problem"
        if temperature is not None:
            self.set_temperature(temperature)
        if temperature_old is not None:
            self.set_temperature_old(temperature_old)

    def set_temperature(self, temperature):
        """The temperature the tangent is linearised at (the operator's
        reference to the current solution)."""
        self._T = np.array(temperature, dtype=np.float64)
        assert self._T.shape == (self.N,)
        self.s.temperature = _dp(self._T)

    def set_heaviside(self, heaviside):
        """Replace the heaviside field (a time-dependent prescribed heaviside)."""
        self._h = np.ascontiguousarray(heaviside, dtype=np.float64)
        assert self._h.shape == (self.N,)
        self.s.heaviside = _dp(self._h)

    def set_temperature_old(self, temperature_old):
        """The old temperature the tangent of a temperature-dependent material
        reads (the operator's reference to temperature_old)."""
        self._To = np.array(temperature_old, dtype=np.float64)
        assert self._To.shape == (self.N,)
        self.s.temperature_old = _dp(self._To)

    def mask(self):
        m = np.zeros(self.N, dtype=np.uint8)
        lib().oracle_constrained_mask(ctypes.byref(self.s), m.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return m.astype(bool)

    def set_dt(self, dt):
        """OperatorBase::reset_time_increment."""
        self.s.dt = float(dt)

    def random(self, seed):
        v = np.zeros(self.N)
        lib().oracle_random_vector(ctypes.byref(self.s), seed, _dp(v))
        return v

    def vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros(self.N)
        assert lib().oracle_vmult(ctypes.byref(self.s), _dp(src), _dp(dst)) == 0
        return dst

    def create_rhs(self, temperature, temperature_old, heat_source=None, dst=None):
        """dst + the residual right-hand side (create_rhs adds into dst)."""
        tn = np.ascontiguousarray(temperature, dtype=np.float64)
        to = np.ascontiguousarray(temperature_old, dtype=np.float64)
        f = None if heat_source is None else np.ascontiguousarray(heat_source, dtype=np.float64)
        out = np.zeros(self.N) if dst is None else np.array(dst, dtype=np.float64)
        assert lib().oracle_create_rhs(ctypes.byref(self.s), _dp(tn), _dp(to),
                                       _dp(f) if f is not None else None, _dp(out)) == 0
        return out

    def l2_norm(self, v):
        """L2 norm of the FE field v with the operator's Gauss quadrature."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(1)
        assert lib().oracle_l2_norm(ctypes.byref(self.s), _dp(v), _dp(out)) == 0
        return float(out[0])

    def diagonal(self):
        d = np.zeros(self.N)
        assert lib().oracle_diagonal(ctypes.byref(self.s), _dp(d)) == 0
        return d

    def inverse_diagonal(self):
        d = np.zeros(self.N)
        assert lib().oracle_inverse_diagonal(ctypes.byref(self.s), _dp(d)) == 0
        return d

    def solve(self, b, solver="CG", rel_tol=1e-8, abs_tol=1e-20, max_it=10000, inv_diag=None,
              restart=30, x0=None, history=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        it = ctypes.c_int()
        r0 = ctypes.c_double()
        rf = ctypes.c_double()
        hist = np.zeros(max(history, 1))
        dptr = _dp(inv_diag) if inv_diag is not None else None
        status = lib().oracle_solve(ctypes.byref(self.s), 0 if solver == "CG" else 1, restart,
                                    _dp(b), _dp(x), dptr, rel_tol, abs_tol, max_it,
                                    ctypes.byref(it), ctypes.byref(r0), ctypes.byref(rf),
                                    _dp(hist), history)
        return {"x": x, "status": status, "iterations": it.value, "initial_res": r0.value,
                "final_res": rf.value, "history": hist[: min(history, it.value + 1)]}
