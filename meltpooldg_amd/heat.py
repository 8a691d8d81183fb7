"""Host-side mirror of MeltPoolDG's heat-operator interface over the C ABI.

Names, argument meaning and error behaviour follow the reference:

* ``HeatDiffuseOperator``  <- HeatDiffuseMultiPhaseOperator<3,double>
  (include/meltpooldg/heat/heat_diffuse_operator.hpp:109-366) as an
  OperatorMatrixFree (include/meltpooldg/core/operator_base.hpp:189-283):
  ``vmult(dst, src)``, ``compute_inverse_diagonal_from_matrixfree(diag)``,
  ``reset_time_increment(dt)``.
* ``JacobiPreconditioner`` <- include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:25-89
* ``LinearSolverData``     <- include/meltpooldg/linear_algebra/linear_solver_data.hpp:37-51
* ``LinearSolver.solve``   <- include/meltpooldg/linear_algebra/linear_solver.hpp:26-134;
  returns the iteration count, raises ``NoConvergence`` where deal.II throws
  SolverControl::NoConvergence.

Vectors are torch tensors on the GPU (torch is only the device-memory and
stream plumbing); every arithmetic operation runs in libmpdg_heat.so.
"""
import ctypes
from dataclasses import dataclass, field

from . import _lib as L


class NoConvergence(RuntimeError):
    """dealii::SolverControl::NoConvergence (linear_solver.hpp:124-129)."""

    def __init__(self, iterations, residual, message=""):
        super().__init__(f"Iterative method reported convergence failure in step {iterations}. "
                         f"The residual in the last step was {residual}. {message}")
        self.last_step = iterations
        self.last_residual = residual


class ZeroTimeIncrement(ValueError):
    """ExcZeroTimeIncrement (include/meltpooldg/core/exceptions.hpp:6)."""


@dataclass
class PhaseData:
    """One phase of MaterialData (density, specific heat capacity, conductivity)."""
    density: float = 1.0
    specific_heat_capacity: float = 1.0
    thermal_conductivity: float = 1.0


MATERIAL_TYPES = {"gas": 0, "gas_liquid": 1, "liquid_solid": 2, "gas_liquid_solid": 3}
APPARENT_CAPACITY_TYPES = {"poly4_bell": 0, "constant": 1, "qlq": 2}


@dataclass
class MaterialData:
    """Material model of include/meltpooldg/core/material.hpp:46-55 ('gas',
    'gas_liquid', 'liquid_solid' or 'gas_liquid_solid') with the MaterialData
    fields it reads (include/meltpooldg/core/material_data.hpp). The
    temperature-dependent types take the solid fraction from the solidus and
    liquidus temperatures ('mushy_zone' or 'sharp') and, with a latent heat of
    fusion > 0, add its apparent capacity ('poly4_bell', 'constant' or 'qlq',
    include/meltpooldg/heat/apparent_capacity.hpp) to c_p."""
    type: str = "gas"
    gas: PhaseData = field(default_factory=PhaseData)
    liquid: PhaseData = field(default_factory=PhaseData)
    solid: PhaseData = field(default_factory=PhaseData)
    two_phase_fluid_properties_transition_type: str = "smooth"
    use_volume_specific_thermal_capacity_for_phase_interpolation: bool = False
    solid_liquid_properties_transition_type: str = "mushy_zone"
    solidus_temperature: float = 0.0
    liquidus_temperature: float = 0.0
    latent_heat_of_fusion: float = 0.0
    apparent_capacity_type: str = "qlq"


@dataclass
class RadiationBC:
    """HeatData::RadiationBC (include/meltpooldg/heat/heat_data.hpp:46-50)."""
    emissivity: float = 0.0
    temperature_infinity: float = 0.0


@dataclass
class ConvectionBC:
    """HeatData::ConvectionBC (include/meltpooldg/heat/heat_data.hpp:52-56)."""
    convection_coefficient: float = 0.0
    temperature_infinity: float = 0.0


def _face_mask(boundary_ids):
    """Colorized boundary ids of the box (2d: x_d = lo, 2d + 1: x_d = hi) as a bit mask."""
    mask = 0
    for b in boundary_ids or ():
        if not 0 <= int(b) < 6:
            raise ValueError(f"boundary id {b} is not a face of the box (0..5)")
        mask |= 1 << int(b)
    return mask


@dataclass
class LinearSolverData:
    """include/meltpooldg/linear_algebra/linear_solver_data.hpp:37-51 (same defaults)."""
    do_matrix_free: bool = True
    preconditioner_type: str = "Identity"
    solver_type: str = "GMRES"
    max_iterations: int = 10000
    rel_tolerance: float = 1e-12
    abs_tolerance: float = 1e-20
    monitor_type: str = "none"
    # engine extensions
    mixed_precision: bool = False
    inner_rel_tolerance: float = 1e-2
    check_interval: int = 0
    gmres_restart: int = 30


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _make_desc(n_cells, degree, material, box, precision, conductivity_amplitude, velocity,
               velocity_w0, dirichlet_faces, stream, convection=None, radiation=None,
               convection_boundary_ids=(), radiation_boundary_ids=(), velocity_value=None,
               neumann_bc=None):
    n = (n_cells,) * 3 if isinstance(n_cells, int) else tuple(n_cells)
    mat = material or MaterialData()
    d = L.HeatDesc()
    d.n_cells[:] = [int(v) for v in n]
    d.degree = int(degree)
    d.box_lo[:] = [float(v) for v in box[0]]
    d.box_hi[:] = [float(v) for v in box[1]]
    d.precision = L.MPDG_FP32 if precision == "fp32" else L.MPDG_FP64
    if mat.type not in MATERIAL_TYPES:
        raise ValueError(f"material type {mat.type!r} is not built ({sorted(MATERIAL_TYPES)})")
    if mat.apparent_capacity_type not in APPARENT_CAPACITY_TYPES:
        raise ValueError(f"apparent capacity type {mat.apparent_capacity_type!r} is not one of "
                         f"{sorted(APPARENT_CAPACITY_TYPES)}")
    if mat.solid_liquid_properties_transition_type not in ("mushy_zone", "sharp"):
        raise ValueError("solid liquid properties transition type must be mushy_zone or sharp")
    d.material = MATERIAL_TYPES[mat.type]
    for dst, src in ((d.gas, mat.gas), (d.liquid, mat.liquid), (d.solid, mat.solid)):
        dst.density = src.density
        dst.specific_heat_capacity = src.specific_heat_capacity
        dst.thermal_conductivity = src.thermal_conductivity
    d.two_phase_sharp = 1 if mat.two_phase_fluid_properties_transition_type == "sharp" else 0
    d.volume_specific_heat = int(mat.use_volume_specific_thermal_capacity_for_phase_interpolation)
    d.solidus_temperature = float(mat.solidus_temperature)
    d.liquidus_temperature = float(mat.liquidus_temperature)
    d.solid_liquid_sharp = int(mat.solid_liquid_properties_transition_type == "sharp")
    d.latent_heat_of_fusion = float(mat.latent_heat_of_fusion)
    d.apparent_capacity_type = APPARENT_CAPACITY_TYPES[mat.apparent_capacity_type]
    d.conductivity_field = 1 if conductivity_amplitude else 0
    d.conductivity_amplitude = float(conductivity_amplitude)
    d.velocity = {None: 0, "none": 0, "vortex": 1, "constant": 2, "nodal": 3}[velocity]
    d.velocity_w0 = float(velocity_w0)
    if velocity == "constant":
        v = tuple(float(c) for c in velocity_value)
        d.velocity_value[:] = list(v) + [0.0] * (3 - len(v))
    d.dirichlet_faces = int(dirichlet_faces)
    conv = convection or ConvectionBC()
    rad = radiation or RadiationBC()
    d.convection_faces = _face_mask(convection_boundary_ids)
    d.convection_coefficient = float(conv.convection_coefficient)
    d.convection_temperature_infinity = float(conv.temperature_infinity)
    d.radiation_faces = _face_mask(radiation_boundary_ids)
    d.emissivity = float(rad.emissivity)
    d.radiation_temperature_infinity = float(rad.temperature_infinity)
    for face, flux in (neumann_bc or {}).items():
        q = [float(c) for c in flux]
        if not 0 <= int(face) < 6 or len(q) != 3:
            raise ValueError("neumann_bc maps a face id 0..5 to a flux vector (qx, qy, qz)")
        d.neumann_faces |= 1 << int(face)
        d.neumann_flux[int(face)][:] = q
    d.stream = stream
    return d


def lattice_index_of_points(n_cells, degree, points, box=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))):
    """Lattice index of each support point (numpy array of shape (n_dofs, 3),
    e.g. DoFTools::map_dofs_to_support_points of the reference's DoFHandler) of
    FE_Q(degree) on the box mesh; numpy uint32 array. Raises on a point that is
    not a support point or a support point given twice. Host computation."""
    import numpy as np
    d = _make_desc(n_cells, degree, None, box, "fp64", 0.0, None, 0.2, 0, None)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(pts.shape[0], dtype=np.uint32)
    lib = L.load()
    L.check(lib.mpdg_lattice_index_of_points(ctypes.byref(d), ctypes.c_void_p(pts.ctypes.data),
                                             pts.shape[0], ctypes.c_void_p(out.ctypes.data)))
    return out


class HeatDiffuseOperator:
    """Matrix-free heat tangent J = rho_cp/dt M + K_k + rho_cp C_u on a box mesh.

    Parameters mirror the reference's set-up: FE_Q(degree) on Gauss-Lobatto
    points with QGauss(degree+1) on ``n_cells`` affine hex cells of the box,
    homogeneous Dirichlet on ``dirichlet_faces`` (bit 2d: x_d = lo, 2d+1: hi).
    Robin faces as the reference attaches them: ``convection_boundary_ids`` /
    ``radiation_boundary_ids`` (the bc_convection_indices / bc_radiation_indices
    of heat_diffuse_operator.cpp:76-77, colorized ids 0..5 of the box) with the
    parameters ``convection`` (ConvectionBC) and ``radiation`` (RadiationBC);
    radiation makes the tangent depend on the temperature registered with
    ``set_temperature``. ``velocity``: None, "vortex" (analytic at q, z component
    ``velocity_w0``), "nodal" (a DoF-vector field given by ``set_velocity``,
    FP64) or "constant" (``velocity_value``, a prescribed_velocity
    field constant in space). ``neumann_bc``: {face id: (qx, qy, qz)}, a heat
    flux constant on the face, whose -(q . n) create_rhs integrates (the
    reference's neumann_bc functions, heat_diffuse_operator.cpp:735-767); not on
    a Robin face.
    """

    def __init__(self, n_cells, degree, material=None, box=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
                 precision="fp64", conductivity_amplitude=0.0, velocity=None, velocity_w0=0.2,
                 dirichlet_faces=0x3F, stream=None, convection=None, radiation=None,
                 convection_boundary_ids=(), radiation_boundary_ids=(), velocity_value=None,
                 neumann_bc=None):
        self._lib = L.load()
        n = (n_cells,) * 3 if isinstance(n_cells, int) else tuple(n_cells)
        mat = material or MaterialData()
        d = _make_desc(n_cells, degree, mat, box, precision, conductivity_amplitude, velocity,
                       velocity_w0, dirichlet_faces, stream if stream is not None else _stream_handle(),
                       convection, radiation, convection_boundary_ids, radiation_boundary_ids,
                       velocity_value, neumann_bc)
        self.desc = d
        h = ctypes.c_void_p()
        L.check(self._lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.n_cells = n
        self.degree = degree
        self.material = mat
        self._precision = "fp32" if precision == "fp32" else "fp64"

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mpdg_heat_destroy(h)
            self._h = None

    # -- OperatorBase ---------------------------------------------------------
    def reset_time_increment(self, dt):
        st = self._lib.mpdg_heat_set_dt(self._h, float(dt))
        if st == L.MPDG_ERR_ZERO_TIME_INCREMENT:
            raise ZeroTimeIncrement(self._lib.mpdg_last_error().decode())
        L.check(st)
        self.time_increment = float(dt)

    def n_dofs(self):
        return int(self._lib.mpdg_heat_n_dofs(self._h))

    def n_free_dofs(self):
        return int(self._lib.mpdg_heat_n_free_dofs(self._h))

    def set_temperature(self, temperature):
        """The temperature the tangent is linearised at: the operator's reference
        to the current solution (heat_diffuse_operator.hpp:128). The engine keeps
        the device pointer, so later changes of ``temperature`` are seen; None
        clears it."""
        if temperature is None:
            L.check(self._lib.mpdg_heat_set_temperature(self._h, None))
            self._temperature = None
            return
        self._check_vec(temperature, "float64")
        L.check(self._lib.mpdg_heat_set_temperature(self._h, _ptr(temperature)))
        self._temperature = temperature

    def set_temperature_old(self, temperature_old):
        """The old temperature a temperature-dependent material's tangent reads
        (the operator's reference to temperature_old, heat_diffuse_operator.hpp:129).
        Kept by pointer like set_temperature; None clears it."""
        if temperature_old is None:
            L.check(self._lib.mpdg_heat_set_temperature_old(self._h, None))
            self._temperature_old = None
            return
        self._check_vec(temperature_old, "float64")
        L.check(self._lib.mpdg_heat_set_temperature_old(self._h, _ptr(temperature_old)))
        self._temperature_old = temperature_old

    def to_lattice(self, lattice, dofs, lattice_index):
        """lattice[lattice_index[i]] = dofs[i] (device tensors; lattice_index an int32
        / uint32 device tensor from lattice_index_of_points)."""
        self._check_vec(lattice, "float64")
        self._check_vec(dofs, "float64")
        self._check_index(lattice_index)
        L.check(self._lib.mpdg_heat_to_lattice(self._h, _ptr(lattice), _ptr(dofs), _ptr(lattice_index)))

    def from_lattice(self, dofs, lattice, lattice_index):
        """dofs[i] = lattice[lattice_index[i]]."""
        self._check_vec(lattice, "float64")
        self._check_vec(dofs, "float64")
        self._check_index(lattice_index)
        L.check(self._lib.mpdg_heat_from_lattice(self._h, _ptr(dofs), _ptr(lattice), _ptr(lattice_index)))

    def _check_index(self, index):
        import torch
        if not (index.is_cuda and index.dtype in (torch.int32, torch.uint32)
                and index.is_contiguous() and index.numel() == self.n_dofs()):
            raise ValueError("lattice_index must be a contiguous int32/uint32 device tensor "
                             "of n_dofs() entries")

    def set_graphs(self, enable=True):
        """Replay the Krylov iteration chunks of this operator's solves as
        captured hipGraphs (default) or launch every kernel eagerly."""
        L.check(self._lib.mpdg_heat_set_graphs(self._h, 1 if enable else 0))

    def enable_history_data(self, length=10001):
        """SolverControl::enable_history_data: the FP64 CG solves record
        ||r_k|| for k = 0 .. iterations into a device buffer of `length`
        entries (get_history_data reads it)."""
        import torch
        self._history = torch.zeros(int(length), dtype=torch.float64, device="cuda")
        L.check(self._lib.mpdg_heat_set_residual_history(self._h, _ptr(self._history),
                                                         int(length)))

    def get_history_data(self, iterations=None):
        """The recorded residual norms of the last FP64 CG solve (host numpy),
        entries 0 .. iterations (default: up to the last nonzero)."""
        import torch
        torch.cuda.synchronize()
        h = self._history.cpu().numpy()
        if iterations is None:
            nz = h.nonzero()[0]
            iterations = int(nz[-1]) if nz.size else 0
        return h[: iterations + 1].copy()

    def set_velocity(self, velocity):
        """The nodal velocity of a ``velocity="nodal"`` operator (the DoF vector
        the reference operator reads with vel_eval.read_dof_values_plain,
        heat_diffuse_operator.cpp:975-980): an FP64 device tensor of 3 n_dofs()
        values, component-blocked [v_x | v_y | v_z] in lattice numbering, or
        of shape (3, n_dofs()). The engine copies it."""
        v = velocity.reshape(-1)
        if not (getattr(v, "is_cuda", False) and str(v.dtype).endswith("float64")
                and v.is_contiguous() and v.numel() == 3 * self.n_dofs()):
            raise ValueError("velocity must be a contiguous FP64 GPU tensor of 3 n_dofs() values")
        L.check(self._lib.mpdg_heat_set_velocity(self._h, _ptr(v)))

    def set_heaviside(self, heaviside, fe="FE_Q"):
        """Register the level-set heaviside field (read plain, interpolated at q).
        fe: the level set's finite element, "FE_Q" (the heat element's space)
        or "FE_Q_iso_Q1" (p linear sub-elements per cell, equidistant nodes in
        lattice order): the latter is first interpolated to the heat element's
        support points, as the reference does when the level set's element is
        FE_Q_iso_Q1 (heat_diffuse_operator.cpp:834-838, 982-993)."""
        self._check_vec(heaviside, "float64")
        if fe == "FE_Q":
            L.check(self._lib.mpdg_heat_set_heaviside(self._h, _ptr(heaviside)))
        elif fe == "FE_Q_iso_Q1":
            L.check(self._lib.mpdg_heat_set_heaviside_iso_q1(self._h, _ptr(heaviside)))
        else:
            raise ValueError(f"level-set element {fe!r}: FE_Q or FE_Q_iso_Q1")
        self._heaviside = heaviside

    # -- OperatorMatrixFree ---------------------------------------------------
    def vmult(self, dst, src):
        if str(src.dtype).endswith("float32"):
            self._check_vec(dst, "float32")
            self._check_vec(src, "float32")
            L.check(self._lib.mpdg_heat_vmult_f32(self._h, _ptr(dst), _ptr(src)))
            return
        self._check_vec(dst, "float64")
        self._check_vec(src, "float64")
        L.check(self._lib.mpdg_heat_vmult(self._h, _ptr(dst), _ptr(src)))

    def compute_inverse_diagonal_from_matrixfree(self, diag):
        self._check_vec(diag, "float64")
        L.check(self._lib.mpdg_heat_inverse_diagonal(self._h, _ptr(diag)))

    _VARIANTS = {"line": L.MPDG_VMULT_LINE, "cell": L.MPDG_VMULT_CELL, "mfma": L.MPDG_VMULT_LINE_MFMA,
                 "cell_mfma": L.MPDG_VMULT_CELL_MFMA}

    def _prec_code(self, precision):
        precision = self._precision if precision is None else precision
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        return L.MPDG_FP32 if precision == "fp32" else L.MPDG_FP64

    @classmethod
    def _tuning_dict(cls, t):
        names = {v: k for k, v in cls._VARIANTS.items()}
        return {"z_layers": t.z_layers, "blocks_per_cu": t.blocks_per_cu,
                "variant": names.get(t.variant, t.variant), "tile_log2": t.tile_log2,
                "schedule": {0: "passes", 1: "ordered", 2: "slabs"}[t.schedule]}

    def create_rhs(self, dst, src, temperature, heat_source=None):
        """HeatDiffuseMultiPhaseOperator::create_rhs(dst, src = temperature_old)
        (source/heat/heat_diffuse_operator.cpp:783-806): adds the residual
        right-hand side to ``dst`` on the free rows. The reference operator holds
        ``temperature`` and ``heat_source`` from its constructor; here they are
        arguments (``heat_source`` None: no source). FP64 vectors."""
        for v in (dst, src, temperature) + ((heat_source,) if heat_source is not None else ()):
            self._check_vec(v, "float64")
        L.check(self._lib.mpdg_heat_create_rhs(self._h, _ptr(dst), _ptr(temperature), _ptr(src),
                                               _ptr(heat_source) if heat_source is not None else None))

    def get_tuning(self, precision=None):
        """Launch tuning of the vmult kernels for one precision (engine extension)."""
        t = L.Tuning()
        L.check(self._lib.mpdg_heat_get_tuning(self._h, self._prec_code(precision), ctypes.byref(t)))
        return self._tuning_dict(t)

    def set_tuning(self, z_layers=None, blocks_per_cu=None, variant=None, tile_log2=None,
                   precision=None, schedule=None):
        cur = self.get_tuning(precision)
        t = L.Tuning()
        t.z_layers = int(cur["z_layers"] if z_layers is None else z_layers)
        t.blocks_per_cu = int(cur["blocks_per_cu"] if blocks_per_cu is None else blocks_per_cu)
        v = cur["variant"] if variant is None else variant
        t.variant = self._VARIANTS[v] if isinstance(v, str) else int(v)
        t.tile_log2 = int(cur["tile_log2"] if tile_log2 is None else tile_log2)
        sc = cur["schedule"] if schedule is None else schedule
        if sc not in ("passes", "ordered", "slabs"):
            raise ValueError("schedule must be 'passes', 'ordered' or 'slabs'")
        t.schedule = {"passes": 0, "ordered": 1, "slabs": 2}[sc]
        L.check(self._lib.mpdg_heat_set_tuning(self._h, self._prec_code(precision), ctypes.byref(t)))

    def autotune(self, precision=None, repeats=5):
        """Times the candidate kernel forms and launch shapes of this operator's
        vmult, installs the fastest and returns it (engine extension)."""
        t = L.Tuning()
        L.check(self._lib.mpdg_heat_autotune(self._h, self._prec_code(precision), int(repeats),
                                             ctypes.byref(t)))
        return self._tuning_dict(t)

    def time_vmult(self, dst, src, repeats=10):
        ms = ctypes.c_double()
        prec = L.MPDG_FP32 if str(src.dtype).endswith("float32") else L.MPDG_FP64
        L.check(self._lib.mpdg_heat_time_vmult(self._h, _ptr(dst), _ptr(src), int(repeats), prec,
                                               ctypes.byref(ms)))
        return ms.value

    def _check_vec(self, v, dtype):
        if not getattr(v, "is_cuda", False):
            raise ValueError("vectors must be GPU tensors")
        if not str(v.dtype).endswith(dtype):
            raise ValueError(f"expected {dtype}, got {v.dtype}")
        if not v.is_contiguous() or v.numel() != self.n_dofs():
            raise ValueError(f"expected a contiguous vector of {self.n_dofs()} entries")


class JacobiPreconditioner:
    """preconditioner_jacobi.hpp:25-89: update() -> inverse diagonal, vmult(dst, src).
    The inverse diagonal starts as a zero vector, as reinit()'s
    initialize_dof_vector leaves it; update() computes it."""

    def __init__(self, operator):
        import torch
        self.eq_operator = operator
        self.inverse_diag = torch.zeros(operator.n_dofs(), dtype=torch.float64, device="cuda")

    def update(self, *_):
        self.eq_operator.compute_inverse_diagonal_from_matrixfree(self.inverse_diag)

    def reinit(self):
        self.inverse_diag.zero_()

    def vmult(self, dst, src):
        op = self.eq_operator
        L.check(op._lib.mpdg_heat_jacobi_vmult(op._h, _ptr(dst), _ptr(src), _ptr(self.inverse_diag)))


def _solver_params(lib, data, use_diag=None):
    prm = L.SolverParams()
    lib.mpdg_solver_params_default(ctypes.byref(prm))
    prm.solver_type = {"CG": 0, "GMRES": 1}[data.solver_type]
    if use_diag is None:
        if data.preconditioner_type not in ("Identity", "Diagonal"):
            raise ValueError(f"preconditioner type {data.preconditioner_type!r} is not built "
                             "(Identity, Diagonal)")
        use_diag = data.preconditioner_type == "Diagonal"
    prm.preconditioner_type = 3 if use_diag else 0
    prm.max_iterations = int(data.max_iterations)
    prm.rel_tolerance = float(data.rel_tolerance)
    prm.abs_tolerance = float(data.abs_tolerance)
    prm.mixed_precision = 1 if data.mixed_precision else 0
    prm.inner_rel_tolerance = float(data.inner_rel_tolerance)
    prm.check_interval = int(data.check_interval)
    prm.gmres_restart = int(getattr(data, "gmres_restart", 30))
    return prm


class LinearSolver:
    """linear_solver.hpp:26-134."""

    last_report = None

    @staticmethod
    def solve(system_matrix, solution, rhs, data, preconditioner=None, identifier=""):
        """Solve system_matrix * solution = rhs; returns the iteration count.

        ``preconditioner`` decides when it is given, as the object passed to the
        reference's LinearSolver::solve does:

        * a JacobiPreconditioner is applied with its inverse diagonal as the
          caller last computed it (``update()``); it is not recomputed here
          (mpdg_heat_solve_preconditioned);
        * the string ``"Identity"`` is PreconditionIdentity.

        Without one, ``data.preconditioner_type`` decides ("Identity" or
        "Diagonal"; others are not built and raise): "Diagonal" recomputes the
        inverse diagonal on the device first, as HeatDiffuseOperation::solve
        updates its preconditioner before every solve
        (source/heat/heat_diffuse_operation.cpp:275-276).
        """
        op = system_matrix
        op._check_vec(solution, "float64")
        op._check_vec(rhs, "float64")
        rep = L.SolverReport()
        if preconditioner is None:
            prm = _solver_params(op._lib, data)
            st = op._lib.mpdg_heat_solve(op._h, _ptr(solution), _ptr(rhs), ctypes.byref(prm),
                                         ctypes.byref(rep))
        elif isinstance(preconditioner, JacobiPreconditioner) or preconditioner == "Identity":
            prm = _solver_params(op._lib, data, use_diag=preconditioner != "Identity")
            inv = None
            if isinstance(preconditioner, JacobiPreconditioner):
                if preconditioner.eq_operator is not op:
                    raise ValueError("the preconditioner belongs to another operator")
                inv = _ptr(preconditioner.inverse_diag)
            st = op._lib.mpdg_heat_solve_preconditioned(op._h, _ptr(solution), _ptr(rhs), inv,
                                                        ctypes.byref(prm), ctypes.byref(rep))
        else:
            raise ValueError(f"preconditioner {preconditioner!r} is not built "
                             "(a JacobiPreconditioner, 'Identity' or None)")
        LinearSolver.last_report = {"iterations": rep.iterations,
                                    "outer_iterations": rep.outer_iterations,
                                    "initial_residual": rep.initial_residual,
                                    "final_residual": rep.final_residual,
                                    "seconds": rep.seconds}
        if st == L.MPDG_ERR_NOT_CONVERGED:
            raise NoConvergence(rep.iterations, rep.final_residual, identifier)
        L.check(st)
        return rep.iterations


@dataclass
class NonlinearSolverData:
    """include/meltpooldg/linear_algebra/nonlinear_solver_data.hpp:10-20 (same defaults)."""
    max_nonlinear_iterations: int = 10
    field_correction_tolerance: float = 1e-10
    residual_tolerance: float = 1e-9
    max_nonlinear_iterations_alt: int = 0
    field_correction_tolerance_alt: float = 1e-9
    residual_tolerance_alt: float = 1e-8
    use_relative_tolerances: bool = False
    verbosity_level: int = -1


class NewtonDidNotConverge(RuntimeError):
    """ExcNewtonDidNotConverge (newton_raphson_solver.cpp:79) /
    ExcHeatTransferNoConvergence (heat_diffuse_operation.cpp:282-285)."""


class HeatDiffuseOperation:
    """Time stepping of the heat equation on the device
    (source/heat/heat_diffuse_operation.cpp):

    * solution history T (current), T_old, T_old_old and the previous time
      increment;
    * ``init_time_advance(dt, dt_old=None)``: reset_time_increment, the
      predictor (Predictor::vmult, include/meltpooldg/linear_algebra/
      predictor.hpp:118-160: "none" keeps the current solution, "zero",
      "linear_extrapolation", the heat default, heat_data.cpp:13,
      predictor_linear.cpp:76-96), homogeneous constraints distributed.
      ``dt_old`` is the time iterator's old increment (default: the previous
      call's dt, 0 on the first call);
    * ``solve()``: NewtonRaphsonSolver::solve on the residual create_rhs with
      the tangent solve (GMRES + Diagonal by default, heat_data.cpp:11-12);
    * ``l2_norm()``: the norm the reference prints after each Newton solve
      (compute_norm of the current solution, heat_diffuse_operation.cpp:112-115).

    ``temperature``: initial condition (FP64 device vector, copied)."""

    def __init__(self, operator, temperature, heat_source=None, linear_solver=None,
                 nlsolve=None, predictor="linear_extrapolation"):
        if predictor not in ("none", "zero", "linear_extrapolation"):
            raise ValueError(f"predictor type {predictor!r} is not built "
                             "(none, zero, linear_extrapolation)")
        self.predictor = predictor
        import torch
        self.op = operator
        self.temperature = temperature.clone()
        self.temperature_old = temperature.clone()
        self.temperature_old_old = temperature.clone()
        self.heat_source = heat_source
        self.linear_solver = linear_solver or LinearSolverData(solver_type="GMRES",
                                                               preconditioner_type="Diagonal")
        self.nlsolve = nlsolve or NonlinearSolverData()
        self.dt = 0.0
        self.dt_old = 0.0
        self.last_report = None
        self._torch = torch

    def init_time_advance(self, dt, dt_old=None):
        lib, h = self.op._lib, self.op._h
        self.temperature_old_old.copy_(self.temperature_old)
        self.temperature_old.copy_(self.temperature)
        self.dt_old, self.dt = (self.dt if dt_old is None else float(dt_old)), float(dt)
        self.op.reset_time_increment(self.dt)
        if self.predictor == "linear_extrapolation":
            L.check(lib.mpdg_heat_linear_predictor(h, _ptr(self.temperature),
                                                   _ptr(self.temperature_old),
                                                   _ptr(self.temperature_old_old), self.dt,
                                                   self.dt_old))
        elif self.predictor == "zero":
            self.temperature.zero_()
        L.check(lib.mpdg_heat_distribute_constraints(h, _ptr(self.temperature)))

    def solve(self):
        lib, h = self.op._lib, self.op._h
        n = self.nlsolve
        np_ = L.NewtonParams(int(n.max_nonlinear_iterations), float(n.field_correction_tolerance),
                             float(n.residual_tolerance), int(n.max_nonlinear_iterations_alt),
                             float(n.field_correction_tolerance_alt),
                             float(n.residual_tolerance_alt), 1 if n.use_relative_tolerances else 0)
        lp = _solver_params(lib, self.linear_solver)
        rep = L.NewtonReport()
        f = self.heat_source
        # the solve points the operator at the solution and the old solution
        self.op._temperature = self.temperature
        self.op._temperature_old = self.temperature_old
        st = lib.mpdg_heat_newton_solve(h, _ptr(self.temperature), _ptr(self.temperature_old),
                                        _ptr(f) if f is not None else None, ctypes.byref(np_),
                                        ctypes.byref(lp), ctypes.byref(rep))
        self.last_report = {"iterations": rep.iterations, "linear_iterations": rep.linear_iterations,
                            "initial_residual": rep.initial_residual, "residual": rep.residual,
                            "update": rep.update, "seconds": rep.seconds}
        if st == L.MPDG_ERR_NOT_CONVERGED:
            raise NewtonDidNotConverge(lib.mpdg_last_error().decode())
        L.check(st)
        return self.last_report

    def advance(self, dt, dt_old=None):
        self.init_time_advance(dt, dt_old)
        return self.solve()

    def l2_norm(self, v=None):
        out = ctypes.c_double()
        L.check(self.op._lib.mpdg_heat_l2_norm(self.op._h, _ptr(self.temperature if v is None else v),
                                               ctypes.byref(out)))
        return out.value


class HeatBatch:
    """Independent heat solves of differing mesh size and degree solved together
    (batched vmult per degree and colour, segmented vector kernels, per-solve
    convergence, hipGraph-captured iteration chunks). ``operators`` are
    HeatDiffuseOperator-style parameter dicts: n_cells, degree and optionally
    material (MaterialData), conductivity_amplitude, velocity, dirichlet_faces."""

    def __init__(self, members, stream=None):
        self._lib = L.load()
        n = len(members)
        descs = (L.HeatDesc * n)()
        for i, m in enumerate(members):
            d = _make_desc(m.get("n_cells"), m.get("degree"), m.get("material"),
                           m.get("box", ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))),
                           m.get("precision", "fp64"), m.get("conductivity_amplitude", 0.0),
                           m.get("velocity"), m.get("velocity_w0", 0.2),
                           m.get("dirichlet_faces", 0x3F), stream if stream is not None else 0)
            descs[i] = d
        h = ctypes.c_void_p()
        L.check(self._lib.mpdg_heat_batch_create(descs, n, ctypes.byref(h)))
        self._h = h
        self.count = n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mpdg_heat_batch_destroy(h)
            self._h = None

    def reset_time_increment(self, dt):
        L.check(self._lib.mpdg_heat_batch_set_dt(self._h, float(dt)))

    def n_dofs(self, i):
        return int(self._lib.mpdg_heat_batch_n_dofs(self._h, i))

    def solve(self, xs, bs, data):
        """Solves member i for xs[i] (initial guess on entry) with rhs bs[i];
        returns the per-member reports (iterations, residuals, seconds to convergence)."""
        n = self.count
        xp = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        bp = (ctypes.c_void_p * n)(*[b.data_ptr() for b in bs])
        prm = L.SolverParams()
        self._lib.mpdg_solver_params_default(ctypes.byref(prm))
        prm.solver_type = 0
        prm.max_iterations = int(data.max_iterations)
        prm.rel_tolerance = float(data.rel_tolerance)
        prm.abs_tolerance = float(data.abs_tolerance)
        prm.mixed_precision = 1 if data.mixed_precision else 0
        prm.inner_rel_tolerance = float(data.inner_rel_tolerance)
        prm.check_interval = int(data.check_interval)
        reps = (L.BatchReport * n)()
        st = self._lib.mpdg_heat_batch_solve(self._h, xp, bp, ctypes.byref(prm), reps)
        out = [{"status": r.status, "iterations": r.iterations, "outer_iterations": r.outer_iterations,
                "initial_residual": r.initial_residual, "final_residual": r.final_residual,
                "seconds": r.seconds} for r in reps]
        if st == L.MPDG_ERR_NOT_CONVERGED:
            raise NoConvergence(max(r["iterations"] for r in out), max(r["final_residual"] for r in out))
        L.check(st)
        return out
