"""JSON front end: the reference's parameter files drive the device heat path.

* ``parse_parameters(source)`` reads a MeltPoolDG parameter file (JSON, the
  format of applications/mp-heat-transfer/cases/*/tests/*.json) into the
  reference's parameter structures, restated with their names and defaults:
  BaseData (source/core/base_data.cpp), TimeSteppingData
  (source/time_integration/time_stepping_data.cpp), HeatData
  (source/heat/heat_data.cpp:8-85) with LinearSolverData
  (source/linear_algebra/linear_solver_data.cpp:9-36), NonlinearSolverData
  (source/linear_algebra/nonlinear_solver_data.cpp) and PredictorData
  (source/linear_algebra/predictor_data.cpp), MaterialData
  (source/core/material_data.cpp), the application-specific and the
  case-specific parameters. An unknown key in a modelled section is an error,
  as in the reference's ParameterHandler; settings that ask for physics this
  engine does not run (cut operator, evaporation, laser, AMR, time-dependent
  boundary conditions, least-squares predictor) raise ``UnsupportedFeature``
  naming the key. The
  ILU preconditioner, which needs an assembled sparse matrix, runs as the
  Jacobi (Diagonal) preconditioner: the Newton solution does not depend on the
  preconditioner beyond the linear tolerance, only the Krylov iteration counts
  do, and ``run`` says so in its log.
* ``setup(parameters)`` builds the simulation case (the reference's
  SimulationCase subclasses): ``unidirectional_heat_transfer``
  (applications/mp-heat-transfer/cases/unidirectional_heat_transfer/
  unidirectional_heat_transfer.cpp) in 1D and 2D, extruded to the engine's 3D
  box (below), with the material type of determine_material_type
  (source/core/material.cpp:7-39): gas, gas_liquid, liquid_solid or
  gas_liquid_solid.
* ``run(parameters)`` is HeatTransferApplication::run
  (applications/mp-heat-transfer/heat_transfer_application.cpp:44-135) on the
  device: the time iterator (source/time_integration/time_iterator.cpp:26-67),
  the prescribed fields (a time-dependent heaviside re-interpolated at each
  step's new time, heat_transfer_application.cpp:67-70, 470-480), and
  HeatDiffuseOperation::solve per step, printing
  the time increment and "Newton Raphson solver converged: ||solution|| = ..."
  lines as the reference does.

Extrusion. The engine is 3D. A 2D case on n x n cells runs on n x n x 1 cells
and a 1D case on n x 1 x 1; fields constant along the added directions solve
the 3D problem exactly (tensor-product mass, stiffness and face terms), so the
discrete solution is the lower-dimensional one. Each added direction has the
depth 1 / sqrt(sum_i w_i^2) (w_i: Gauss-Lobatto weights of the FE degree on
[0, 1]), which makes the 3D residual vector's Euclidean norm equal to the
lower-dimensional one, so Newton's absolute residual tolerance acts as in the
reference. Norms are reported for the lower-dimensional field (the 3D L2 norm
divided by sqrt(depth) per added direction).

Usage: python -m meltpooldg_amd.frontend parameters.json
"""
import argparse
import json
import math
import sys
from dataclasses import dataclass, field, fields, is_dataclass

import numpy as np

from .heat import ConvectionBC, LinearSolverData, NonlinearSolverData, RadiationBC


class UnsupportedFeature(ValueError):
    """A parameter asks for physics or numerics this engine does not run."""


class UnknownParameter(KeyError):
    """A key the reference's parameter set for this application does not have."""


# ---- parameter structures (names and defaults of the reference) -----------

@dataclass
class FEData:
    type: str = "FE_Q"
    degree: int = 1


@dataclass
class BaseParameters:
    case_name: str = "not_initialized"
    dimension: int = 2
    number: str = "double"
    global_refinements: int = 1
    do_print_parameters: bool = True
    verbosity_level: int = 1
    fe: FEData = field(default_factory=FEData)


@dataclass
class TimeSteppingParameters:
    start_time: float = 0.0
    end_time: float = 1.0
    time_step_size: float = 0.01
    max_n_steps: int = 10000000
    time_step_size_function: str = "0.0*t"


@dataclass
class DiffuseParameters:
    use_volume_specific_thermal_capacity_for_phase_interpolation: bool = False


@dataclass
class PredictorParameters:
    type: str = "linear_extrapolation"  # HeatData's default (heat_data.cpp:13)
    n_old_solutions: int = 2


@dataclass
class HeatParameters:
    fe: FEData = field(default_factory=lambda: FEData("not_initialized", -1))
    operator_type: str = "diffuse"
    cut: dict = field(default_factory=dict)  # CutFEM settings, kept only to name the operator
    enable_time_dependent_bc: bool = False
    diffuse: DiffuseParameters = field(default_factory=DiffuseParameters)
    radiation: RadiationBC = field(default_factory=RadiationBC)
    convection: ConvectionBC = field(default_factory=ConvectionBC)
    verbosity_level: int = -1
    nlsolve: NonlinearSolverData = field(default_factory=NonlinearSolverData)
    # HeatData's solver defaults (heat_data.cpp:11-12)
    linear_solver: LinearSolverData = field(
        default_factory=lambda: LinearSolverData(solver_type="GMRES", preconditioner_type="Diagonal"))
    predictor: PredictorParameters = field(default_factory=PredictorParameters)


@dataclass
class MaterialPhaseParameters:
    thermal_conductivity: float = 0.0
    specific_heat_capacity: float = 0.0
    density: float = 0.0
    dynamic_viscosity: float = 0.0


@dataclass
class MaterialParameters:
    material_template: str = "none"
    gas: MaterialPhaseParameters = field(default_factory=MaterialPhaseParameters)
    liquid: MaterialPhaseParameters = field(default_factory=MaterialPhaseParameters)
    solid: MaterialPhaseParameters = field(default_factory=MaterialPhaseParameters)
    solidus_temperature: float = 0.0
    liquidus_temperature: float = 0.0
    apparent_capacity_type: str = "qlq"
    latent_heat_of_fusion: float = 0.0
    boiling_temperature: float = 0.0
    latent_heat_of_evaporation: float = 0.0
    molar_mass: float = 0.0
    specific_enthalpy_reference_temperature: float = -1e100
    two_phase_fluid_properties_transition_type: str = "smooth"
    solid_liquid_properties_transition_type: str = "mushy_zone"


@dataclass
class ApplicationSpecificParameters:
    do_solidification: bool = False
    amr_strategy: str = "KellyErrorEstimator"


@dataclass
class Parameters:
    base: BaseParameters = field(default_factory=BaseParameters)
    time_stepping: TimeSteppingParameters = field(default_factory=TimeSteppingParameters)
    heat: HeatParameters = field(default_factory=HeatParameters)
    material: MaterialParameters = field(default_factory=MaterialParameters)
    application_specific: ApplicationSpecificParameters = field(
        default_factory=ApplicationSpecificParameters)
    case_specific: dict = field(default_factory=dict)


# JSON section / key -> attribute where the attribute name is not the key with
# spaces and hyphens turned into underscores
_RENAME = {
    "time stepping": "time_stepping",
    "radiative boundary condition": "radiation",
    "convective boundary condition": "convection",
    "linear solver": "linear_solver",
    "application specific": "application_specific",
    "n old solutions": "n_old_solutions",
}
# sections that only configure output or diagnostics
_IGNORED = {"output", "profiling"}
# sections of other physics: accepted only while that physics stays off, i.e.
# no "enable" / "do amr" flag in them is true and the laser has no power
_OFF_SECTIONS = {"adaptive meshing", "laser", "evaporation", "rte"}


def _attr(key):
    return _RENAME.get(key, key.replace(" ", "_").replace("-", "_"))


def _convert(value, current, path):
    if isinstance(current, bool):
        if isinstance(value, str):
            if value.lower() not in ("true", "false"):
                raise ValueError(f"{path}: expected a boolean, got {value!r}")
            return value.lower() == "true"
        return bool(value)
    if isinstance(current, int):
        return int(value)
    if isinstance(current, float):
        return float(value)
    return str(value)


def _assign(obj, data, path):
    for key, value in data.items():
        name = _attr(key)
        names = {f.name for f in fields(obj)}
        if name not in names:
            raise UnknownParameter(f"{'/'.join(path + (key,))}")
        current = getattr(obj, name)
        if isinstance(current, dict):
            setattr(obj, name, dict(value))
        elif is_dataclass(current):
            if not isinstance(value, dict):
                raise ValueError(f"{'/'.join(path + (key,))}: expected a section")
            _assign(current, value, path + (key,))
        else:
            setattr(obj, name, _convert(value, current, "/".join(path + (key,))))


def _switched_on(section, path):
    """First flag path in a section that turns its physics on, or None."""
    for key, value in section.items():
        if isinstance(value, dict):
            hit = _switched_on(value, path + (key,))
            if hit:
                return hit
        elif key in ("enable", "do amr") and str(value).lower() == "true":
            return "/".join(path + (key,))
        elif key == "power" and float(value) > 0.0:
            return "/".join(path + (key,))
    return None


def _check_off_section(key, value):
    hit = _switched_on(value, (key,)) if isinstance(value, dict) else None
    if hit:
        raise UnsupportedFeature(f"{hit}: {key} is not part of the built heat path")


def parse_parameters(source):
    """Parameters from a JSON file path, JSON text or an already loaded dict."""
    if isinstance(source, dict):
        data = dict(source)
    elif isinstance(source, str) and source.lstrip().startswith("{"):
        data = json.loads(source)
    else:
        with open(source) as f:
            data = json.load(f)
    data.pop("$schema", None)
    prm = Parameters()
    for key, value in data.items():
        if key in _IGNORED:
            continue
        if key in _OFF_SECTIONS:
            _check_off_section(key, value)
            continue
        if key == "case specific":
            prm.case_specific = dict(value)
            continue
        name = _attr(key)
        if name not in {f.name for f in fields(prm)}:
            raise UnknownParameter(key)
        _assign(getattr(prm, name), value, (key,))
    check_supported(prm)
    return prm


def check_supported(prm):
    h = prm.heat
    if h.operator_type != "diffuse":
        raise UnsupportedFeature(f"heat/operator type = {h.operator_type!r}: only the diffuse "
                                 "operator is built")
    if h.enable_time_dependent_bc:
        raise UnsupportedFeature("heat/enable time dependent bc: not built")
    if prm.application_specific.do_solidification:
        m = prm.material
        if m.latent_heat_of_fusion > 0.0:
            if m.apparent_capacity_type not in ("poly4_bell", "constant", "qlq"):
                raise UnsupportedFeature(f"material/apparent capacity type = "
                                         f"{m.apparent_capacity_type!r}")
            if prm.heat.diffuse.use_volume_specific_thermal_capacity_for_phase_interpolation:
                raise UnsupportedFeature("material/latent heat of fusion with the volume-specific "
                                         "heat capacity: the reference's apparent capacity is not "
                                         "implemented for it")
        if m.solid_liquid_properties_transition_type not in ("mushy_zone", "sharp"):
            raise UnsupportedFeature("material/solid liquid properties transition type = "
                                     f"{m.solid_liquid_properties_transition_type!r}")
        if (m.solid_liquid_properties_transition_type == "mushy_zone" and
                not m.liquidus_temperature > m.solidus_temperature):
            raise UnsupportedFeature("material: the mushy zone needs liquidus temperature > "
                                     "solidus temperature")
    if h.predictor.type not in ("none", "zero", "linear_extrapolation"):
        raise UnsupportedFeature(f"heat/predictor/type = {h.predictor.type!r}: not built")
    ls = h.linear_solver
    if ls.solver_type not in ("CG", "GMRES"):
        raise UnsupportedFeature(f"heat/linear solver/solver type = {ls.solver_type!r}")
    if ls.preconditioner_type not in ("Identity", "Diagonal", "ILU"):
        raise UnsupportedFeature(f"heat/linear solver/preconditioner type = "
                                 f"{ls.preconditioner_type!r}: Identity and Diagonal are built "
                                 "(ILU runs as Diagonal)")
    if not ls.do_matrix_free:
        raise UnsupportedFeature("heat/linear solver/do matrix free = false: the engine is "
                                 "matrix-free")
    if prm.material.material_template != "none":
        raise UnsupportedFeature(f"material/material template = "
                                 f"{prm.material.material_template!r}: not built")
    if prm.material.two_phase_fluid_properties_transition_type not in ("smooth", "sharp"):
        raise UnsupportedFeature("material/two phase fluid properties transition type = "
                                 f"{prm.material.two_phase_fluid_properties_transition_type!r}")
    f = prm.time_stepping.time_step_size_function.replace(" ", "")
    if f not in ("0.0*t", "0*t", "0", "0.0"):
        raise UnsupportedFeature("time stepping/time step size function: only constant time "
                                 "steps are built")
    if prm.base.number != "double":
        raise UnsupportedFeature("base/number: the parameter files run in double")


# ---- the simulation case ---------------------------------------------------

def gll_weights(p):
    """Gauss-Lobatto weights on [0, 1] of the FE_Q(p) support points."""
    from numpy.polynomial import legendre as Lg
    inner = np.sort(Lg.Legendre.basis(p).deriv().roots().real) if p > 1 else np.array([])
    z = np.concatenate([[-1.0], inner, [1.0]])
    w = 2.0 / (p * (p + 1) * Lg.legval(z, [0] * p + [1]) ** 2)
    return 0.5 * w, 0.5 * (z + 1.0)


def smoothed_heaviside(distance, eps):
    """CharacteristicFunctions::smoothed_heaviside
    (include/meltpooldg/utilities/characteristic_functions.hpp:17-28)."""
    d = np.asarray(distance, dtype=float)
    return np.where(d > eps, 1.0, np.where(d <= -eps, 0.0,
                    (d + eps) / (2.0 * eps) + np.sin(np.pi * d / eps) / (2.0 * np.pi)))


@dataclass
class CaseSetup:
    """What a simulation case hands the heat operation, on the engine's 3D box."""
    dimension: int
    n_cells: tuple
    degree: int
    box: tuple
    extrusion: float           # depth of each added direction
    material: object           # meltpooldg_amd.MaterialData
    two_phase: bool
    heaviside: object          # lattice vector or None
    velocity: object           # (vx, vy, vz) or None
    convection: ConvectionBC
    radiation: RadiationBC
    convection_faces: int
    radiation_faces: int
    dirichlet_faces: int
    temperature0: object       # lattice vector
    coords: tuple              # lattice coordinates per direction
    heaviside_at: object = None  # t -> lattice vector, for a time-dependent heaviside
    # {face id: (qx, qy, qz)}: a constant heat flux per face, the Neumann
    # conditions a case attaches (heat_diffuse_operator.cpp:78, 735-767)
    neumann_bc: dict = None

    def norm_scale(self):
        """Factor from the 3D L2 norm to the case's own dimension."""
        return 1.0 / math.sqrt(self.extrusion ** (3 - self.dimension))


def _lattice_coords(n_cells, degree, box):
    _, gll = gll_weights(degree)
    out = []
    for d in range(3):
        lo, hi = box[0][d], box[1][d]
        h = (hi - lo) / n_cells[d]
        c = [lo + (i + gll[:-1]) * h for i in range(n_cells[d])]
        out.append(np.concatenate(c + [[hi]]))
    return tuple(out)


def _unidirectional_heat_transfer(prm):
    """SimulationUnidirectionalHeatTransfer (unidirectional_heat_transfer.cpp):
    the box [0, 0.1]^dim, refined globally; Robin conditions on the right face
    (dim 1: the lower one); T0 linear from 1960 to 1980 in x with
    solidification, else T0 = 1000 with a Robin face, else T0 = x; the
    prescribed velocity (-v, 0) if v != 0; for two phases the heaviside of the
    horizontal level set 0.05 - y, or with solidification of the covected
    vertical one 0.1 * 2/3 - x - v t (time-dependent), both with eps = 0.01.
    The material type follows determine_material_type
    (source/core/material.cpp:7-39) from the heaviside's presence and the
    application's do solidification."""
    from .heat import MaterialData, PhaseData
    cs = {"do solidification": False, "do two phase": False, "velocity": 0.0}
    for k, v in prm.case_specific.items():
        if k not in cs:
            raise UnknownParameter(f"case specific/{k}")
        cs[k] = v
    dim = prm.base.dimension
    if dim not in (1, 2):
        raise UnsupportedFeature(f"base/dimension = {dim}: the case is defined in 1D and 2D")
    x_min, x_max = 0.0, 0.1
    degree = prm.heat.fe.degree if prm.heat.fe.degree > 0 else prm.base.fe.degree
    if prm.heat.fe.degree > 0 and prm.heat.fe.degree != prm.base.fe.degree and cs["do two phase"]:
        raise UnsupportedFeature("heat/fe/degree different from base/fe/degree with two phases: "
                                 "the heaviside lives on the base FE")
    n = 2 ** prm.base.global_refinements
    w, _ = gll_weights(degree)
    depth = 1.0 / math.sqrt(float(np.sum(w ** 2)))
    if dim == 2:
        n_cells = (n, n, 1)
        box = ((x_min, 0.0, 0.0), (x_max, 0.1, depth))
    else:
        n_cells = (n, 1, 1)
        box = ((x_min, 0.0, 0.0), (x_max, depth, depth))
    # get_colorized_rectangle_boundary_ids: right = 1 (2D), lower = 0 (1D)
    robin_face = 1 << (0 if dim == 1 else 1)
    rad, conv = prm.heat.radiation, prm.heat.convection
    rad_faces = robin_face if rad.emissivity > 0.0 else 0
    conv_faces = robin_face if conv.convection_coefficient > 0.0 else 0
    coords = _lattice_coords(n_cells, degree, box)
    Z, Y, X = np.meshgrid(coords[2], coords[1], coords[0], indexing="ij")
    if cs["do solidification"]:
        # LinearTemp(1960, 1980)
        T0 = (1960.0 + (1980.0 - 1960.0) * (X - x_min) / (x_max - x_min)).reshape(-1)
    elif rad.emissivity > 0.0 or conv.convection_coefficient > 0.0:
        T0 = np.full(X.size, 1000.0)
    else:
        # LinearTemp(0, 0.1): left + (right - left) (x - x_min) / (x_max - x_min)
        T0 = (0.0 + (0.1 - 0.0) * (X - x_min) / (x_max - x_min)).reshape(-1)
    v = float(cs["velocity"])
    velocity = (-v, 0.0, 0.0) if v != 0.0 else None
    two_phase = bool(cs["do two phase"])
    heaviside = None
    heaviside_at = None
    if two_phase:
        if cs["do solidification"]:
            # CovectedVerticalLevelSetHeaviside(v): level 0.1 * 2/3, moving with v t
            Xf = X.reshape(-1)

            def heaviside_at(t):
                return smoothed_heaviside(x_max * 2.0 / 3.0 - Xf - v * t, 0.01)

            heaviside = heaviside_at(prm.time_stepping.start_time)
        else:
            if dim != 2:
                raise UnsupportedFeature("do two phase in 1D: the horizontal level set needs y")
            heaviside = smoothed_heaviside(x_max / 2 - Y, 0.01).reshape(-1)
    solid = bool(prm.application_specific.do_solidification)
    m = prm.material
    phase = lambda ph: PhaseData(ph.density, ph.specific_heat_capacity, ph.thermal_conductivity)
    mtype = {(False, False): "gas", (True, False): "gas_liquid", (False, True): "liquid_solid",
             (True, True): "gas_liquid_solid"}[(two_phase, solid)]
    material = MaterialData(
        type=mtype, gas=phase(m.gas), liquid=phase(m.liquid), solid=phase(m.solid),
        two_phase_fluid_properties_transition_type=m.two_phase_fluid_properties_transition_type,
        use_volume_specific_thermal_capacity_for_phase_interpolation=(
            prm.heat.diffuse.use_volume_specific_thermal_capacity_for_phase_interpolation),
        solid_liquid_properties_transition_type=m.solid_liquid_properties_transition_type,
        solidus_temperature=m.solidus_temperature, liquidus_temperature=m.liquidus_temperature,
        latent_heat_of_fusion=m.latent_heat_of_fusion,
        apparent_capacity_type=m.apparent_capacity_type)
    return CaseSetup(dimension=dim, n_cells=n_cells, degree=degree, box=box, extrusion=depth,
                     material=material, two_phase=two_phase, heaviside=heaviside,
                     velocity=velocity, convection=conv, radiation=rad,
                     convection_faces=conv_faces, radiation_faces=rad_faces, dirichlet_faces=0,
                     temperature0=T0, coords=coords, heaviside_at=heaviside_at)


CASES = {"unidirectional_heat_transfer": _unidirectional_heat_transfer}


def setup(prm):
    case = prm.base.case_name
    if case not in CASES:
        raise UnsupportedFeature(f"base/case name = {case!r}: built cases are {sorted(CASES)}")
    return CASES[case](prm)


# ---- the application loop --------------------------------------------------

class TimeIterator:
    """TimeIterator (source/time_integration/time_iterator.cpp:26-67) for a
    constant time step: is_finished, compute_next_time_increment (the last
    increment is cut to the end time), the old increment as the predictor
    reads it."""

    def __init__(self, ts):
        self.data = ts
        self.current_time = ts.start_time
        self.increment = ts.time_step_size
        self.old_increment = 0.0
        self.n_steps = 0

    def is_finished(self):
        if self.n_steps >= self.data.max_n_steps:
            return True
        if self.current_time > self.data.end_time:
            return True
        return abs(self.current_time - self.data.end_time) <= np.finfo(float).eps

    def compute_next_time_increment(self):
        self.n_steps += 1
        if self.current_time + self.increment > self.data.end_time:
            self.increment = self.data.end_time - self.current_time
        if self.n_steps > 1:
            self.old_increment = self.increment
        self.current_time += self.increment
        return self.increment


def make_operator(case, stream=None):
    """The device heat operator of a case set-up."""
    from .heat import HeatDiffuseOperator
    conv_ids = [f for f in range(6) if (case.convection_faces >> f) & 1]
    rad_ids = [f for f in range(6) if (case.radiation_faces >> f) & 1]
    op = HeatDiffuseOperator(case.n_cells, case.degree, material=case.material, box=case.box,
                             velocity="constant" if case.velocity else None,
                             velocity_value=case.velocity, dirichlet_faces=case.dirichlet_faces,
                             stream=stream, convection=case.convection, radiation=case.radiation,
                             convection_boundary_ids=conv_ids, radiation_boundary_ids=rad_ids,
                             neumann_bc=case.neumann_bc)
    return op


def run(prm, log=print, max_steps=None, graphs=True, case=None):
    """HeatTransferApplication::run on the device. Returns one record per time
    step: step, time, dt, the L2 norm of the solution (in the case's own
    dimension) and the Newton report. ``graphs``: replay the Krylov restart
    cycles / iteration chunks as captured hipGraphs (False: eager launches).
    ``case``: a CaseSetup to run instead of ``setup(prm)`` (a case with its
    own boundary conditions, e.g. Neumann faces)."""
    import torch
    from .heat import HeatDiffuseOperation
    if not isinstance(prm, Parameters):
        prm = parse_parameters(prm)
    if case is None:
        case = setup(prm)
    op = make_operator(case)
    op.set_graphs(graphs)
    if case.heaviside is not None:
        op.set_heaviside(torch.tensor(case.heaviside, dtype=torch.float64, device="cuda"))
    T0 = torch.tensor(case.temperature0, dtype=torch.float64, device="cuda")
    ls = prm.heat.linear_solver
    if ls.preconditioner_type == "ILU":
        from dataclasses import replace
        log("note: heat/linear solver/preconditioner type ILU runs as Diagonal (Jacobi)")
        ls = replace(ls, preconditioner_type="Diagonal")
    heat = HeatDiffuseOperation(op, T0, linear_solver=ls, nlsolve=prm.heat.nlsolve,
                                predictor=prm.heat.predictor.type)
    it = TimeIterator(prm.time_stepping)
    out = []
    while not it.is_finished():
        if max_steps is not None and len(out) >= max_steps:
            break
        t0 = it.current_time
        dt = it.compute_next_time_increment()
        log(f"| Time increment # {it.n_steps:<6d} t = {t0:.4e} to {it.current_time:.4e} "
            f"( dt = {dt:.4e} )")
        if case.heaviside_at is not None:
            op.set_heaviside(torch.tensor(case.heaviside_at(it.current_time), dtype=torch.float64,
                                          device="cuda"))
        rep = heat.advance(dt, it.old_increment)
        norm = heat.l2_norm() * case.norm_scale()
        log(f"| Newton Raphson solver converged: ||solution|| = {norm:.5e}")
        out.append({"step": it.n_steps, "time": it.current_time, "dt": dt, "l2_norm": norm,
                    "newton": rep})
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("parameters", help="MeltPoolDG parameter file (JSON)")
    ap.add_argument("--max-steps", type=int, default=None)
    args = ap.parse_args(argv)
    try:
        prm = parse_parameters(args.parameters)
    except (UnsupportedFeature, UnknownParameter) as e:
        print(f"error: {e}", file=sys.stderr)
        return 2
    run(prm, max_steps=args.max_steps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
