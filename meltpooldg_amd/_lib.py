"""ctypes binding of the C ABI in include/mpdg_heat.h (libmpdg_heat.so).

The library is the only compute path: if it is missing or fails to load the
import raises, there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpdg_heat.so")

MPDG_OK = 0
MPDG_ERR_ARGUMENT = 1
MPDG_ERR_HIP = 2
MPDG_ERR_NOT_CONVERGED = 3
MPDG_ERR_UNSUPPORTED = 4
MPDG_ERR_ZERO_TIME_INCREMENT = 5

MPDG_FP64 = 0
MPDG_FP32 = 1


class Phase(ctypes.Structure):
    _fields_ = [("density", ctypes.c_double), ("specific_heat_capacity", ctypes.c_double),
                ("thermal_conductivity", ctypes.c_double)]


class HeatDesc(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int * 3),
        ("degree", ctypes.c_int),
        ("box_lo", ctypes.c_double * 3),
        ("box_hi", ctypes.c_double * 3),
        ("precision", ctypes.c_int),
        ("material", ctypes.c_int),
        ("gas", Phase),
        ("liquid", Phase),
        ("two_phase_sharp", ctypes.c_int),
        ("volume_specific_heat", ctypes.c_int),
        ("solid", Phase),
        ("solidus_temperature", ctypes.c_double),
        ("liquidus_temperature", ctypes.c_double),
        ("solid_liquid_sharp", ctypes.c_int),
        ("latent_heat_of_fusion", ctypes.c_double),
        ("apparent_capacity_type", ctypes.c_int),
        ("conductivity_field", ctypes.c_int),
        ("conductivity_amplitude", ctypes.c_double),
        ("velocity", ctypes.c_int),
        ("velocity_w0", ctypes.c_double),
        ("velocity_value", ctypes.c_double * 3),
        ("dirichlet_faces", ctypes.c_uint),
        ("convection_faces", ctypes.c_uint),
        ("convection_coefficient", ctypes.c_double),
        ("convection_temperature_infinity", ctypes.c_double),
        ("radiation_faces", ctypes.c_uint),
        ("emissivity", ctypes.c_double),
        ("radiation_temperature_infinity", ctypes.c_double),
        ("stream", ctypes.c_void_p),
        ("neumann_faces", ctypes.c_uint),
        ("neumann_flux", (ctypes.c_double * 3) * 6),
    ]


class SolverParams(ctypes.Structure):
    _fields_ = [
        ("solver_type", ctypes.c_int),
        ("preconditioner_type", ctypes.c_int),
        ("max_iterations", ctypes.c_int),
        ("rel_tolerance", ctypes.c_double),
        ("abs_tolerance", ctypes.c_double),
        ("gmres_restart", ctypes.c_int),
        ("mixed_precision", ctypes.c_int),
        ("inner_rel_tolerance", ctypes.c_double),
        ("check_interval", ctypes.c_int),
    ]


class SolverReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("outer_iterations", ctypes.c_int),
        ("initial_residual", ctypes.c_double),
        ("final_residual", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]


class BatchReport(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("outer_iterations", ctypes.c_int),
        ("initial_residual", ctypes.c_double),
        ("final_residual", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]


class NewtonParams(ctypes.Structure):
    _fields_ = [
        ("max_nonlinear_iterations", ctypes.c_int),
        ("field_correction_tolerance", ctypes.c_double),
        ("residual_tolerance", ctypes.c_double),
        ("max_nonlinear_iterations_alt", ctypes.c_int),
        ("field_correction_tolerance_alt", ctypes.c_double),
        ("residual_tolerance_alt", ctypes.c_double),
        ("use_relative_tolerances", ctypes.c_int),
    ]


class NewtonReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("linear_iterations", ctypes.c_int),
        ("initial_residual", ctypes.c_double),
        ("residual", ctypes.c_double),
        ("update", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]


MPDG_VMULT_LINE = 0
MPDG_VMULT_CELL = 1
MPDG_VMULT_LINE_MFMA = 2
MPDG_VMULT_CELL_MFMA = 3


class Tuning(ctypes.Structure):
    _fields_ = [("z_layers", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("variant", ctypes.c_int), ("tile_log2", ctypes.c_int),
                ("schedule", ctypes.c_int)]


# name -> (restype, argtypes); exactly the entry points include/mpdg_heat.h declares
_VP = ctypes.c_void_p
SIGNATURES = {
    "mpdg_heat_create": (ctypes.c_int, [ctypes.POINTER(HeatDesc), ctypes.POINTER(_VP)]),
    "mpdg_heat_destroy": (ctypes.c_int, [_VP]),
    "mpdg_heat_n_dofs": (ctypes.c_int64, [_VP]),
    "mpdg_heat_n_free_dofs": (ctypes.c_int64, [_VP]),
    "mpdg_heat_set_dt": (ctypes.c_int, [_VP, ctypes.c_double]),
    "mpdg_heat_set_heaviside": (ctypes.c_int, [_VP, _VP]),
    "mpdg_heat_set_heaviside_iso_q1": (ctypes.c_int, [_VP, _VP]),
    "mpdg_heat_set_velocity": (ctypes.c_int, [_VP, _VP]),
    "mpdg_heat_set_temperature": (ctypes.c_int, [_VP, _VP]),
    "mpdg_heat_set_temperature_old": (ctypes.c_int, [_VP, _VP]),
    "mpdg_heat_set_graphs": (ctypes.c_int, [_VP, ctypes.c_int]),
    "mpdg_lattice_index_of_points": (ctypes.c_int, [_VP, _VP, ctypes.c_int64, _VP]),
    "mpdg_heat_to_lattice": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "mpdg_heat_from_lattice": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "mpdg_heat_vmult": (ctypes.c_int, [_VP, _VP, _VP]),
    "mpdg_heat_vmult_f32": (ctypes.c_int, [_VP, _VP, _VP]),
    "mpdg_heat_inverse_diagonal": (ctypes.c_int, [_VP, _VP]),
    "mpdg_heat_create_rhs": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "mpdg_newton_params_default": (None, [ctypes.POINTER(NewtonParams)]),
    "mpdg_heat_newton_solve": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.POINTER(NewtonParams),
                                              ctypes.POINTER(SolverParams),
                                              ctypes.POINTER(NewtonReport)]),
    "mpdg_heat_linear_predictor": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_double,
                                                  ctypes.c_double]),
    "mpdg_heat_l2_norm": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_double)]),
    "mpdg_heat_distribute_constraints": (ctypes.c_int, [_VP, _VP]),
    "mpdg_heat_l2_norm_noalloc": (ctypes.c_int, [_VP, _VP, _VP]),
    "mpdg_heat_jacobi_vmult": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "mpdg_heat_solve": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(SolverParams),
                                       ctypes.POINTER(SolverReport)]),
    "mpdg_heat_solve_preconditioned": (ctypes.c_int, [_VP, _VP, _VP, _VP,
                                                      ctypes.POINTER(SolverParams),
                                                      ctypes.POINTER(SolverReport)]),
    "mpdg_heat_set_residual_history": (ctypes.c_int, [_VP, _VP, ctypes.c_int]),
    "mpdg_solver_params_default": (None, [ctypes.POINTER(SolverParams)]),
    "mpdg_last_error": (ctypes.c_char_p, []),
    "mpdg_heat_batch_create": (ctypes.c_int, [ctypes.POINTER(HeatDesc), ctypes.c_int,
                                              ctypes.POINTER(_VP)]),
    "mpdg_heat_batch_destroy": (ctypes.c_int, [_VP]),
    "mpdg_heat_batch_set_dt": (ctypes.c_int, [_VP, ctypes.c_double]),
    "mpdg_heat_batch_n_dofs": (ctypes.c_int64, [_VP, ctypes.c_int]),
    "mpdg_heat_batch_solve": (ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                                             ctypes.POINTER(SolverParams), ctypes.POINTER(BatchReport)]),
    "mpdg_heat_get_tuning": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(Tuning)]),
    "mpdg_heat_set_tuning": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(Tuning)]),
    "mpdg_heat_autotune": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Tuning)]),
    "mpdg_heat_time_vmult": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


def load():
    """Load libmpdg_heat.so (built by meltpooldg_amd.build). Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m meltpooldg_amd.build` "
                          "(the HIP library is the only compute path)")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and refers
    # to it as libamdhip64.so, so if this library loaded /opt/rocm's copy first,
    # torch would load a second runtime (and the first one would see no device).
    # Loading torch first makes this library bind to the runtime already present.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MpdgError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"mpdg status {status}: {message}")
        self.status = status


def check(status):
    if status != MPDG_OK:
        raise MpdgError(status, load().mpdg_last_error().decode())
