"""meltpooldg_amd: MI355X-native matrix-free heat-operator engine.

The compute path is libmpdg_heat.so (HIP kernels for gfx950 behind the C ABI
of include/mpdg_heat.h). This package is the host-side mirror of MeltPoolDG's
heat operator / Jacobi / LinearSolver interface on top of it.
"""
from .heat import (ConvectionBC, HeatBatch, HeatDiffuseOperation, HeatDiffuseOperator, JacobiPreconditioner,
                   LinearSolver, LinearSolverData, MaterialData, NewtonDidNotConverge, NoConvergence,
                   NonlinearSolverData, PhaseData, RadiationBC, ZeroTimeIncrement,
                   lattice_index_of_points)

__all__ = ["ConvectionBC", "HeatBatch", "HeatDiffuseOperation", "HeatDiffuseOperator", "JacobiPreconditioner",
           "LinearSolver", "LinearSolverData", "MaterialData", "NewtonDidNotConverge", "NoConvergence",
           "NonlinearSolverData", "PhaseData", "RadiationBC", "ZeroTimeIncrement",
           "lattice_index_of_points"]
