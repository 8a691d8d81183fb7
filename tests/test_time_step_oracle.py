"""The oracle's heat time step (linear predictor, Newton with create_rhs and
the Jacobi-GMRES Jacobian solve, L2 norm of the FE field) against the
reference's own regression output of the unidirectional_heat_transfer case.

The reference case is 2D (32 x 32 Q1 cells on [0, 0.1]^2, insulated, initial
T = x). Its solution does not depend on y; the 3D operator on the extruded
mesh of 32 x 1 x 1 cells on [0, 0.1] x [0, 0.1] x [0, 1] has the same
discrete solution (tensor-product mass and stiffness act as the identity on
fields constant in y and z), and with a z-extent of 1 the same L2 norm. The
norms are printed with 6 significant digits.
"""
import json
import os

import numpy as np

from oracle import cpu
from oracle.newton import NewtonData, linear_predictor, newton_time_step

HERE = os.path.dirname(os.path.abspath(__file__))


def unidirectional_problem():
    g = json.load(open(os.path.join(HERE, "golden", "unidirectional_heat_transfer.json")))
    s = g["setup"]
    n = 2 ** s["global_refinements"]
    pb = cpu.Problem((n, 1, 1), s["degree"], dt=s["time_step"], lo=(0.0, 0.0, 0.0),
                     hi=(0.1, 0.1, 1.0), dirichlet_faces=0)
    x = np.linspace(0.0, 0.1, n * s["degree"] + 1)
    T0 = np.tile(x, 4)  # nodal interpolation of T = x on the 33 x 2 x 2 lattice
    return g, pb, T0


def test_l2_norm_closed_forms():
    g, pb, T0 = unidirectional_problem()
    assert abs(pb.l2_norm(T0) - np.sqrt(0.1 * 0.1 ** 3 / 3.0)) < 1e-15
    pb2 = cpu.Problem((3, 2, 4), 3, lo=(0.0, -1.0, 0.5), hi=(2.0, 1.0, 1.5), dirichlet_faces=0)
    assert abs(pb2.l2_norm(np.full(pb2.N, 3.0)) - 3.0 * np.sqrt(4.0)) < 1e-13


def test_oracle_time_steps_match_reference_output():
    g, pb, T0 = unidirectional_problem()
    dt = g["setup"]["time_step"]
    data = NewtonData(residual_tolerance=g["setup"]["nlsolve"]["residual_tolerance"])
    T_old, T_old_old, dt_old = T0, T0, 0.0
    for step, ref in enumerate(g["l2_norms"]):
        pred = linear_predictor(T_old, T_old_old, dt, dt_old)
        T, rep = newton_time_step(pb, pred, T_old, data=data, solver="GMRES", rel_tol=1e-12)
        assert rep["converged"]
        norm = pb.l2_norm(T)
        assert float(f"{norm:.5e}") == ref, (step, norm, ref)
        T_old_old, T_old, dt_old = T_old, T, dt


def test_linear_predictor():
    a, b = np.array([1.0, 2.0]), np.array([0.5, 1.0])
    assert np.allclose(linear_predictor(a, b, 2.0, 1.0), 3.0 * a - 2.0 * b)
    assert np.array_equal(linear_predictor(a, b, 1.0, 0.0), a)
