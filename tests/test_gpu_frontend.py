"""The JSON front end on the device: every regression input of the reference's
unidirectional_heat_transfer case that needs only the built physics runs
through meltpooldg_amd.frontend (parameters -> case -> HeatTransferApplication
loop on the GPU) and prints the norms the reference printed (to the printed
digits, with the float-noise allowance explained in tests/test_frontend.py),
and each step's norm agrees with the oracle's time step through the same case
set-up within 1e-11 relative (Newton increments equal up to one, see below).
The command-line entry point runs one input in-process. A case set-up with
Neumann faces (a heat flux through x = x_min and y = y_min) runs through the
same loop and matches the oracle's time steps.
"""
import json
import os

import numpy as np
import pytest

from tests.test_frontend import CASES, SUPPORTED, matches_printed, oracle_problem

pytestmark = pytest.mark.gpu


def oracle_norms(prm, cs):
    from meltpooldg_amd import frontend as F
    from oracle.newton import NewtonData, linear_predictor, newton_time_step
    nl = prm.heat.nlsolve
    data = NewtonData(nl.max_nonlinear_iterations, nl.field_correction_tolerance,
                      nl.residual_tolerance, nl.max_nonlinear_iterations_alt,
                      nl.field_correction_tolerance_alt, nl.residual_tolerance_alt,
                      nl.use_relative_tolerances)
    ls = prm.heat.linear_solver
    it = F.TimeIterator(prm.time_stepping)
    pb = oracle_problem(cs, prm.time_stepping.time_step_size)
    T = cs.temperature0.copy()
    T_old = T.copy()
    out = []
    while not it.is_finished():
        dt = it.compute_next_time_increment()
        pb.set_dt(dt)
        if cs.heaviside_at is not None:
            pb.set_heaviside(cs.heaviside_at(it.current_time))
        T_old_old, T_old = T_old, T
        T, rep = newton_time_step(pb, linear_predictor(T_old, T_old_old, dt, it.old_increment),
                                  T_old, data=data, solver=ls.solver_type,
                                  rel_tol=ls.rel_tolerance, abs_tol=ls.abs_tolerance,
                                  max_it=ls.max_iterations)
        out.append((pb.l2_norm(T) * cs.norm_scale(), rep["iterations"]))
    return out


@pytest.mark.parametrize("name", SUPPORTED)
def test_reference_inputs_on_the_device(name):
    from meltpooldg_amd import frontend as F
    prm = F.parse_parameters(CASES[name]["input"])
    lines = []
    rec = F.run(prm, log=lines.append)
    ref = CASES[name]["l2_norms"]
    orc = oracle_norms(prm, F.setup(prm))
    print(name, [r["l2_norm"] for r in rec], ref, flush=True)
    assert len(rec) == len(ref) == len(orc)
    for r, printed, (o, o_it) in zip(rec, ref, orc):
        assert matches_printed(r["l2_norm"], printed), (name, r["step"], r["l2_norm"], printed)
        assert abs(r["l2_norm"] - o) <= 1e-11 * abs(o), (name, r["step"], r["l2_norm"], o)
        # after an increment the residual sits at its round-off floor, close to
        # the absolute tolerance (5e-9) of these inputs: the two summation
        # orders may stop one increment apart (the norms above agree to 1e-11)
        assert abs(r["newton"]["iterations"] - o_it) <= 1
    assert sum("Newton Raphson solver converged: ||solution|| = " in s for s in lines) == len(ref)
    ilu = prm.heat.linear_solver.preconditioner_type == "ILU"
    assert any("ILU runs as Diagonal" in s for s in lines) == ilu


@pytest.mark.parametrize("name", ["unidirectional_heat_transfer",
                                  "unidirectional_heat_transfer_with_two_phase",
                                  "unidirectional_heat_transfer_with_radiation_and_convection"])
def test_case_with_neumann_faces(name):
    from dataclasses import replace
    from meltpooldg_amd import frontend as F
    prm = F.parse_parameters(CASES[name]["input"])
    # x = x_min and y = y_min (the Robin faces of the with_* inputs are x = x_max)
    cs = replace(F.setup(prm), neumann_bc={0: (-2.0, 0.0, 0.0), 2: (0.0, 0.5, 0.0)})
    rec = F.run(prm, log=lambda s: None, max_steps=4, case=cs)
    orc = oracle_norms(prm, cs)[:len(rec)]
    print(name, "neumann", [r["l2_norm"] for r in rec], orc, flush=True)
    plain = F.run(prm, log=lambda s: None, max_steps=1)
    assert len(rec) == 4 and rec[0]["l2_norm"] != plain[0]["l2_norm"]
    for r, (o, o_it) in zip(rec, orc):
        assert abs(r["l2_norm"] - o) <= 1e-11 * abs(o), (name, r["step"], r["l2_norm"], o)
        assert abs(r["newton"]["iterations"] - o_it) <= 1


def test_command_line_entry_point(tmp_path, capsys):
    from meltpooldg_amd import frontend as F
    name = "unidirectional_heat_transfer_with_radiation"
    path = tmp_path / "input.json"
    path.write_text(json.dumps(CASES[name]["input"]))
    assert F.main([str(path), "--max-steps", "2"]) == 0
    out = capsys.readouterr().out
    norms = [float(l.split("= ")[-1]) for l in out.splitlines() if "||solution||" in l]
    assert norms == CASES[name]["l2_norms"][:2]
    bad = tmp_path / "cut.json"
    bad.write_text(json.dumps(CASES["unidirectional_heat_transfer_two_phase_cut"]["input"]))
    assert F.main([str(bad)]) == 2
