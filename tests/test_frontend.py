"""The JSON front end (meltpooldg_amd.frontend) on the CPU.

* Every regression input of the reference's unidirectional_heat_transfer case
  (tests/golden/unidirectional_cases.json, written by
  tests/golden/make_unidirectional_cases.py from the reference's own test JSONs
  and outputs) parses; the nine that need only the built physics set up
  (two of them with solidification, the covected time-dependent heaviside
  and the ILU preconditioner, which runs as Diagonal), the eight that need
  the cut operator, evaporation or AMR are rejected with UnsupportedFeature
  naming the key.
* The parameter structures' defaults equal the ones
  applications/mp-heat-transfer/parameters.schema.json lists (kept in the same
  fixture), except where the reference's own constructors override them
  (HeatData: GMRES + Diagonal + linear extrapolation, heat_data.cpp:8-14).
* Unknown keys are errors, as in the reference's ParameterHandler.
* The time iterator follows time_iterator.cpp (end-time cut, old increment).
* Through the front end's case set-up (extrusion to 3D included), the oracle's
  time step reproduces the norms the reference printed for all seven supported
  inputs, to the printed 6 digits. The reference's printed norm carries its
  own float rounding: VectorTools::compute_norm stores each cell's L2 norm in
  a Vector<float> (source/utilities/vector_tools.cpp:151) and sums their
  squares from there, which moves a norm of ~100 by about 1e-7 (measured by
  rounding this solution's cell norms to float) up to a few 1e-6 (a plain
  float sum). So a value within 1e-8 relative of a rounding boundary of the
  printed digits may print either way; everywhere else the digits must match.
"""
import json
import os

import numpy as np
import pytest

from meltpooldg_amd import frontend as F

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = json.load(open(os.path.join(HERE, "golden", "unidirectional_cases.json")))
CASES = FIXTURE["cases"]
SUPPORTED = ["unidirectional_heat_transfer", "unidirectional_heat_transfer_with_convection",
             "unidirectional_heat_transfer_with_radiation",
             "unidirectional_heat_transfer_with_radiation_and_convection",
             "unidirectional_heat_transfer_with_two_phase",
             "unidirectional_heat_transfer_with_two_phase_rho_times_cp",
             "unidirectional_heat_transfer_with_velocity",
             "unidirectional_heat_transfer_with_two_phase_and_solidification",
             "unidirectional_heat_transfer_with_velocity_and_solidification"]
REJECTED = {
    "unidirectional_heat_transfer_one_phase_cut": "operator type",
    "unidirectional_heat_transfer_one_phase_cut_evaporative_heat_loss": "evaporation",
    "unidirectional_heat_transfer_one_phase_cut_with_velocity": "operator type",
    "unidirectional_heat_transfer_two_phase_cut": "operator type",
    "unidirectional_heat_transfer_two_phase_cut_evaporative_heat_loss": "evaporation",
    "unidirectional_heat_transfer_two_phase_cut_solidification": "operator type",
    "unidirectional_heat_transfer_two_phase_cut_with_velocity": "operator type",
    "unidirectional_heat_transfer_with_radiation_and_convection_amr": "do amr",
}


def test_every_reference_input_is_classified():
    assert set(CASES) == set(SUPPORTED) | set(REJECTED)
    for name in SUPPORTED:
        F.setup(F.parse_parameters(CASES[name]["input"]))
    for name, why in REJECTED.items():
        with pytest.raises(F.UnsupportedFeature) as e:
            F.setup(F.parse_parameters(CASES[name]["input"]))
        assert why in str(e.value), (name, str(e.value))


def _get(obj, path):
    for key in path.split("/"):
        obj = getattr(obj, F._attr(key))
    return obj


# the reference constructors that override the schema's member defaults
OVERRIDDEN = {"heat/linear solver/solver type": "GMRES",
              "heat/linear solver/preconditioner type": "Diagonal",
              "heat/predictor/type": "linear_extrapolation"}


def test_defaults_follow_the_reference_schema():
    prm = F.Parameters()
    checked = 0
    for path, default in FIXTURE["schema_defaults"].items():
        try:
            ours = _get(prm, path)
        except AttributeError:
            continue  # a key of a section the front end does not model
        expect = OVERRIDDEN.get(path, default)
        if isinstance(expect, float) or isinstance(ours, float):
            assert float(ours) == float(expect), path
        else:
            assert ours == expect, path
        checked += 1
    assert checked >= 45


def test_unknown_keys_and_types():
    base = CASES["unidirectional_heat_transfer"]["input"]
    bad = json.loads(json.dumps(base))
    bad["heat"]["linear solver"]["rel tolerence"] = 1e-9
    with pytest.raises(F.UnknownParameter):
        F.parse_parameters(bad)
    bad = json.loads(json.dumps(base))
    bad["case specific"] = {"do three phase": True}
    with pytest.raises(F.UnknownParameter):
        F.setup(F.parse_parameters(bad))
    ok = json.loads(json.dumps(base))
    ok["heat"]["nlsolve"]["max nonlinear iterations"] = "7"  # strings are converted
    ok["heat"]["diffuse"] = {"use volume-specific thermal capacity for phase interpolation": "true"}
    prm = F.parse_parameters(ok)
    assert prm.heat.nlsolve.max_nonlinear_iterations == 7
    assert prm.heat.diffuse.use_volume_specific_thermal_capacity_for_phase_interpolation is True
    for section, value in (("laser", {"power": 100.0}), ("rte", {"enable": True})):
        bad = json.loads(json.dumps(base))
        bad[section] = value
        with pytest.raises(F.UnsupportedFeature):
            F.parse_parameters(bad)


def test_time_iterator_cuts_the_last_step_and_keeps_the_old_increment():
    ts = F.TimeSteppingParameters(start_time=0.0, end_time=0.25, time_step_size=0.1)
    it = F.TimeIterator(ts)
    seen = []
    while not it.is_finished():
        dt = it.compute_next_time_increment()
        seen.append((round(it.current_time, 12), round(dt, 12), round(it.old_increment, 12)))
    assert seen == [(0.1, 0.1, 0.0), (0.2, 0.1, 0.1), (0.25, 0.05, 0.05)]
    it = F.TimeIterator(F.TimeSteppingParameters(time_step_size=0.1, max_n_steps=2))
    n = 0
    while not it.is_finished():
        it.compute_next_time_increment()
        n += 1
    assert n == 2


def test_extrusion_depth_and_norm_scale():
    for p in (1, 2, 4):
        w, x = F.gll_weights(p)
        assert abs(w.sum() - 1.0) < 1e-14 and abs(x[0]) < 1e-15 and abs(x[-1] - 1.0) < 1e-15
        # GLL quadrature is exact to degree 2p - 1
        assert abs(w @ x ** (2 * p - 1) - 1.0 / (2 * p)) < 1e-14
    prm = F.parse_parameters(CASES["unidirectional_heat_transfer"]["input"])
    cs = F.setup(prm)
    assert cs.n_cells == (32, 32, 1) and abs(cs.extrusion - np.sqrt(2.0)) < 1e-15
    # T0 = x: its 2D L2 norm over [0, 0.1]^2 is sqrt(0.1 * 0.1^3 / 3)
    from oracle import cpu
    pb = cpu.Problem(cs.n_cells, cs.degree, lo=cs.box[0], hi=cs.box[1], dirichlet_faces=0)
    assert abs(pb.l2_norm(cs.temperature0) * cs.norm_scale() - np.sqrt(1e-4 / 3.0)) < 1e-15


def matches_printed(value, printed, rel_noise=1e-8):
    """value prints as `printed` (%.5e), or lies within rel_noise * |value| of
    the rounding interval of `printed`."""
    if float(f"{value:.5e}") == printed:
        return True
    exp = int(np.floor(np.log10(abs(printed))))
    half = 0.5 * 10.0 ** (exp - 5)
    dist = max(abs(value - printed) - half, 0.0)
    return dist <= rel_noise * abs(value)


def test_matches_printed():
    assert matches_printed(99.99974989, 99.9998)      # 1.1e-9 relative below the boundary
    assert not matches_printed(99.99974, 99.9998)      # 1e-7 relative below
    assert matches_printed(5.641984e-3, 5.64198e-3)
    assert not matches_printed(5.64199e-3, 5.64198e-3)


def oracle_problem(cs, dt):
    from oracle import cpu
    m = cs.material
    mat = {"type": m.type, "gas_rho": m.gas.density, "gas_cp": m.gas.specific_heat_capacity,
           "gas_k": m.gas.thermal_conductivity, "liq_rho": m.liquid.density,
           "liq_cp": m.liquid.specific_heat_capacity, "liq_k": m.liquid.thermal_conductivity,
           "sol_rho": m.solid.density, "sol_cp": m.solid.specific_heat_capacity,
           "sol_k": m.solid.thermal_conductivity, "solidus": m.solidus_temperature,
           "liquidus": m.liquidus_temperature,
           "solid_liquid_sharp": int(m.solid_liquid_properties_transition_type == "sharp"),
           "sharp": int(m.two_phase_fluid_properties_transition_type == "sharp"),
           "volume_specific": int(m.use_volume_specific_thermal_capacity_for_phase_interpolation)}
    robin = dict(convection_faces=cs.convection_faces,
                 convection_coefficient=cs.convection.convection_coefficient,
                 convection_t_inf=cs.convection.temperature_infinity,
                 radiation_faces=cs.radiation_faces, emissivity=cs.radiation.emissivity,
                 radiation_t_inf=cs.radiation.temperature_infinity)
    return cpu.Problem(cs.n_cells, cs.degree, dt=dt, material=mat, heaviside=cs.heaviside,
                       lo=cs.box[0], hi=cs.box[1], dirichlet_faces=cs.dirichlet_faces,
                       robin=robin, velocity="constant" if cs.velocity else None,
                       velocity_value=cs.velocity, neumann=cs.neumann_bc)


@pytest.mark.parametrize("name", SUPPORTED)
def test_oracle_reproduces_reference_norms_through_the_front_end(name):
    from oracle.newton import NewtonData, linear_predictor, newton_time_step
    prm = F.parse_parameters(CASES[name]["input"])
    cs = F.setup(prm)
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
    norms = []
    while not it.is_finished():
        dt = it.compute_next_time_increment()
        pb.set_dt(dt)
        if cs.heaviside_at is not None:
            pb.set_heaviside(cs.heaviside_at(it.current_time))
        T_old_old, T_old = T_old, T
        pred = linear_predictor(T_old, T_old_old, dt, it.old_increment)
        T, rep = newton_time_step(pb, pred, T_old, data=data, solver=ls.solver_type,
                                  rel_tol=ls.rel_tolerance, abs_tol=ls.abs_tolerance,
                                  max_it=ls.max_iterations)
        assert rep["converged"], (name, it.n_steps, rep)
        norms.append(pb.l2_norm(T) * cs.norm_scale())
    ref = CASES[name]["l2_norms"]
    assert len(norms) == len(ref)
    for step, (v, r) in enumerate(zip(norms, ref)):
        assert matches_printed(v, r), (name, step + 1, v, r)


def test_solidification_case_set_up():
    """do solidification: T0 = LinearTemp(1960, 1980), the material type of
    determine_material_type, and for two phases the covected vertical level
    set 0.1 * 2/3 - x - v t, re-evaluated at each step's time (its zero
    level, where the smoothed heaviside is 1/2, moves by -v t in x)."""
    prm = F.parse_parameters(CASES["unidirectional_heat_transfer_with_two_phase_and_solidification"]["input"])
    cs = F.setup(prm)
    assert cs.material.type == "gas_liquid_solid" and cs.material.solid.thermal_conductivity == 17.0
    assert (cs.material.solidus_temperature, cs.material.liquidus_temperature) == (1966.0, 1974.0)
    X = np.broadcast_to(cs.coords[0][None, None, :], (len(cs.coords[2]), len(cs.coords[1]),
                                                      len(cs.coords[0]))).reshape(-1)
    assert np.allclose(cs.temperature0, 1960.0 + 200.0 * X, rtol=0, atol=1e-12)
    v = 1e-4
    for t in (0.0, 1.0, 5.0):
        H = cs.heaviside_at(t)
        x0 = 0.1 * 2.0 / 3.0 - v * t
        # the heaviside is 1 left of the moving level, 0 right of it, 1/2 on it
        assert np.all(H[X < x0 - 0.01] == 1.0) and np.all(H[X >= x0 + 0.01] == 0.0)
        mid = F.smoothed_heaviside(np.array([0.0]), 0.01)[0]
        assert mid == pytest.approx(0.5, abs=1e-15)
    assert not np.array_equal(cs.heaviside_at(0.0), cs.heaviside_at(5.0))
    assert np.array_equal(cs.heaviside, cs.heaviside_at(prm.time_stepping.start_time))
    prm = F.parse_parameters(CASES["unidirectional_heat_transfer_with_velocity_and_solidification"]["input"])
    cs = F.setup(prm)
    assert cs.material.type == "liquid_solid" and cs.heaviside is None and cs.heaviside_at is None
    assert cs.velocity == (-0.005, 0.0, 0.0)
    base = CASES["unidirectional_heat_transfer_with_velocity_and_solidification"]["input"]
    # a latent heat of fusion adds its apparent capacity (default type qlq)
    latent = json.loads(json.dumps(base))
    latent["material"]["latent heat of fusion"] = 2.5e5
    latent["material"]["apparent capacity type"] = "poly4_bell"
    cs = F.setup(F.parse_parameters(latent))
    assert cs.material.latent_heat_of_fusion == 2.5e5
    assert cs.material.apparent_capacity_type == "poly4_bell"
    # ... but, as in the reference, not with the volume-specific heat capacity
    bad = json.loads(json.dumps(latent))
    bad["heat"]["diffuse"] = {"use volume-specific thermal capacity for phase interpolation": True}
    with pytest.raises(F.UnsupportedFeature, match="latent heat"):
        F.parse_parameters(bad)
