"""Writes unidirectional_heat_transfer_robin.json from the reference's own
regression inputs and outputs of the unidirectional_heat_transfer case with
convective and/or radiative faces (run in a checkout that holds the reference
at REF; the GPU tests only read the written fixture).

Per case it keeps the Newton-converged ||solution|| of each time step (the
L2 norm of the temperature field, printed with 6 significant digits) and the
parameters of the heat section. The case itself (unidirectional_heat_transfer
.cpp: 2D box [0, 0.1]^2, colorized ids, initial T = 1000 when a Robin face is
present, the Robin conditions on the right face x = 0.1, id 1) fixes the rest.
"""
import json
import os
import re
import sys

REF = os.environ.get("MPDG_REFERENCE", "/root/reference")
CASE = "applications/mp-heat-transfer/cases/unidirectional_heat_transfer/tests"
HERE = os.path.dirname(os.path.abspath(__file__))
RIGHT_FACE = 1  # colorized boundary id of x = x_max (bit 1 of a face mask)


def main():
    out = {"source": f"{CASE}/unidirectional_heat_transfer_<case>.json and .mpirun=4.output",
           "cases": {}}
    setup = None
    for case in ("with_convection", "with_radiation", "with_radiation_and_convection"):
        base = os.path.join(REF, CASE, f"unidirectional_heat_transfer_{case}")
        prm = json.load(open(base + ".json"))
        text = open(base + ".mpirun=4.output").read()
        norms = [float(v) for v in re.findall(r"\|\|solution\|\| = ([0-9.eE+-]+)", text)]
        heat = prm["heat"]
        conv = heat.get("convective boundary condition", {})
        rad = heat.get("radiative boundary condition", {})
        alpha = float(conv.get("convection coefficient", 0.0))
        eps = float(rad.get("emissivity", 0.0))
        mat = prm["material"]["gas"]
        s = {"dimension": prm["base"]["dimension"],
             "global_refinements": prm["base"]["global refinements"],
             "degree": prm["base"]["fe"]["degree"],
             "time_step": prm["time stepping"]["time step size"],
             "steps": prm["time stepping"]["max n steps"],
             "box": [[0.0, 0.0], [0.1, 0.1]],
             "initial_temperature": 1000.0,
             "density": mat["density"], "specific_heat_capacity": mat["specific heat capacity"],
             "thermal_conductivity": mat["thermal conductivity"],
             "linear_solver": {"rel_tolerance": heat["linear solver"]["rel tolerance"]},
             "nlsolve": {"residual_tolerance": heat["nlsolve"]["residual tolerance"]}}
        assert setup is None or setup == s, "cases differ in their common setup"
        setup = s
        out["cases"][case] = {
            "convection_faces": (1 << RIGHT_FACE) if alpha > 0.0 else 0,
            "convection_coefficient": alpha,
            "convection_temperature_infinity": float(conv.get("temperature infinity", 0.0)),
            "radiation_faces": (1 << RIGHT_FACE) if eps > 0.0 else 0,
            "emissivity": eps,
            "radiation_temperature_infinity": float(rad.get("temperature infinity", 0.0)),
            "l2_norms": norms,
        }
    out["setup"] = setup
    path = os.path.join(HERE, "unidirectional_heat_transfer_robin.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    sys.exit(main())
