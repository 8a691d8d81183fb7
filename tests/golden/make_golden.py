"""Writes the golden fixtures of tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

* heat_*.npz: vmult and diagonal of the constrained heat tangent on tiny
  meshes, from the dense brute-force assembly in oracle/brute.py (explicit
  3D quadrature, no sum factorisation), with the seeded inputs.
* material_known_answers.json: the values printed by the reference's own
  unit test unit_tests/material.cc (unit_tests/material.output), transcribed
  as data: MaterialData gas k=10 cp=0 rho=1000, liquid k=40 cp=1000 rho=8000,
  solid k=30 cp=900 rho=7000, mushy zone between solidus 0 and liquidus 100,
  smooth two-phase blend, evaluated at heaviside 0.2 / temperature 20
  (double) and heaviside 0.6 / temperature 60 (VectorizedArray), with the
  temperature derivatives of the liquid_solid and gas_liquid_solid rows.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import brute  # noqa: E402
from oracle.rng import random_vector  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

GAS = {"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0}
TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}

CASES = [
    # name, n, p, dt, material, k_amplitude, velocity, faces, heaviside
    ("q2_gas", (2, 2, 2), 2, 1e-3, GAS, 0.0, False, 0x3F, False),
    ("q3_two_phase", (2, 3, 2), 3, 1e-3, TWO, 0.0, False, 0x3F, True),
    ("q4_sine_vortex", (2, 2, 2), 4, 5e-4, dict(GAS, gas_k=1e-2), 0.5, True, 0x3F, False),
    ("q1_partial_faces", (3, 2, 1), 1, 1e-2, GAS, 0.0, False, 0b010110, False),
    ("q5_sharp_volspec_vortex", (1, 1, 1), 5, 1e-3,
     dict(TWO, sharp=1, volume_specific=1, liq_cp=3.0), 0.0, True, 0x3F, True),
    ("q2_noslip_none", (3, 1, 2), 2, 1e-1, dict(GAS, gas_rho=2.0, gas_cp=3.0, gas_k=0.5), 0.0, False, 0, False),
]


def main():
    for name, n, p, dt, mat, amp, vel, faces, use_h in CASES:
        H = None
        if use_h:
            H = 0.5 * (random_vector(n, p, 7, dirichlet_faces=0) + 1.0)
        A, mask = brute.assemble(n, p, dt, mat, heaviside=H, k_amplitude=amp, velocity=vel,
                                 dirichlet_faces=faces)
        src = random_vector(n, p, 0, dirichlet_faces=faces)
        dst = brute.vmult(A, mask, src)
        diag = brute.diagonal(A, mask)
        np.savez(os.path.join(OUT, f"heat_{name}.npz"), n=np.array(n), degree=p, dt=dt,
                 material=json.dumps(mat), k_amplitude=amp, velocity=int(vel), faces=faces,
                 heaviside=H if H is not None else np.zeros(0), src=src, dst=dst, diag=diag,
                 mask=mask)
        print(name, src.size)


def write_known_answers():
    known = {
        "material_data": {
            "gas": {"thermal_conductivity": 10.0, "specific_heat_capacity": 0.0, "density": 1000.0},
            "liquid": {"thermal_conductivity": 40.0, "specific_heat_capacity": 1000.0, "density": 8000.0},
            "solid": {"thermal_conductivity": 30.0, "specific_heat_capacity": 900.0, "density": 7000.0},
            "solid_liquid_properties_transition_type": "mushy_zone",
            "two_phase_fluid_properties_transition_type": "smooth",
            "solidus_temperature": 0.0,
            "liquidus_temperature": 100.0,
        },
        "source": "unit_tests/material.cc + unit_tests/material.output",
        "rows": [
            {"type": "gas", "heaviside": 0.0, "k": 10, "cp": 0, "rho": 1000},
            {"type": "liquid", "heaviside": 0.0, "k": 40, "cp": 1000, "rho": 8000},
            {"type": "gas_liquid", "heaviside": 0.2, "k": 16, "cp": 200, "rho": 2400},
            {"type": "gas_liquid", "heaviside": 0.6, "k": 28, "cp": 600, "rho": 5200},
            {"type": "gas_liquid_consistent_with_evaporation", "heaviside": 0.2, "k": 16, "cp": 200,
             "rho": 1212.12},
            {"type": "gas_liquid_consistent_with_evaporation", "heaviside": 0.6, "k": 28, "cp": 600,
             "rho": 2105.26},
            {"type": "liquid_solid", "heaviside": 0.0, "temperature": 20.0, "k": 31.04, "cp": 910.4,
             "rho": 7104, "d_k": 0.096, "d_cp": 0.96, "d_rho": 9.6},
            {"type": "liquid_solid", "heaviside": 0.0, "temperature": 60.0, "k": 36.48, "cp": 964.8,
             "rho": 7648, "d_k": 0.144, "d_cp": 1.44, "d_rho": 14.4},
            {"type": "gas_liquid_solid", "heaviside": 0.2, "temperature": 20.0, "k": 14.208,
             "cp": 182.08, "rho": 2220.8, "d_k": 0.0192, "d_cp": 0.192, "d_rho": 1.92},
            {"type": "gas_liquid_solid", "heaviside": 0.6, "temperature": 60.0, "k": 25.888,
             "cp": 578.88, "rho": 4988.8, "d_k": 0.0864, "d_cp": 0.864, "d_rho": 8.64},
        ],
        "print_precision_digits": 6,
    }
    with open(os.path.join(OUT, "material_known_answers.json"), "w") as f:
        json.dump(known, f, indent=1)


if __name__ == "__main__":
    main()
    write_known_answers()
