"""Writes particle_config5.json: the 200 spheres of the dense particle-laden
interface workload drawn by symmetry_amd.ic.particle_config5 (seed 99):
centres, densities and contact angles, committed so that the draws are fixed
even if the generator's implementation changes."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from symmetry_amd import ic  # noqa: E402

if __name__ == "__main__":
    c = ic.particle_config5()
    with open(os.path.join(HERE, "particle_config5.json"), "w") as f:
        json.dump({"seed": 99, "generator": "numpy.random.default_rng (PCG64)", "n": c["nx"], "rad": c["rad"],
                   "centers": c["centers"].tolist(), "rho_p": c["rho_p"].tolist(), "cita_deg": c["cita_deg"].tolist()},
                  f, indent=0)
