"""Writes ensemble_config4.json: the 64 boxes of the batched-ensemble workload
drawn by symmetry_amd.ic.ensemble_config4 (seed 7), committed so that the
draws are fixed even if the generator's implementation changes."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from symmetry_amd import ic  # noqa: E402

if __name__ == "__main__":
    boxes = ic.ensemble_config4()
    keep = [{k: (list(v) if isinstance(v, tuple) else v) for k, v in b.items()} for b in boxes]
    with open(os.path.join(HERE, "ensemble_config4.json"), "w") as f:
        json.dump({"seed": 7, "generator": "numpy.random.default_rng (PCG64)", "boxes": keep}, f, indent=0)
