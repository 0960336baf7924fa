"""Writes uneven_batch_draws.json: the 48 draws of the uneven batch of Newton
searches (SplitMix64 seed 0), kept as a fixture so that a change of the draw
procedure shows up in the tests."""
import json
import os
import sys
from dataclasses import asdict

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from kalnajs_log_spirals_amd.UnevenBatch import draw_batch  # noqa: E402

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "uneven_batch_draws.json")
    with open(out, "w") as f:
        json.dump({"seed": 0, "draws": [asdict(d) for d in draw_batch(48, 0)]}, f, indent=1)
    print("wrote", out)
