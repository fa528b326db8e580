"""Writes the golden find fixtures: deterministic synthetic scenes (by
generator and seed, with SHA-256 of every buffer) and the oracle's find
results on them, including the coarse and prepared candidate counts.

Run from the repository root: python tests/golden/make_golden.py
The fixtures pin the oracle against drift (tests/test_oracle.py) and give the
GPU tests expected results that do not depend on the oracle build."""
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mvkit_amd import synth  # noqa: E402
from oracle import pyoracle  # noqa: E402

# name -> (scene factory, model (levels, start, extent), find (min_score, num_matches, max_overlap))
CASES = {
    "plain640": ("workload_plain_640", {}, (3, 0.0, 0.0), (0.8, 3, 0.5)),
    "small_rot": ("small_rot", {}, (0, -math.radians(20), math.radians(40)), (0.7, 2, 0.5)),
    "small_rot_all": ("small_rot", {}, (0, -math.radians(20), math.radians(40)), (0.5, 0, 0.5)),
    "small_rot_frame2": ("small_rot", {"frame": 2}, (0, -math.radians(20), math.radians(40)), (0.6, 0, 0.5)),
    "rot2048": ("workload_rot_2048", {}, (4, -math.pi / 6, math.pi / 3), (0.7, 5, 0.5)),
}


def make(name):
    factory, kw, (levels, start, extent), (ms, nm, ov) = CASES[name]
    s = getattr(synth, factory)(**kw)
    m = pyoracle.Model(s.template, levels=levels, start=start, extent=extent)
    res, tr = m.find(s.image, start, extent, ms, nm, ov, trace=True)
    return {
        "scene": {"factory": factory, "kwargs": kw, "seed": s.seed, "sha256": s.sha256(),
                  "image_shape": list(s.image.shape), "template_shape": list(s.template.shape)},
        "model": {"levels": levels, "angle_start": start, "angle_extent": extent},
        "find": {"min_score": ms, "num_matches": nm, "max_overlap": ov, "subpixel": True},
        "n_coarse": int(tr["info"].n_coarse), "n_prepared": int(tr["info"].n_prepared),
        "results": [[float(v) for v in r] for r in res],
    }


def main():
    pyoracle.build()
    for name in CASES:
        out = make(name)
        with open(os.path.join(HERE, f"find_{name}.json"), "w") as f:
            json.dump(out, f, indent=1)
        print(name, len(out["results"]), out["n_coarse"], out["n_prepared"])


if __name__ == "__main__":
    main()
