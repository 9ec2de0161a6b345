"""Writes the regression fixture small_256_ll.json: the oracle's images of
the 256x256 workload (64 spheres, radius [0.02, 0.08], alpha [0.2, 0.6],
seed 1, Linked List, tail blending on, pool of every fragment) at K = 8 (no
pixel holds more) and K = 4 (the composite tail engages), each as a SHA-256
of its float32 bits plus a few pixels in full, so that a later
change of the oracle or of the scene generator shows up as a diff of this
file. The reference ships no golden vectors of its own; this fixture is
produced by the oracle and pins nothing against the reference by itself.

Run from the repository root: python tests/golden/make_golden.py"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

PIXELS = [0, 32896, 33153, 40000, 65535]


def compute():
    import oit_oracle
    from vk_order_independent_transparency_amd import scene

    p = scene.SMALL_256
    tg, d, c = scene.generate(p)
    res = {
        "workload": "256x256, 64 spheres, radius [0.02, 0.08], alpha [0.2, 0.6], seed 1; Linked List, tail on",
        "fragments": int(len(tg)),
        "fragments_sha256": hashlib.sha256(tg.tobytes() + d.tobytes() + c.tobytes()).hexdigest(),
        "max_list": int(np.bincount(tg, minlength=p.targets).max()),
        "images": {},
    }
    for K in (8, 4):
        cfg = oit_oracle.make_config(oit_oracle.LINKEDLIST, K, 1, p.width, p.height, ll_capacity=len(tg) + 1)
        out = np.ascontiguousarray(oit_oracle.run(cfg, tg, d, c)["out"], dtype=np.float32)
        res["images"][f"K{K}"] = {
            "sha256": hashlib.sha256(out.tobytes()).hexdigest(),
            "pixels": {str(i): [float(x) for x in out[i]] for i in PIXELS},
        }
    return res


if __name__ == "__main__":
    with open(os.path.join(HERE, "small_256_ll.json"), "w") as f:
        json.dump(compute(), f, indent=1)
        f.write("\n")
