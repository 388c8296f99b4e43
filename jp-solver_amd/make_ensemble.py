#!/usr/bin/env python3
"""Draws the 32-member TGV ensemble and writes it as JSON (ensemble32.json next to this file).

Members: edge N uniform over {32, 48, ..., 160}; Re log-uniform in [400, 3200]; u0 uniform in
[0.03, 0.08]; phase shifts uniform in [0, 2 pi)^3; members 0-15 uniform lattices, 16-31 3-level AMR
(coarse N0 = N, sensor refinement on both levels, frac1 0.3, frac2 0.1, regrid every 10 coarse steps).
The drawn list is committed, so the workload does not depend on the generator's RNG implementation.
"""
import json
import math
import os

import numpy as np


def draw(count=32, seed=1600):
    rng = np.random.default_rng(seed)
    sizes = list(range(32, 161, 16))
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        re = float(math.exp(rng.uniform(math.log(400.0), math.log(3200.0))))
        u0 = float(rng.uniform(0.03, 0.08))
        phase = [float(x) for x in rng.uniform(0.0, 2 * math.pi, 3)]
        m = {"n": n, "re": re, "u0": u0, "phase": phase, "amr": int(i >= count // 2)}
        if m["amr"]:
            m.update({"l1": 2, "l2": 2, "frac1": 0.3, "frac2": 0.1, "adapt": 10})
        out.append(m)
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ensemble32.json")
    with open(path, "w") as f:
        json.dump({"seed": 1600, "members": draw()}, f, indent=1)
    print(path)
