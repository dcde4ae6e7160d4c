"""Writes tests/golden/c1_trajectories23.json: trajectories 2 and 3 of the
reference's example workload at DIM 4 (NC 8, 10^4, BETA 20, DT 0.01, LEVEL 8,
cold start, UPDATE2 1000) as the askja program runs them after the first:
momenta from Philox-4x32-10 (key = run seed 120351616, counter = trajectory
number - 1, reproducible Box-Muller), 1000 drifts, the Metropolis uniform from
Philox (purpose 3, link 0, the same counter).  Computed by the CPU oracle,
which restates sun.cpp's FP64 operation order (sun.cpp:1249-1318 for the
loop, 672-801 for the MD).

The chain state a trajectory starts from is the previous one's outcome; the
first trajectory (tests/golden/c1_trajectory1.json) ends with dH = 110.9, so
it is rejected and trajectory 2 starts cold again.  Stored per trajectory:
H before and after, the uniform, the decision, SHA-256 of the proposed links
and momenta (U-*.conf layout, little-endian doubles).  Takes ~20 min on 8
cores.
"""
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

NC, EXT, BETA, DT, LEVEL, SEED, NDRIFT = 8, [10, 10, 10, 10], 20.0, 0.01, 8, 120351616, 1000


def uniform(seed, counter):
    ctr = [0, 0, counter & 0xFFFFFFFF, (counter >> 32) ^ (3 << 24)]
    r = O.philox4x32_10(ctr, [seed & 0xFFFFFFFF, seed >> 32])
    return (float(r[0]) + 0.5) * 2.3283064365386963e-10


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    t0 = time.time()
    first = json.load(open(os.path.join(HERE, "c1_trajectory1.json")))
    u0 = uniform(SEED, 0)
    acc1 = first["h_new"] < first["h_old"] or u0 < math.exp(first["h_old"] - first["h_new"])
    assert not acc1, "trajectory 1 is accepted: trajectory 2 would start from its links"
    o = O.Oracle(NC, EXT, beta=BETA, level=LEVEL)
    o.initializeU()
    trajs = []
    for counter in (1, 2):
        o.take()
        o.mom_philox(SEED, counter)
        h0 = o.hamiltonian("pi")
        o.md_trajectory(DT, NDRIFT)
        h1 = o.hamiltonian("Pi")
        u = uniform(SEED, counter)
        acc = h1 < h0 or u < math.exp(h0 - h1)
        trajs.append({"trajectory": counter + 1, "counter": counter, "h_old": h0, "h_new": h1, "uniform": u,
                      "accepted": bool(acc), "links_sha256": sha(o.field("u")), "momenta_sha256": sha(o.field("Pi"))})
        print(json.dumps(trajs[-1]), flush=True)
        if not acc:
            o.load()
    out = {
        "workload": "NC 8, DIM 4, 10^4, BETA 20, DT 0.01, LEVEL 8, cold, Philox seed 120351616, 1000 drifts",
        "nc": NC, "extents": EXT, "beta": BETA, "dt": DT, "level": LEVEL, "seed": SEED, "ndrift": NDRIFT,
        "trajectories": trajs, "seconds": time.time() - t0,
    }
    with open(os.path.join(HERE, "c1_trajectories23.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
