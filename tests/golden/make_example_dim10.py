"""Writes tests/golden/example_dim10_trajectory1.json: the first trajectory
of the reference's own example params file (params_N8S10T10B20k.txt:3-20:
NC 8, DIM 10, N0 = N1 = 10 and eight directions of extent 1, BETA 20000,
DT 0.01, LEVEL 8, UPDATE2 1000, cold start, RANDOM 0 so the momenta come
from mt19937 seed 120351616 and the first trajectory has 999 drifts),
computed by the CPU oracle in sun.cpp's FP64 operation order.  The
size-1 directions make every neighbour step along them a self-reference.

Stored: SHA-256 of the final links and momenta (U-*.conf layout), H before and
after, and sample links.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

NC, EXT, BETA, DT, LEVEL, SEED, NDRIFT = 8, [10, 10] + [1] * 8, 20000.0, 0.01, 8, 120351616, 999
SAMPLES = [(0, 0), (17, 1), (42, 5), (99, 9)]


def main():
    t0 = time.time()
    o = O.Oracle(NC, EXT, beta=BETA, level=LEVEL)
    o.initializeU()
    o.initializeE(SEED)
    h0 = o.hamiltonian("pi")
    o.md_trajectory(DT, NDRIFT)
    h1 = o.hamiltonian("Pi")
    u = o.field("u")
    pi = o.field("Pi")
    out = {
        "workload": "params_N8S10T10B20k.txt: NC 8, DIM 10, 10x10x1^8, BETA 20000, DT 0.01, LEVEL 8, cold, "
                    "mt19937 seed 120351616, 999 drifts",
        "nc": NC, "extents": EXT, "beta": BETA, "dt": DT, "level": LEVEL, "seed": SEED, "ndrift": NDRIFT,
        "h_old": h0, "h_new": h1,
        "links_sha256": hashlib.sha256(np.ascontiguousarray(u).tobytes()).hexdigest(),
        "momenta_sha256": hashlib.sha256(np.ascontiguousarray(pi).tobytes()).hexdigest(),
        "samples": [{"site": s, "dir": d, "re": u[s, d].real.ravel().tolist(), "im": u[s, d].imag.ravel().tolist()}
                    for s, d in SAMPLES],
        "seconds": time.time() - t0,
    }
    with open(os.path.join(HERE, "example_dim10_trajectory1.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("h_old", "h_new", "links_sha256", "seconds")}))


if __name__ == "__main__":
    main()
