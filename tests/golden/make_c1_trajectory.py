"""Writes tests/golden/c1_trajectory1.json: the first trajectory of the
reference's example workload at DIM 4 (NC 8, 10^4, BETA 20, DT 0.01,
LEVEL 8, cold start, RANDOM 0 so the momenta come from mt19937 seed
120351616, UPDATE2 1000 so the first trajectory has 999 drifts), computed by
the CPU oracle, which restates sun.cpp's FP64 operation order.

Stored: SHA-256 of the final links and momenta (U-*.conf layout, little-endian
doubles), H before and after, and a few sample links.  Takes ~10 min on 8
cores.
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

NC, EXT, BETA, DT, LEVEL, SEED, NDRIFT = 8, [10, 10, 10, 10], 20.0, 0.01, 8, 120351616, 999
SAMPLES = [(0, 0), (1234, 1), (5678, 2), (9999, 3)]


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
        "workload": "NC 8, DIM 4, 10^4, BETA 20, DT 0.01, LEVEL 8, cold, mt19937 seed 120351616, 999 drifts",
        "nc": NC, "extents": EXT, "beta": BETA, "dt": DT, "level": LEVEL, "seed": SEED, "ndrift": NDRIFT,
        "h_old": h0, "h_new": h1,
        "links_sha256": hashlib.sha256(np.ascontiguousarray(u).tobytes()).hexdigest(),
        "momenta_sha256": hashlib.sha256(np.ascontiguousarray(pi).tobytes()).hexdigest(),
        "samples": [{"site": s, "dir": d, "re": u[s, d].real.ravel().tolist(), "im": u[s, d].imag.ravel().tolist()}
                    for s, d in SAMPLES],
        "seconds": time.time() - t0,
    }
    with open(os.path.join(HERE, "c1_trajectory1.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("h_old", "h_new", "links_sha256", "seconds")}))


if __name__ == "__main__":
    main()
