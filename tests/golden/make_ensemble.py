"""Writes tests/golden/ensemble_su3_4x4.json: ensemble averages (plaquette,
volume-averaged |Polyakov loop|, and measurements4d.cpp's spatial Wilson line
W_1 and Polyakov loop P-bar) of plain HMC from the CPU oracle (SU(3), D=4, 4^4, beta 5.7, dt 0.05, 20 MD steps
(trajectory length 1), LEVEL 8, Haar start, counter-based momenta, 20000
trajectories after 200 for thermalisation, reunitarised chain) with their
autocorrelation-corrected errors.  The GPU chain (independent seed) must
agree within these errors.  Takes about 10 min on 8 cores.
"""
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

NC, EXT, BETA, DT, NDRIFT, LEVEL = 3, [4, 4, 4, 4], 5.7, 0.05, 20, 8
NTHERM, NTRAJ, SEED = 200, 20000, 777


def tau_int(x, S=1.5):
    x = np.asarray(x, np.float64)
    n = len(x)
    d = x - x.mean()
    var = np.dot(d, d) / n
    f = np.fft.rfft(d, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / (n - np.arange(n)) / var
    tint = 0.5
    for W in range(1, n // 2):
        tint += acf[W]
        tau = S / math.log((2 * tint + 1) / (2 * tint - 1)) if tint > 0.5 else 1e-9
        if math.exp(-W / tau) - tau / math.sqrt(W * n) < 0:
            return 2 * tint
    return 2 * tint


def main():
    t0 = time.time()
    o = O.Oracle(NC, EXT, beta=BETA, level=LEVEL)
    o.haar(SEED, 0)
    o.hmc(DT, NDRIFT, NTHERM, SEED, 0, True)
    rec = o.hmc(DT, NDRIFT, NTRAJ, SEED, NTHERM, True)
    out = {"workload": f"SU(3) D=4 4^4 beta {BETA}, dt {DT}, {NDRIFT} MD steps, LEVEL {LEVEL}",
           "ntraj": NTRAJ, "acceptance": float(rec[:, 1].mean())}
    for name, col in (("plaquette", 2), ("abs_polyakov", 5), ("wilson_measure_x1", 8), ("polyakov_measure", 9)):
        x = rec[:, col]
        tau = tau_int(x)
        out[name] = {"mean": float(x.mean()), "err": float(x.std() * math.sqrt(tau / len(x))), "tau": tau}
    out["seconds"] = time.time() - t0
    with open(os.path.join(HERE, "ensemble_su3_4x4.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
