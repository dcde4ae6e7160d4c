"""Writes exact_2d_plaquette.json: <ReTr U_p / N> of 2D SU(N) at the
couplings the GPU 2D chains run (oracle/exact2d.py, Bessel-determinant sum),
each evaluated at two working precisions that must agree."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import exact2d  # noqa: E402

POINTS = [(3, 6.0), (64, 4096.0)]


def main():
    out = []
    for n, beta in POINTS:
        a = exact2d.plaquette(n, beta, dps=120)
        b = exact2d.plaquette(n, beta, dps=160)
        assert abs(a - b) < 1e-14, (n, beta, a, b)
        out.append({"nc": n, "beta": beta, "plaquette": b})
        print(n, beta, repr(b), flush=True)
    with open(os.path.join(HERE, "exact_2d_plaquette.json"), "w") as f:
        json.dump({"source": "oracle/exact2d.py (Bessel-determinant sum, two precisions)", "points": out}, f, indent=1)


if __name__ == "__main__":
    main()
