"""One fixed-momentum trajectory of the reference's example workload at
DIM 4 (NC 8, 10^4, BETA 20, DT 0.01, LEVEL 8, cold start, mt19937 seed
120351616 momenta, 999 drifts) against the oracle's result committed in
tests/golden/c1_trajectory1.json (written by make_c1_trajectory.py), and the
same for the reference's own example params file as it stands (DIM 10,
10 x 10 x 1^8, BETA 20000; tests/golden/example_dim10_trajectory1.json,
written by make_example_dim10.py).

FP64_EXACT must reproduce the oracle's links and momenta bit for bit (equal
SHA-256); FP64 (FMA, factored staples) must agree link by link to 1e-10.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import askja_amd as A
from conftest import ROOT

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_trajectory1.json")))


def _run(prec):
    lat = A.Lattice(GOLD["nc"], GOLD["extents"], prec)
    lat.cold()
    lat.mom_gsl(GOLD["seed"], GOLD["beta"])
    e = lat.md_trajectory(GOLD["dt"], GOLD["ndrift"], GOLD["level"])
    return lat, e


def test_c1_trajectory_exact_bitwise():
    lat, e = _run(A.FP64_EXACT)
    u = lat.links()
    assert hashlib.sha256(np.ascontiguousarray(u).tobytes()).hexdigest() == GOLD["links_sha256"]
    p = lat.momenta()
    assert hashlib.sha256(np.ascontiguousarray(p).tobytes()).hexdigest() == GOLD["momenta_sha256"]
    for s in GOLD["samples"]:
        want = np.array(s["re"]) + 1j * np.array(s["im"])
        assert np.array_equal(u[s["site"], s["dir"]].ravel(), want)
    beta, n2 = GOLD["beta"], 2.0 * GOLD["nc"]
    assert abs(beta * (e[0] + e[1] / n2) - GOLD["h_old"]) < 1e-9 * GOLD["h_old"]
    assert abs(beta * (e[2] + e[3] / n2) - GOLD["h_new"]) < 1e-9 * GOLD["h_old"]


def test_c1_trajectory_fast_link_by_link():
    ex, _ = _run(A.FP64_EXACT)
    fast, e = _run(A.FP64)
    d = np.max(np.abs(fast.links() - ex.links()))
    assert d <= 1e-10, d
    beta, n2 = GOLD["beta"], 2.0 * GOLD["nc"]
    assert abs(beta * (e[2] + e[3] / n2) - GOLD["h_new"]) < 1e-6


PHILOX = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_trajectories23.json")))


def test_c1_trajectories_2_3_exact_bitwise():
    """Trajectories 2 and 3 of the same run (tests/golden/c1_trajectories23.json,
    written by make_c1_philox.py): Philox momenta (counter = trajectory - 1),
    1000 drifts, the Metropolis draw.  The device's momenta, proposed links and
    decisions equal the oracle's bit for bit (the Box-Muller steps are
    reproducible IEEE sequences); every trajectory is rejected, so each starts
    from the cold links again."""
    lat = A.Lattice(PHILOX["nc"], PHILOX["extents"], A.FP64_EXACT)
    beta, n2 = PHILOX["beta"], 2.0 * PHILOX["nc"]
    for tr in PHILOX["trajectories"]:
        lat.cold()
        lat.mom_refresh(PHILOX["seed"], tr["counter"], beta)
        e = lat.md_trajectory(PHILOX["dt"], PHILOX["ndrift"], PHILOX["level"])
        assert hashlib.sha256(np.ascontiguousarray(lat.links()).tobytes()).hexdigest() == tr["links_sha256"]
        assert hashlib.sha256(np.ascontiguousarray(lat.momenta()).tobytes()).hexdigest() == tr["momenta_sha256"]
        h0, h1 = beta * (e[0] + e[1] / n2), beta * (e[2] + e[3] / n2)
        assert abs(h0 - tr["h_old"]) < 1e-9 * tr["h_old"] and abs(h1 - tr["h_new"]) < 1e-9 * tr["h_old"]
        assert lat.uniform(PHILOX["seed"], tr["counter"]) == tr["uniform"]
        assert (h1 < h0 or tr["uniform"] < np.exp(h0 - h1)) == tr["accepted"]


EXAMPLE = json.load(open(os.path.join(ROOT, "tests", "golden", "example_dim10_trajectory1.json")))


def test_example_params_trajectory_exact_bitwise():
    lat = A.Lattice(EXAMPLE["nc"], EXAMPLE["extents"], A.FP64_EXACT)
    lat.cold()
    lat.mom_gsl(EXAMPLE["seed"], EXAMPLE["beta"])
    e = lat.md_trajectory(EXAMPLE["dt"], EXAMPLE["ndrift"], EXAMPLE["level"])
    u = lat.links()
    assert hashlib.sha256(np.ascontiguousarray(u).tobytes()).hexdigest() == EXAMPLE["links_sha256"]
    assert hashlib.sha256(np.ascontiguousarray(lat.momenta()).tobytes()).hexdigest() == EXAMPLE["momenta_sha256"]
    for s in EXAMPLE["samples"]:
        assert np.array_equal(u[s["site"], s["dir"]].ravel(), np.array(s["re"]) + 1j * np.array(s["im"]))
    beta, n2 = EXAMPLE["beta"], 2.0 * EXAMPLE["nc"]
    assert abs(beta * (e[2] + e[3] / n2) - EXAMPLE["h_new"]) < 1e-9 * abs(EXAMPLE["h_old"])
    fast = A.Lattice(EXAMPLE["nc"], EXAMPLE["extents"], A.FP64)
    fast.cold()
    fast.mom_gsl(EXAMPLE["seed"], EXAMPLE["beta"])
    fast.md_trajectory(EXAMPLE["dt"], EXAMPLE["ndrift"], EXAMPLE["level"])
    assert np.max(np.abs(fast.links() - u)) <= 1e-10
