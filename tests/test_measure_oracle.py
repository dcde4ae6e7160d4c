"""Pins oracle/measure_oracle.py (the checker of askja-measure) without its
own formulas: the line averages and plaquette sums against the C oracle's
independent restatements, the eigen-phases against matrices built with known
spectra (including the winding shift of a centre element), slice traces of
cold links, and the two-pass statistics on repeated configurations."""
import math
import os

import numpy as np
import pytest

from conftest import random_su

mo = pytest.importorskip("oracle.measure_oracle")


def test_lines_and_plaquettes_vs_c_oracle(oracle, tmp_path):
    n, ext = 3, [4, 3, 2, 2]
    o = oracle.Oracle(n, ext)
    rng = np.random.default_rng(3)
    u = random_su(oracle, rng, (o.V, o.dim), n)
    o.field("u")[...] = u
    assert abs(mo.line_mean(u, ext, 0) - o.polyakov_measure()) < 1e-13
    for d in range(1, 4):
        assert abs(mo.line_mean(u, ext, d) - o.wilson_measure(d)) < 1e-13
    ut, us = mo.plaq_exp(u, ext)
    ot, os_ = o.plaq_exp()
    assert abs(ut - ot) < 1e-11 and abs(us - os_) < 1e-11


def test_eigen_phases_known_spectra(oracle):
    rng = np.random.default_rng(4)
    q = random_su(oracle, rng, (), 4)
    th = np.array([-2.5, -0.3, 0.9, 1.9])  # sums to 0: no winding shift
    w = q @ np.diag(np.exp(1j * th)) @ q.conj().T
    assert np.allclose(mo.evs(w), np.sort(th), atol=1e-12)
    z = np.exp(2j * np.pi / 3) * np.eye(3)  # phases 2pi/3 each, winding 1
    assert np.allclose(mo.evs(z), 0.0, atol=1e-12)


def test_slice_traces_and_two_pass_statistics(tmp_path):
    n, ext = 2, [4, 2, 2]
    V = int(np.prod(ext))
    cold = np.broadcast_to(np.eye(n, dtype=complex), (V, len(ext), n, n)).copy()
    assert np.allclose(mo.slice_traces(cold, ext), 1.0)
    path = tmp_path / "c.conf"
    cold.tofile(path)
    r = mo.measure([str(path)] * 3, n, ext, 4.0)
    assert r["mean"]["P"] == 1.0 and r["mean"]["Ut"] == 0.0
    assert all(v == 0.0 for v in r["err"].values()) and all(c == 0.0 for c in r["corr"])
    assert len(r["running"]) == 2
    width = math.pi / (2 * n)
    zero_bin = int(math.floor(math.pi / width))
    assert r["hist"][zero_bin][1] == n and r["hist"][zero_bin][2] == n  # n phases at 0 per configuration
    assert sum(h[1] for h in r["hist"]) == n
