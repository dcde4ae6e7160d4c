"""Pins oracle/exact2d.py (the exact 2D SU(N) plaquette used to check the GPU
chains at D = 2) without its own formula: direct integration over the
eigenphases with the Weyl measure for SU(2) and SU(3) (periodic trapezoid
rule, exponentially convergent), the strong-coupling expansion, and the
committed fixture against the Gross-Witten large-N value."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

exact2d = pytest.importorskip("oracle.exact2d")


def weyl_plaquette(n, beta, m=256):
    x = beta / n
    t = 2 * np.pi * np.arange(m) / m
    if n == 2:
        th = [t, -t]
    else:
        a, b = np.meshgrid(t, t, indexing="ij")
        th = [a, b, -a - b]
    meas = np.ones_like(th[0])
    for i in range(n):
        for j in range(i + 1, n):
            meas = meas * np.abs(np.exp(1j * th[i]) - np.exp(1j * th[j])) ** 2
    re = sum(np.cos(p) for p in th)
    w = meas * np.exp(x * (re - n))
    return float((w * re).sum() / w.sum() / n)


@pytest.mark.parametrize("n,beta", [(2, 1.0), (2, 4.0), (3, 2.0), (3, 6.0), (3, 15.0)])
def test_bessel_sum_equals_weyl_integral(n, beta):
    assert abs(exact2d.plaquette(n, beta, dps=40) - weyl_plaquette(n, beta)) < 1e-12


def test_strong_coupling():
    # <ReTr U / N> = x <(ReTr U)^2> / N + O(x^2) with <(ReTr U)^2> = 1/2 (N >= 3)
    n, beta = 4, 1e-3
    assert abs(exact2d.plaquette(n, beta, dps=40) - beta / (2 * n * n)) < 1e-8


def test_fixture():
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "exact_2d_plaquette.json")))
    pts = {(p["nc"], p["beta"]): p["plaquette"] for p in gold["points"]}
    assert abs(pts[(3, 6.0)] - weyl_plaquette(3, 6.0)) < 1e-12
    # beta = N^2 is the Gross-Witten point lambda = 2 N^2 / beta = 2, where
    # <ReTr U / N> -> 1/2 as N -> infinity
    assert abs(pts[(64, 4096.0)] - 0.5) < 0.02
