"""Device side of the measurement program against numpy restatements of
measurements4d.cpp on the same links: per-slice traces (scalarLoopcorr,
357-408) and the sorted, winding-recentred eigen-phases of line products
(getevs, 185-309; the reference uses Eigen's ComplexEigenSolver, here
numpy.linalg.eigvals)."""
import numpy as np
import pytest

import askja_amd as A
from conftest import random_su

pytestmark = pytest.mark.gpu


def lex_sites_with(ext, fixed):
    """lexicographic indices of the sites whose coordinates match `fixed` (dict dim -> value)"""
    return [i for i, x in enumerate(np.ndindex(*ext)) if all(x[d] == v for d, v in fixed.items())]


def getevs_line(u, o, start, d, n):
    w = np.eye(n, dtype=complex)
    pos = start
    for _ in range(o.ext[d]):
        w = w @ u[pos, d]
        pos = o.getplus(pos, d)
    th = np.angle(np.linalg.eigvals(w))
    wind = th.sum() / (2 * np.pi)
    if abs(wind) > 0.1:
        th = th - wind * 2 * np.pi / n
    return np.sort(th)


@pytest.mark.parametrize("n,ext", [(2, [4, 3, 2]), (3, [4, 4, 2, 2]), (5, [3, 4, 2, 2]), (12, [2, 3, 2]),
                                   (40, [3, 2])], ids=["N2", "N3", "N5", "N12", "N40"])
def test_line_phases_and_slice_traces(oracle, n, ext):
    o = oracle.Oracle(n, ext)
    rng = np.random.default_rng(n)
    u = random_su(oracle, rng, (o.V, o.dim), n)
    lat = A.Lattice(n, ext, A.FP64)
    lat.set_links(u)
    for d in (0, 1):
        starts = lex_sites_with(ext, {k: 0 for k in range(len(ext)) if k not in (0 if d else 1, d)} | {d: 0})
        got = lat.line_phases(d, starts)
        for q, s in enumerate(starts):
            want = getevs_line(u, o, s, d, n)
            assert np.max(np.abs(got[q] - want)) < 1e-10, (d, q, got[q], want)
    for d in range(len(ext)):
        want = [np.mean([np.trace(u[s, d]).real / n for s in lex_sites_with(ext, {0: x0})]) for x0 in range(ext[0])]
        assert np.max(np.abs(lat.slice_traces(d) - want)) < 1e-13


def test_line_phases_degenerate_spectra(oracle):
    """Cold links (all phases 0) and centre elements (threefold degenerate
    phase 2 pi / 3 per link: winding recentring) in SU(3)."""
    n, ext = 3, [4, 2, 2, 2]
    o = oracle.Oracle(n, ext)
    lat = A.Lattice(n, ext, A.FP64)
    lat.cold()
    assert np.max(np.abs(lat.line_phases(0, [0, 1, 2]))) < 1e-12
    z = np.exp(2j * np.pi / 3)
    u = np.broadcast_to(z * np.eye(3), (o.V, o.dim, 3, 3)).copy()
    lat.set_links(u)
    got = lat.line_phases(1, [0, 4])
    for q, s in enumerate([0, 4]):
        assert np.max(np.abs(got[q] - getevs_line(u, o, s, 1, n))) < 1e-10
