"""askja_hmc_trajectory_noalloc (the allocation-free trajectory) through the
C ABI against askja_hmc_trajectory: the same decisions, energies and links,
bit for bit, on two calls in a row, the second on a smaller lattice; a
lattice whose partial sums exceed the static buffer is refused."""
import ctypes as C

import numpy as np
import pytest

import askja_amd as A

pytestmark = pytest.mark.gpu


def _pair(n, ext, prec, seed):
    a, b = A.Lattice(n, ext, prec), A.Lattice(n, ext, prec)
    a.haar(seed, 0)
    b.set_links(a.links())
    return a, b


def _noalloc(lat, beta, dt, nd, level, seed, counter, reunit):
    p = A.HmcParams(beta, dt, nd, level, 1 if reunit else 0, seed)
    r = A.HmcResult()
    rc = A.lib().askja_hmc_trajectory_noalloc(lat._h, C.byref(p), counter, C.byref(r))
    assert rc == 0, A.lib().askja_last_error(lat._h)
    return r


@pytest.mark.parametrize("n,big,small,prec", [
    (3, [8, 16, 16, 16], [4, 4, 4, 4], A.FP64),
    (3, [8, 8, 8, 8], [4, 4, 2, 2], A.FP32),
    (16, [4, 6, 6, 6], [2, 4, 2, 2], A.FP64),
    (8, [6, 6, 6, 6], [4, 4, 4, 2], A.FP32),
], ids=["su3-fp64", "su3-fp32", "su16-fp64", "su8-fp32"])
def test_noalloc_matches_trajectory(n, big, small, prec):
    beta = 5.9 if n == 3 else 0.4 * n * n
    for ext, counter in ((big, 3), (small, 4)):
        ref, nal = _pair(n, ext, prec, 11 + counter)
        for k in range(2):
            r0 = ref.hmc_trajectory(beta, 0.02, 10, 6, 5, counter + k, reunit=True)
            r1 = _noalloc(nal, beta, 0.02, 10, 6, 5, counter + k, True)
            for f in ("h_old", "h_new", "dh", "uniform", "action", "accepted", "accept_case"):
                assert getattr(r0, f) == getattr(r1, f), (ext, k, f, getattr(r0, f), getattr(r1, f))
            assert np.array_equal(ref.links(), nal.links()), (ext, k)
        # the ordinary entry point still works on the context after the static one
        r2 = nal.hmc_trajectory(beta, 0.02, 10, 6, 5, counter + 2, reunit=True)
        r3 = ref.hmc_trajectory(beta, 0.02, 10, 6, 5, counter + 2, reunit=True)
        assert r2.dh == r3.dh and np.array_equal(ref.links(), nal.links())


def test_noalloc_refuses_oversized_lattice():
    # SU(9) 20^4: one partial-sum slot of 6 doubles per link, 3.84 M > 2^21
    lat = A.Lattice(9, [20, 20, 20, 20], A.FP32)
    p = A.HmcParams(30.0, 0.02, 2, 6, 0, 1)
    r = A.HmcResult()
    rc = A.lib().askja_hmc_trajectory_noalloc(lat._h, C.byref(p), 0, C.byref(r))
    assert rc == -1
    assert b"static buffer" in A.lib().askja_last_error(lat._h)
