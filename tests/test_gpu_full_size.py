"""The workloads at their full sizes through the C ABI: SU(3) 8x32^3 and 16^4,
SU(16) 6x24^3, SU(64) 128^2.

- SU(3): one kick+drift of the exact mode against the oracle, bit for bit
  (the oracle takes well under a second per full-size step with OpenMP), and
  the FP64 and FP32 modes within their per-step bars (1e-13 and 2e-5
  relative).
- SU(16) and SU(64): the oracle would need minutes per step, so the MFMA
  modes are held to the exact mode (the VALU kernels that are bit-identical
  to the oracle at every N on the small lattices of the parity tests) within
  the same bars (BF16x3: 2e-4 N), and the FP64 action pass with its
  temporal / spatial split to the exact mode's plaquette sums within 1e-12.
- Properties that do not depend on the lattice size: Polyakov lines of Haar
  links are themselves Haar, so E|Tr P|^2 = 1; the MD is reversible; the
  energy error of a fixed-length trajectory falls as dt^2; the action summed
  inside the kicks equals the plaquette reduction.
"""
import numpy as np
import pytest

import askja_amd as A

pytestmark = pytest.mark.gpu

SEED = 120351616


def _start(n, ext, prec, beta, stream=2):
    lat = A.Lattice(n, ext, prec)
    lat.haar(SEED, stream)
    lat.mom_refresh(SEED, 0, beta)
    return lat


def _bar(lat, u_ref, p_ref, tol, n):
    scale_p = np.max(np.abs(p_ref))
    dp = np.max(np.abs(lat.momenta() - p_ref)) / scale_p
    du = np.max(np.abs(lat.links() - u_ref))
    assert dp <= tol * n and du <= tol * n, (dp, du)


@pytest.mark.parametrize("ext,level", [([8, 32, 32, 32], 6), ([16, 16, 16, 16], 8)], ids=["8x32^3", "16^4"])
def test_su3_full_size_kick_drift_vs_oracle(oracle, ext, level):
    n, beta, eps = 3, 5.9, 0.02
    exact = _start(n, ext, A.FP64_EXACT, beta)
    u, p = exact.links(), exact.momenta()
    o = oracle.Oracle(n, ext, beta=beta, level=level)
    for name, v in (("u", u), ("U", u), ("pi", p), ("Pi", p)):
        o.field(name)[...] = v
    o.updateE(eps)
    o.updateUandcopyDown(eps)
    u1, p1 = o.field("u").copy(), o.field("pi").copy()
    del o
    exact.kick_drift(eps, eps, level)
    assert np.array_equal(exact.momenta(), p1)
    assert np.array_equal(exact.links(), u1)
    exact.close()
    for prec, tol in ((A.FP64, 1e-13), (A.FP32, 2e-5)):
        lat = A.Lattice(n, ext, prec)
        lat.set_links(u)
        lat.set_momenta(p)
        lat.kick_drift(eps, eps, level)
        _bar(lat, u1, p1, tol, n)
        lat.close()


@pytest.mark.parametrize("n,ext,level,beta", [(16, [6, 24, 24, 24], 8, 102.4), (64, [128, 128], 8, 4096.0)],
                         ids=["SU16-6x24^3", "SU64-128^2"])
def test_mfma_full_size_vs_exact_mode(n, ext, level, beta):
    eps = 0.01 if n == 16 else 0.005
    exact = _start(n, ext, A.FP64_EXACT, beta)
    u, p = exact.links(), exact.momenta()
    t_ex, s_ex = exact.plaq_split()
    exact.kick_drift(eps, eps, level)
    u1, p1 = exact.links(), exact.momenta()
    exact.close()
    for prec, tol in ((A.FP64, 1e-13), (A.FP32, 2e-5), (A.BF16X3, 2e-4)):
        lat = A.Lattice(n, ext, prec)
        lat.set_links(u)
        if prec == A.FP64:
            # the MFMA action pass and its split from the direction-0 links
            t, s = lat.plaq_split()
            assert abs(t - t_ex) <= 1e-12 * abs(t_ex) and abs(s - s_ex) <= 1e-12 * abs(s_ex), (t, t_ex, s, s_ex)
        lat.set_momenta(p)
        lat.kick_drift(eps, eps, level)
        _bar(lat, u1, p1, tol, n)
        lat.close()


@pytest.mark.parametrize("n,ext,window", [(3, [8, 32, 32, 32], 0.04), (3, [16, 16, 16, 16], 0.1),
                                          (16, [6, 24, 24, 24], 0.06)],
                         ids=["SU3-8x32^3", "SU3-16^4", "SU16-6x24^3"])
def test_polyakov_lines_of_haar_links(n, ext, window):
    """Each Polyakov line is a product of independent Haar links, hence Haar:
    E|Tr P|^2 = 1 and E Tr P = 0 (for these N, |Tr P|^2 has variance 1, so the
    windows are 5-6 standard errors of the mean over the spatial sites)."""
    lat = A.Lattice(n, ext, A.FP64)
    lat.haar(SEED, 2)
    w = lat.wilson_lines(0) * n
    assert w.shape == (np.prod(ext) // ext[0],)
    assert abs(np.mean(np.abs(w) ** 2) - 1.0) < window
    assert abs(np.mean(w)) < window
    lat.close()


def test_full_size_reversibility_and_energy():
    """SU(3) 8x32^3, FP64: 10 MD steps forward, momenta negated, 10 back
    return the start; the action the kicks sum equals the plaquette
    reduction; halving dt at fixed trajectory length divides dH by 4."""
    n, ext, beta = 3, [8, 32, 32, 32], 5.9
    lat = _start(n, ext, A.FP64, beta)
    u0, p0 = lat.links(), lat.momenta()
    t, s = lat.plaq_split()
    e = lat.md_trajectory(0.02, 10, 6)
    assert abs(e[0] - (t + s)) <= 1e-12 * (t + s)
    t1, s1 = lat.plaq_split()
    assert abs(e[2] - (t1 + s1)) <= 1e-12 * (t1 + s1)
    lat.mom_negate()
    lat.md_trajectory(0.02, 10, 6)
    assert np.max(np.abs(lat.links() - u0)) < 1e-10
    assert np.max(np.abs(lat.momenta() + p0)) < 1e-9

    def dh(dt, steps):
        lat.set_links(u0)
        lat.set_momenta(p0)
        e = lat.md_trajectory(dt, steps, 8)
        return beta * ((e[2] - e[0]) + (e[3] - e[1]) / (2 * n))

    d1, d2 = dh(0.01, 20), dh(0.005, 40)
    assert 3.2 < d1 / d2 < 4.8, (d1, d2)
    lat.close()
