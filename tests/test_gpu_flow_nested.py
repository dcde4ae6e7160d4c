"""The two-time-scale field-transformation HMC trajectory
(askja_flow_hmc_trajectory_nested): the plain Wilson action on the fine step
(the fused kick+drift), the flow's correction Phi - beta S_W on the coarse
step.  With one inner step it is the one-scale trajectory up to rounding
(same links, dH, accept decision); its energy error scales as dt^2 when both
time scales are halved, with the plain force unscaled and scaled
(alpha != 1); a non-divisor inner count is refused; and its chain samples
the target: with the spectral map on 32^2 SU(3) at beta 6, with the plain
force scaled by 0.8, the plaquette of U = F(V) agrees with the exact
two-dimensional value (tests/golden/exact_2d_plaquette.json)."""
import json
import math
import os

import numpy as np
import pytest

import askja_amd as A
from askja_amd import learned as L

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stout(seed, ext=(8, 8, 8, 8), beta=6.0, scale=0.3):
    lat = A.Lattice(3, list(ext), A.FP64)
    lat.haar(seed, 0)
    for t in range(3):
        lat.hmc_trajectory(beta, 0.02, 20, 8, seed, t)
    mu, par, blk = L.sweep_masks(4, 1)
    lat.flow_define(mu, par, blk, 1, 0.15, kind="stout")
    rng = np.random.default_rng(seed)
    lat.flow_set_weights((rng.standard_normal(lat.flow_bsize) * scale).astype(np.float32))
    return lat


def _spectral(n, ext, beta, seed, ntraj, dt, nd, scale):
    lat = A.Lattice(n, list(ext), A.FP64)
    lat.haar(seed, 0)
    for t in range(ntraj):
        lat.hmc_trajectory(beta, dt, nd, 8, seed, t)
    mu, off, blk = L.spectral_masks(1)
    lat.flow_define(mu, off, blk, 1, 0.4, kind="spectral")
    rng = np.random.default_rng(seed)
    lat.flow_set_weights((rng.standard_normal(161) * scale).astype(np.float32))
    return lat


@pytest.mark.parametrize("kind", ["stout", "spectral"])
def test_one_inner_step_is_the_one_scale_trajectory(kind):
    if kind == "stout":
        lat, beta, dt, nd = _stout(3), 6.0, 0.02, 10
    else:
        lat, beta, dt, nd = _spectral(3, (16, 16), 6.0, 4, 5, 0.05, 10, 0.5), 6.0, 0.05, 10
    V0 = lat.links()
    r1 = lat.flow_hmc_trajectory(beta, dt, nd, 8, 21, 5, reunit=False)
    V1 = lat.links()
    lat.set_links(V0)
    r2 = lat.flow_hmc_trajectory(beta, dt, nd, 8, 21, 5, reunit=False, inner=1)
    V2 = lat.links()
    print(f"{kind}: dH one-scale {r1.dh:.3e}, nested(1) {r2.dh:.3e}")
    assert r1.accepted == r2.accepted
    assert abs(r1.h_old - r2.h_old) < 1e-9 * abs(r1.h_old)
    assert abs(r1.dh - r2.dh) < 1e-8
    assert np.max(np.abs(V1 - V2)) < 1e-11


@pytest.mark.parametrize("alpha", [1.0, 0.9])
def test_nested_energy_error_scales_as_dt_squared(alpha):
    """Fine step and coarse step halved together (4 fine steps per coarse
    step, trajectory length 0.2): |dH| drops about fourfold."""
    lat = _stout(5)
    V0 = lat.links()
    dh = []
    for dt, nd in ((0.04, 5 * 4), (0.02, 10 * 4)):
        lat.set_links(V0)
        r = lat.flow_hmc_trajectory(6.0, dt / 4, nd, 8, 7, 11, reunit=False, inner=4, alpha=alpha)
        dh.append(abs(r.dh))
    print(f"nested FTHMC (alpha {alpha}) |dH| at coarse step 0.04 / 0.02: {dh}")
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_inner_must_divide_ndrift():
    lat = _stout(6, ext=(4, 4, 4, 4))
    with pytest.raises(A.AskjaError, match="divide"):
        lat.flow_hmc_trajectory(6.0, 0.02, 10, 8, 1, 0, reunit=False, inner=3)
    A.lib().askja_clear_error(lat._h)
    lat.flow_hmc_trajectory(6.0, 0.02, 10, 8, 1, 0, reunit=False, inner=5)


def _mean_err(x):
    x = np.asarray(x)
    t = A.tau_int(x)[0] if len(x) >= 8 else 1.0
    t = 1.0 if not np.isfinite(t) or t < 1 else t
    return x.mean(), x.std() * math.sqrt(t / len(x))


def test_nested_spectral_su3_exact_plaquette():
    """32^2 SU(3) at beta 6, one spectral sweep (random weights), trajectory
    length 1.0 as 40 fine steps of 0.025 with the flow's correction every 4
    (11 flow gradients per trajectory) against 0.8 times the plain force
    (any scale keeps the chain exact): the plaquette of U = F(V) against
    the exact value."""
    n, beta = 3, 6.0
    lat = _spectral(n, (32, 32), beta, 11, 60, 0.1, 10, 0.2)
    alpha = 0.8
    print(f"least-squares plain-force scale of this map: {L.plain_force_scale(lat, beta):.4f}")
    for t in range(20):
        lat.flow_hmc_trajectory(beta, 0.025, 40, 8, 12, t, inner=4, alpha=alpha)
    plaq, acc = [], 0
    for t in range(500):
        r = lat.flow_hmc_trajectory(beta, 0.025, 40, 8, 12, 100 + t, inner=4, alpha=alpha)
        acc += r.accepted
        plaq.append(1.0 - r.action / lat.n_plaquettes())
    m, e = _mean_err(plaq)
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "exact_2d_plaquette.json")))
    want = [p["plaquette"] for p in gold["points"] if p["nc"] == n and p["beta"] == beta][0]
    print(f"SU(3) 32^2 nested FTHMC (alpha {alpha:.4f}): plaquette {m:.6f} +- {e:.6f} (exact {want:.6f}), "
          f"acceptance {acc / 500:.2f}")
    assert acc / 500 > 0.5
    assert abs(m - want) < 4 * e + 1e-4, (m, e, want)
