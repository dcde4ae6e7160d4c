"""The spectral map (askja_flow_define_spectral, any N, any D) against its
CPU restatement (oracle/spectral_oracle.py, pinned in
tests/test_spectral_oracle.py): forward links and log J for N = 2 ... 64 in
D = 2 and for N = 2, 4, 9 in D = 3 and 4, the
identity at zero weights against the plain Wilson force, the hand-written
backward against finite differences of the oracle's Phi = beta S_W(F(V)) -
log J(V) (links and MLP weights, eps = 1e-6, FP64), field-transformation HMC
reversibility and dH ~ dt^2, one sub-layer at the 128^2 SU(64) size, and
FTHMC ensembles against the exact two-dimensional plaquette."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg

import askja_amd as A
from askja_amd import learned as L
from oracle import spectral_oracle as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TMAX = 0.4
SUBS = [(0, 0, 0), (1, 1, 0), (0, 2, 1), (1, 3, 1), (0, 1, 0), (1, 0, 1)]


def _setup(n, seed, ext=(8, 8), subs=SUBS, scale=0.5, nblocks=2):
    rng = np.random.default_rng(seed)
    V = S.random_field(ext, n, rng)
    lat = A.Lattice(n, list(ext), A.FP64)
    lat.set_links(V)
    bsize = lat.flow_define([s[0] for s in subs], [s[1] for s in subs], [s[2] for s in subs], nblocks, TMAX,
                            kind="spectral")
    assert bsize == 32 * 3 + 65
    w = (rng.standard_normal((nblocks, bsize)) * scale).astype(np.float32)
    lat.flow_set_weights(w.ravel())
    return rng, V, lat, w


def _ta(m):
    h = (m - m.conj().T) / 2j
    return h - np.trace(h) / m.shape[0] * np.eye(m.shape[0])


def plain_gradient(U, ext, beta):
    """Lie gradient of beta S_W: beta TA(U_l sum of staples) / (2N), D = 2."""
    n = U.shape[-1]
    g = np.zeros_like(U)
    for x in range(U.shape[0]):
        for mu in (0, 1):
            nu = 1 - mu
            xp, xpn, xmn = S.shift(ext, x, mu), S.shift(ext, x, nu), S.shift(ext, x, nu, -1)
            xpmn = S.shift(ext, xp, nu, -1)
            st = (U[xp, nu] @ U[xpn, mu].conj().T @ U[x, nu].conj().T
                  + U[xpmn, nu].conj().T @ U[xmn, mu].conj().T @ U[xmn, nu])
            g[x, mu] = beta * _ta(U[x, mu] @ st) / (2 * n)
    return g


@pytest.mark.parametrize("n", [2, 3, 5, 8, 16, 24, 64])
def test_spectral_forward_matches_oracle(n):
    rng, V, lat, w = _setup(n, n)
    lj = lat.flow_apply()
    U = lat.flow_output()
    Uo, ljo = S.forward(V, (8, 8), SUBS, w, TMAX)
    err = np.max(np.abs(U - Uo))
    print(f"SU({n}) 8x8 spectral forward: max |U - U_oracle| = {err:.2e}, log J = {lj:.10f} ({ljo:.10f})")
    assert err < (1e-12 if n <= 8 else 1e-10)
    assert abs(lj - ljo) < 1e-10 * max(1.0, abs(ljo)) and abs(ljo) > 1e-2
    assert np.max(np.abs(U - V)) > 1e-3


GEOMS_D34 = {"D3": ((4, 8, 4), [(0, 0, 0), (1, 1, 0), (2, 2, 1), (0, 3, 1), (2, 0, 0)]),
             "D4": ((4, 4, 4, 4), [(0, 0, 0), (1, 1, 0), (2, 2, 1), (3, 3, 1), (0, 1, 0), (3, 2, 1)])}


@pytest.mark.parametrize("geom,n", [("D3", 4), ("D4", 2), ("D4", 4), ("D4", 9)])
def test_spectral_forward_matches_oracle_d34(geom, n):
    ext, subs = GEOMS_D34[geom]
    rng, V, lat, w = _setup(n, 40 + n, ext=ext, subs=subs)
    lj = lat.flow_apply()
    U = lat.flow_output()
    Uo, ljo = S.forward(V, ext, subs, w, TMAX)
    err = np.max(np.abs(U - Uo))
    print(f"SU({n}) {geom} spectral forward: max |U - U_oracle| = {err:.2e}, log J = {lj:.10f} ({ljo:.10f})")
    assert err < 1e-12
    assert abs(lj - ljo) < 1e-10 * max(1.0, abs(ljo)) and abs(ljo) > 1e-2


@pytest.mark.parametrize("geom,n", [("D3", 3), ("D4", 2)])
def test_spectral_link_gradient_vs_finite_differences_d34(geom, n):
    ext, subs = GEOMS_D34[geom]
    beta = 1.5 * n
    rng, V, lat, w = _setup(n, 50 + n, ext=ext, subs=subs)
    s, lj = lat.flow_gradient(beta)
    Uo, ljo = S.forward(V, ext, subs, w, TMAX)
    assert abs(s - S.wilson_action(Uo, ext)) < 1e-10 * max(1.0, abs(s))
    g = lat.flow_link_gradient()
    gens = S.su_generators(n)
    eps = 1e-6
    for trial in range(8):
        x, d = int(rng.integers(V.shape[0])), int(rng.integers(len(ext)))
        X = sum(c * T for c, T in zip(rng.standard_normal(len(gens)), gens))
        vals = []
        for sgn in (1, -1):
            W = V.copy()
            W[x, d] = scipy.linalg.expm(1j * sgn * eps * X) @ V[x, d]
            vals.append(S.phi_action(W, ext, subs, w, TMAX, beta))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = 2 * np.trace(g[x, d] @ X).real
        assert abs(fd - an) < 2e-6 * max(1.0, abs(an)), (x, d, fd, an)


def test_spectral_fthmc_reversible_d4():
    n, ext, beta = 4, [8, 8, 4, 4], 8.0
    lat = A.Lattice(n, ext, A.FP64)
    lat.haar(6, 0)
    for t in range(3):
        lat.hmc_trajectory(beta, 0.05, 10, 8, 6, t)
    mu, off, blk = L.spectral_masks(1, 4)
    lat.flow_define(mu, off, blk, 1, TMAX, kind="spectral")
    lat.flow_set_weights((np.random.default_rng(6).standard_normal(161) * 0.5).astype(np.float32))
    V0 = lat.links()
    lat.mom_refresh(9, 0, beta)
    cf = 2 * n / beta

    def kick(eps):
        lat.flow_gradient(beta)
        lat.set_momenta(lat.momenta() - eps * cf * lat.flow_link_gradient())

    def md(k, dt):
        kick(dt / 2)
        for j in range(k):
            lat.updateUandcopyDown(dt, 8)
            kick(dt if j < k - 1 else dt / 2)

    p0 = lat.momenta()
    md(5, 0.05)
    assert np.max(np.abs(lat.links() - V0)) > 1e-3
    lat.mom_negate()
    md(5, 0.05)
    du, dp = np.max(np.abs(lat.links() - V0)), np.max(np.abs(lat.momenta() + p0))
    print(f"SU({n}) D=4 FTHMC forward + reversed: max |V - V0| = {du:.2e}, max |P + P0| = {dp:.2e}")
    assert du < 3e-11 * n and dp < 3e-10 * n


def test_spectral_identity_is_plain_force():
    n, ext, beta = 3, (8, 8), 4.5
    rng, V, lat, w = _setup(n, 4, scale=0.0)
    s, lj = lat.flow_gradient(beta)
    assert lj == 0.0
    assert abs(s - S.wilson_action(V, ext)) < 1e-10
    g = lat.flow_link_gradient()
    assert np.max(np.abs(g - plain_gradient(V, ext, beta))) < 1e-12
    assert np.max(np.abs(lat.flow_output() - V)) < 1e-13


@pytest.mark.parametrize("n", [2, 3, 8])
def test_spectral_link_gradient_vs_finite_differences(n):
    ext, beta = (8, 8), 1.5 * n
    rng, V, lat, w = _setup(n, 20 + n)
    s, lj = lat.flow_gradient(beta)
    Uo, ljo = S.forward(V, ext, SUBS, w, TMAX)
    assert abs(s - S.wilson_action(Uo, ext)) < 1e-10 * max(1.0, abs(s))
    assert abs(lj - ljo) < 1e-10 * max(1.0, abs(ljo))
    g = lat.flow_link_gradient()
    gens = S.su_generators(n)
    eps = 1e-6
    for trial in range(10):
        x, d = int(rng.integers(V.shape[0])), int(rng.integers(2))
        X = sum(c * T for c, T in zip(rng.standard_normal(len(gens)), gens))
        vals = []
        for sgn in (1, -1):
            W = V.copy()
            W[x, d] = scipy.linalg.expm(1j * sgn * eps * X) @ V[x, d]
            vals.append(S.phi_action(W, ext, SUBS, w, TMAX, beta))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = 2 * np.trace(g[x, d] @ X).real
        assert abs(fd - an) < 2e-6 * max(1.0, abs(an)), (x, d, fd, an)
    assert np.allclose(g, np.conj(np.swapaxes(g, -1, -2)), atol=1e-14)
    assert np.allclose(np.trace(g, axis1=-2, axis2=-1), 0, atol=1e-11)


def test_spectral_weight_gradient_vs_finite_differences():
    n, ext, beta = 3, (8, 8), 4.5
    rng, V, lat, w = _setup(n, 7)
    lat.flow_gradient(beta, wgrad=True)
    gw = lat.flow_weight_gradient()
    bsize = w.shape[1]
    h = np.float32(2.0 ** -9)
    for idx in [0, 5, 96 + 3, 96 + 32 + 7, 160, bsize + 11, bsize + 96 + 20, bsize + 160]:
        vals = []
        for sgn in (1, -1):
            ww = w.ravel().copy()
            ww[idx] = np.float32(ww[idx] + sgn * h)
            vals.append(S.phi_action(V, ext, SUBS, ww.reshape(w.shape), TMAX, beta))
        fd = (vals[0] - vals[1]) / (2 * float(h))
        assert abs(fd - gw[idx]) < 1e-4 * max(1.0, abs(fd)), (idx, fd, gw[idx])


def _thermal(n, ext, beta, seed, ntraj, dt, nd, layers=1, scale=0.5):
    lat = A.Lattice(n, list(ext), A.FP64)
    lat.haar(seed, 0)
    for t in range(ntraj):
        lat.hmc_trajectory(beta, dt, nd, 8, seed, t)
    mu, off, blk = L.spectral_masks(layers)
    lat.flow_define(mu, off, blk, layers, TMAX, kind="spectral")
    rng = np.random.default_rng(seed)
    lat.flow_set_weights((rng.standard_normal(layers * 161) * scale).astype(np.float32))
    return lat


@pytest.mark.parametrize("n,ext,beta", [(3, (16, 16), 6.0), (16, (8, 8), 25.6)])
def test_spectral_fthmc_reversible(n, ext, beta):
    lat = _thermal(n, ext, beta, 6, 3, 0.05, 10)
    V0 = lat.links()
    lat.mom_refresh(9, 0, beta)
    cf = 2 * n / beta

    def kick(eps):
        lat.flow_gradient(beta)
        lat.set_momenta(lat.momenta() - eps * cf * lat.flow_link_gradient())

    def md(k, dt):
        kick(dt / 2)
        for j in range(k):
            lat.updateUandcopyDown(dt, 8)
            kick(dt if j < k - 1 else dt / 2)

    p0 = lat.momenta()
    md(5, 0.05)
    assert np.max(np.abs(lat.links() - V0)) > 1e-3
    lat.mom_negate()
    md(5, 0.05)
    du, dp = np.max(np.abs(lat.links() - V0)), np.max(np.abs(lat.momenta() + p0))
    print(f"SU({n}) FTHMC forward + reversed: max |V - V0| = {du:.2e}, max |P + P0| = {dp:.2e}")
    # round-off of 10 MD steps of N x N products grows about linearly in N
    assert du < 3e-11 * n
    assert dp < 3e-10 * n


@pytest.mark.parametrize("n,ext,beta", [(3, (16, 16), 6.0), (64, (8, 8), 4096.0)])
def test_spectral_fthmc_energy_error_scales_as_dt_squared(n, ext, beta):
    dt0 = 0.1 if n == 3 else 0.02
    lat = _thermal(n, ext, beta, 5, 3, dt0 / 2, 20)
    V0 = lat.links()
    dh = []
    for dt, nd in ((dt0, 5), (dt0 / 2, 10)):
        lat.set_links(V0)
        r = lat.flow_hmc_trajectory(beta, dt, nd, 8, 7, 11, reunit=False)
        dh.append(abs(r.dh))
    print(f"SU({n}) FTHMC |dH| at dt {dt0}, {dt0 / 2}: {dh}")
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_spectral_sublayer_at_128x128_su64_matches_oracle():
    """One sub-layer of the map on the 128^2 SU(64) lattice: every updated
    link against the oracle's eigen-decomposition route (a sample of 48 links
    for the links themselves, all 4096 for log J), every other link unchanged."""
    n, ext = 64, (128, 128)
    lat = A.Lattice(n, list(ext), A.FP64)
    lat.haar(3, 0)
    V = lat.links()
    rng = np.random.default_rng(8)
    lat.flow_define([1], [2], [0], 1, TMAX, kind="spectral")
    w = (rng.standard_normal(161) * 0.5).astype(np.float32)
    lat.flow_set_weights(w)
    lj = lat.flow_apply()
    U = lat.flow_output()
    act = S.active_sites(ext, 1, 2)
    mask = np.ones(U.shape[:2], bool)
    mask[act, 1] = False
    assert np.array_equal(U[mask], V[mask])
    ljo, worst = 0.0, 0.0
    sample = set(rng.choice(len(act), 48, replace=False).tolist())
    for k, x in enumerate(act):
        St, f = S.link_inputs(V, ext, x, 1)
        t = TMAX * np.tanh(S.mlp(w, f))
        Wp, l = S.spectral_map(V[x, 1] @ St, t)
        ljo += l
        if k in sample:
            worst = max(worst, np.max(np.abs(U[x, 1] - Wp @ St.conj().T)))
    print(f"128^2 SU(64) sub-layer: max |U - U_oracle| = {worst:.2e} (48 links), log J = {lj:.8f} ({ljo:.8f})")
    assert worst < 1e-10
    assert abs(lj - ljo) < 1e-9 * abs(ljo) and abs(ljo) > 1.0


def _mean_err(x):
    x = np.asarray(x)
    t = A.tau_int(x)[0] if len(x) >= 8 else 1.0
    t = 1.0 if not np.isfinite(t) or t < 1 else t
    return x.mean(), x.std() * math.sqrt(t / len(x))


def _exact(n, beta):
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "exact_2d_plaquette.json")))
    return [p["plaquette"] for p in gold["points"] if p["nc"] == n and p["beta"] == beta][0]


def test_spectral_fthmc_su3_exact_plaquette():
    """FTHMC with the spectral map (one sweep, random weights) on 32^2 SU(3)
    at beta 6: the plaquette of U = F(V) against the exact value."""
    n, beta = 3, 6.0
    lat = _thermal(n, (32, 32), beta, 11, 60, 0.1, 10, scale=0.2)
    for t in range(20):
        lat.flow_hmc_trajectory(beta, 0.05, 20, 8, 12, t)
    plaq, acc = [], 0
    for t in range(500):
        r = lat.flow_hmc_trajectory(beta, 0.05, 20, 8, 12, 100 + t)
        acc += r.accepted
        plaq.append(1.0 - r.action / lat.n_plaquettes())
    m, e = _mean_err(plaq)
    want = _exact(n, beta)
    print(f"SU(3) 32^2 FTHMC: plaquette {m:.6f} +- {e:.6f} (exact {want:.6f}), acceptance {acc / 500:.2f}")
    assert acc / 500 > 0.5
    assert abs(m - want) < 4 * e + 1e-4, (m, e, want)


def test_spectral_fthmc_su64_exact_plaquette():
    """FTHMC with the spectral map on 16^2 SU(64) at beta = N^2 (the
    Gross-Witten point): the plaquette of U = F(V) against the exact value."""
    n, beta = 64, 4096.0
    lat = _thermal(n, (16, 16), beta, 13, 40, 0.025, 40, scale=0.3)
    for t in range(10):  # the V chain's own equilibrium differs from the plain chain's
        lat.flow_hmc_trajectory(beta, 0.025, 40, 8, 14, 1000 + t)
    plaq, acc = [], 0
    for t in range(100):
        r = lat.flow_hmc_trajectory(beta, 0.025, 40, 8, 14, t)
        acc += r.accepted
        plaq.append(1.0 - r.action / lat.n_plaquettes())
    m, e = _mean_err(plaq)
    want = _exact(n, beta)
    print(f"SU(64) 16^2 FTHMC: plaquette {m:.7f} +- {e:.7f} (exact {want:.7f}), acceptance {acc / 100:.2f}")
    assert acc / 100 > 0.5
    assert abs(m - want) < 4 * e + 2e-4, (m, e, want)


@pytest.mark.parametrize("n,beta,beta_q", [(3, 6.0, 5.0), (8, 51.2, 44.8)])
def test_trained_spectral_map_lowers_heldout_objective(n, beta, beta_q):
    """What the trained network gains over the method without it, for the
    spectral map.  Untrained (w2 = b2 = 0) the map is the identity and the
    objective E_q[beta S_W(F(V)) - log J(V)] is beta E_q[S_W(V)] (plain HMC);
    it is KL(q || p~) up to a constant, q the plain-HMC distribution at
    beta_q and p~ the learned chain's distribution in V at beta.  Trained on
    8 plain-HMC configurations of a 16^2 lattice (one sweep, 8 sub-layers)
    the map must lower it on 4 held-out configurations.  Runs in a
    child process: torch must own the HIP runtime before the library loads."""
    import subprocess
    import sys

    code = r"""
import sys, json, numpy as np, torch
torch.cuda.init()
sys.path.insert(0, %r)
import askja_amd as A
from askja_amd import learned as L
n, beta, bq = %d, %r, %r
lat = A.Lattice(n, [16, 16], A.FP64)
lat.haar(5, 3)
L.plain_chain(lat, bq, 0.05, 20, 8, 5, 0, 50, measure=False)
configs = []
for k in range(12):
    L.plain_chain(lat, bq, 0.05, 20, 8, 5, 1000 + 3 * k, 3, measure=False)
    configs.append(lat.links())
mu, off, blk = L.spectral_masks(1)
bsize = lat.flow_define(mu, off, blk, 1, 0.4, kind="spectral")
w0 = L.init_weights(1, bsize, 3, seed=1)
def objective(w, cs):
    lat.flow_set_weights(w)
    out = []
    for V in cs:
        lat.set_links(V)
        s, lj = lat.flow_gradient(beta)
        out.append(beta * s - lj)
    return np.array(out)
wt, losses = L.train(lat, configs[:8], beta, w0, epochs=10, lr=1e-2)
u, t = objective(w0, configs[8:]), objective(wt, configs[8:])
print(json.dumps({"untrained": u.tolist(), "gain": (u - t).tolist(), "train": [losses[0], losses[-1]]}))
""" % (ROOT, n, beta, beta_q)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    gain = np.asarray(out["gain"])
    print(f"SU({n}) 16^2 spectral map: held-out objective gain per configuration {gain} (nats), "
          f"training loss {out['train'][0]:.3f} -> {out['train'][1]:.3f}")
    assert out["train"][-1] < out["train"][0]
    assert gain.mean() > 0.0, gain
