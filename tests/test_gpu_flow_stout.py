"""The stout form of the learned map (askja_flow_define_stout: U' =
exp(-i rho TAH(U S)) U per masked link, exact 8 x 8 log det of the su(3)
Jacobian) against its CPU restatement (oracle/flow_oracle.py, whose log det
is pinned to a numerical Jacobian in test_flow_oracle.py): forward links and
log J, the hand-written backward against finite differences of the oracle's
Phi = beta S_W(F(V)) - log J(V) (links and MLP weights, eps = 1e-6, FP64),
rho = 0 against the plain force, field-transformation HMC's dH ~ dt^2 and
reversibility, and the forward at the 16^4 size."""
import numpy as np
import pytest
import scipy.linalg

import askja_amd as A
from askja_amd import learned as L
from conftest import random_su

pytestmark = pytest.mark.gpu

RHO, BETA = 0.15, 5.7
GEOMS = [([4, 4, 2, 2], [(0, 0, 0), (0, 1, 0), (1, 0, 1), (2, 1, 1), (3, 0, 0), (1, 1, 1), (3, 1, 0), (2, 0, 1)]),
         ([4, 4, 4], [(0, 0, 0), (1, 1, 0), (2, 0, 1), (0, 1, 1), (1, 0, 0), (2, 1, 1)])]
IDS = ["D4", "D3"]


@pytest.fixture(scope="module")
def fo():
    from oracle import flow_oracle
    return flow_oracle


def _setup(oracle, seed, ext, subs, scale=0.5):
    o = oracle.Oracle(3, ext)
    rng = np.random.default_rng(seed)
    V = random_su(oracle, rng, (o.V, o.dim), 3)
    lat = A.Lattice(3, ext, A.FP64)
    lat.set_links(V)
    bsize = lat.flow_define([s[0] for s in subs], [s[1] for s in subs], [s[2] for s in subs], 2, RHO, kind="stout")
    w = (rng.standard_normal(2 * bsize) * scale).astype(np.float32)
    lat.flow_set_weights(w)
    nf = 3 * (o.dim - 1) * (o.dim - 2) // 2
    return o, rng, V, lat, w, bsize, nf


@pytest.mark.parametrize("ext,subs", GEOMS, ids=IDS)
def test_stout_forward_matches_oracle(oracle, fo, ext, subs):
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 1, ext, subs)
    lj = lat.flow_apply()
    U = lat.flow_output()
    Uo, ljo = fo.stout_forward_batched(ext, V, subs, w, bsize, nf, RHO)
    assert np.max(np.abs(U - Uo)) < 1e-12
    assert abs(lj - ljo) < 1e-11 * max(1.0, abs(ljo)) and abs(ljo) > 1e-2
    assert np.max(np.abs(U - V)) > 1e-3


@pytest.mark.parametrize("ext,subs", GEOMS, ids=IDS)
def test_stout_link_gradient_vs_finite_differences(oracle, fo, ext, subs):
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 2, ext, subs)
    s, lj = lat.flow_gradient(BETA)
    g = lat.flow_link_gradient()
    gens = oracle.generators(3)
    eps = 1e-6
    for trial in range(8):
        x, d = int(rng.integers(o.V)), int(rng.integers(o.dim))
        X = np.einsum("a,aij->ij", rng.standard_normal(8), gens)
        vals = []
        for sgn in (1, -1):
            W = V.copy()
            W[x, d] = scipy.linalg.expm(1j * sgn * eps * X) @ V[x, d]
            vals.append(fo.stout_phi(ext, W, subs, w, bsize, nf, RHO, BETA))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = 2 * np.trace(g[x, d] @ X).real
        assert abs(fd - an) < 2e-6 * max(1.0, abs(an)), (x, d, fd, an)
    assert np.allclose(g, np.conj(np.swapaxes(g, -1, -2)), atol=1e-14)
    assert np.allclose(np.trace(g, axis1=-2, axis2=-1), 0, atol=1e-12)


@pytest.mark.parametrize("ext,subs", GEOMS, ids=IDS)
def test_stout_weight_gradient_vs_finite_differences(oracle, fo, ext, subs):
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 3, ext, subs)
    lat.flow_gradient(BETA, wgrad=True)
    gw = lat.flow_weight_gradient()
    h = np.float32(2.0 ** -9)
    for idx in [0, 5, 32 * nf + 3, 32 * nf + 32 + 7, 32 * nf + 64, bsize + 11, bsize + 32 * nf + 64]:
        vals = []
        for sgn in (1, -1):
            ww = w.copy()
            ww[idx] = np.float32(w[idx] + sgn * h)
            vals.append(fo.stout_phi(ext, V, subs, ww, bsize, nf, RHO, BETA))
        fd = (vals[0] - vals[1]) / (2 * float(h))
        assert abs(fd - gw[idx]) < 1e-4 * max(1.0, abs(fd)), (idx, fd, gw[idx])


def test_stout_identity_is_plain_force(oracle):
    ext, subs = GEOMS[0]
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 4, ext, subs, scale=0.0)
    s, lj = lat.flow_gradient(BETA)
    assert lj == 0.0
    g = lat.flow_link_gradient()
    o.field("u")[...] = V
    o.field("pi")[...] = 0
    o.updateE(-1.0)
    assert np.max(np.abs(g - BETA / 6.0 * o.field("Pi"))) < 1e-12
    assert abs(s - o.energy_mag()) < 1e-10


def _stout_lattice(ext, layers, seed, scale=0.3):
    lat = A.Lattice(3, ext, A.FP64)
    lat.haar(seed, 0)
    for t in range(3):
        lat.hmc_trajectory(6.0, 0.02, 20, 8, seed, t)
    mu, par, blk = L.sweep_masks(len(ext), layers)
    bsize = lat.flow_define(mu, par, blk, layers, RHO, kind="stout")
    rng = np.random.default_rng(seed)
    lat.flow_set_weights((rng.standard_normal(layers * bsize) * scale).astype(np.float32))
    return lat, list(zip(mu, par, blk)), bsize


def test_stout_fthmc_energy_error_scales_as_dt_squared():
    lat, subs, bsize = _stout_lattice([8, 8, 8, 8], 2, 5)
    V0 = lat.links()
    dh = []
    for dt, nd in ((0.04, 5), (0.02, 10)):
        lat.set_links(V0)
        r = lat.flow_hmc_trajectory(6.0, dt, nd, 8, 7, 11, reunit=False)
        dh.append(abs(r.dh))
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_stout_fthmc_reversible():
    lat, subs, bsize = _stout_lattice([8, 8, 8, 8], 2, 6)
    V0 = lat.links()
    lat.mom_refresh(9, 0, 6.0)
    cf = 2 * 3 / 6.0

    def kick(eps):
        lat.flow_gradient(6.0)
        lat.set_momenta(lat.momenta() - eps * cf * lat.flow_link_gradient())

    def md(n, dt):
        kick(dt / 2)
        for k in range(n):
            lat.updateUandcopyDown(dt, 8)
            kick(dt if k < n - 1 else dt / 2)

    p0 = lat.momenta()
    md(5, 0.02)
    assert np.max(np.abs(lat.links() - V0)) > 1e-3
    lat.mom_negate()
    md(5, 0.02)
    assert np.max(np.abs(lat.links() - V0)) < 1e-10
    assert np.max(np.abs(lat.momenta() + p0)) < 1e-9


def test_stout_forward_16x4_matches_oracle(fo):
    """One sweep (8 sub-layers) of the stout map on 16^4 against the batched
    oracle: every link and log J."""
    ext = [16, 16, 16, 16]
    lat, subs, bsize = _stout_lattice(ext, 1, 7)
    V = lat.links()
    lj = lat.flow_apply()
    U = lat.flow_output()
    # the weights _stout_lattice drew
    w = (np.random.default_rng(7).standard_normal(bsize) * 0.3).astype(np.float32)
    Uo, ljo = fo.stout_forward_batched(ext, V, subs, w, bsize, 9, RHO)
    print(f"16^4 stout forward: max |U - U_oracle| = {np.max(np.abs(U - Uo)):.2e}, log J = {lj:.6f} ({ljo:.6f})")
    assert np.max(np.abs(U - Uo)) < 1e-11
    assert abs(lj - ljo) < 1e-10 * abs(ljo) and abs(ljo) > 1.0
