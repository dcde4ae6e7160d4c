"""HIP learned field transformation (flow.hip) against its CPU restatement
(oracle/flow_oracle.py): forward links and log J, the hand-written backward
(gradient with respect to the links and to the MLP weights) against finite
differences of the oracle's Phi = beta S_W(F(V)) - log J(V) (eps = 1e-6,
FP64), the rho = 0 reduction to the plain-HMC force, and energy conservation
of field-transformation HMC."""
import numpy as np
import pytest
import scipy.linalg

import askja_amd as A
from conftest import random_su

pytestmark = pytest.mark.gpu

EXT = [4, 4, 2, 2]
SUBS = [(0, 0, 0), (0, 1, 0), (1, 0, 1), (2, 1, 1), (3, 0, 0), (1, 1, 1), (3, 1, 0), (2, 0, 1), (0, 0, 1), (1, 0, 0)]
# D = 3: one plane without mu per link, so the forward and backward kernels'
# two wavefronts each take that plane's loops (and the staples) alone
EXT3 = [4, 4, 4]
SUBS3 = [(0, 0, 0), (1, 1, 0), (2, 0, 1), (0, 1, 1), (1, 0, 0), (2, 1, 1)]
GEOMS = [(EXT, SUBS), (EXT3, SUBS3)]
GEOM_IDS = ["D4", "D3"]
RHO, DELTA, BETA = 0.3, 1e-3, 5.7


@pytest.fixture(scope="module")
def fo():
    from oracle import flow_oracle
    return flow_oracle


def _setup(oracle, seed, scale=0.5, prec=A.FP64, ext=EXT, subs=SUBS):
    o = oracle.Oracle(3, ext)
    rng = np.random.default_rng(seed)
    V = random_su(oracle, rng, (o.V, o.dim), 3)
    lat = A.Lattice(3, ext, prec)
    lat.set_links(V)
    bsize = lat.flow_define([s[0] for s in subs], [s[1] for s in subs], [s[2] for s in subs], 2, RHO, DELTA)
    w = (rng.standard_normal(2 * bsize) * scale).astype(np.float32)
    lat.flow_set_weights(w)
    nf = 3 * (o.dim - 1) * (o.dim - 2) // 2  # plaquette and two rectangles per plane
    return o, rng, V, lat, w, bsize, nf


@pytest.mark.parametrize("ext,subs", GEOMS, ids=GEOM_IDS)
def test_forward_matches_oracle(oracle, fo, ext, subs):
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 1, ext=ext, subs=subs)
    lj = lat.flow_apply()
    U = lat.flow_output()
    Uo, ljo = fo.forward(o, V, subs, w, bsize, nf, RHO, DELTA)
    assert np.max(np.abs(U - Uo)) < 1e-12
    assert abs(lj - ljo) < 1e-11 * max(1.0, abs(ljo))
    assert abs(ljo) > 1e-2
    assert np.max(np.abs(U - V)) > 1e-3  # the map is not the identity


@pytest.mark.parametrize("ext,subs", GEOMS, ids=GEOM_IDS)
def test_link_gradient_vs_finite_differences(oracle, fo, ext, subs):
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 2, ext=ext, subs=subs)
    s, lj = lat.flow_gradient(BETA)
    g = lat.flow_link_gradient()
    gens = oracle.generators(3)
    eps = 1e-6
    for trial in range(6):
        x, d = int(rng.integers(o.V)), int(rng.integers(o.dim))
        X = np.einsum("a,aij->ij", rng.standard_normal(8), gens)
        vals = []
        for sgn in (1, -1):
            W = V.copy()
            W[x, d] = scipy.linalg.expm(1j * sgn * eps * X) @ V[x, d]
            vals.append(fo.phi(o, W, subs, w, bsize, nf, RHO, DELTA, BETA))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = 2 * np.trace(g[x, d] @ X).real
        assert abs(fd - an) < 2e-6 * max(1.0, abs(an)), (x, d, fd, an)
    # gradients are Hermitian traceless
    assert np.allclose(g, np.conj(np.swapaxes(g, -1, -2)), atol=1e-14)
    assert np.allclose(np.trace(g, axis1=-2, axis2=-1), 0, atol=1e-12)


@pytest.mark.parametrize("ext,subs", GEOMS, ids=GEOM_IDS)
def test_weight_gradient_vs_finite_differences(oracle, fo, ext, subs):
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 3, ext=ext, subs=subs)
    lat.flow_gradient(BETA, wgrad=True)
    gw = lat.flow_weight_gradient()
    h = np.float32(2.0 ** -9)
    for idx in [0, 5, 32 * nf + 3, 32 * nf + 32 + 7, 32 * nf + 64, bsize + 11, bsize + 32 * nf + 64]:
        vals = []
        for sgn in (1, -1):
            ww = w.copy()
            ww[idx] = np.float32(w[idx] + sgn * h)
            vals.append(fo.phi(o, V, subs, ww, bsize, nf, RHO, DELTA, BETA))
        fd = (vals[0] - vals[1]) / (2 * float(h))
        assert abs(fd - gw[idx]) < 1e-4 * max(1.0, abs(fd)), (idx, fd, gw[idx])


def test_identity_reduces_to_plain_force(oracle):
    o, rng, V, lat, w, bsize, nf = _setup(oracle, 4, scale=0.0)
    s, lj = lat.flow_gradient(BETA)
    assert lj == 0.0
    g = lat.flow_link_gradient()
    o.field("u")[...] = V
    o.field("pi")[...] = 0
    o.updateE(-1.0)  # Pi = F, the plain-HMC force: dS/dt along X = Tr(X F)/N
    F = o.field("Pi")
    assert np.max(np.abs(g - BETA / 6.0 * F)) < 1e-12
    assert abs(s - o.energy_mag()) < 1e-10


def test_flow_hmc_conserves_energy():
    dh = []
    for dt, nd in ((0.04, 5), (0.02, 10)):
        lat = A.Lattice(3, [4, 4, 4, 4], A.FP64)
        lat.haar(7, 0)
        for _ in range(5):
            lat.hmc_trajectory(6.0, 0.05, 10, 8, 1, _)
        lat.flow_define([0, 0, 1, 1, 2, 2, 3, 3], [0, 1] * 4, [0] * 8, 1, 0.3, 1e-3)
        rng = np.random.default_rng(0)
        lat.flow_set_weights((rng.standard_normal(lat.flow_bsize) * 0.3).astype(np.float32))
        r = lat.flow_hmc_trajectory(6.0, dt, nd, 8, 5, 100, reunit=False)
        dh.append(abs(r.dh))
    assert dh[1] < dh[0] and 2.5 < dh[0] / dh[1] < 6.0, dh


def test_autograd_function_trains(tmp_path):
    """FlowAction (torch.autograd.Function over the HIP forward/backward)
    gives torch the same gradient as the C ABI and lowers the loss.  Runs in a
    child process: torch must own the HIP runtime before the library loads."""
    import subprocess
    import sys

    code = r"""
import sys, json, numpy as np, torch
torch.cuda.init()
sys.path.insert(0, %r)
import askja_amd as A
from askja_amd import learned as L
lat = A.Lattice(3, [4, 4, 4, 4], A.FP64)
lat.haar(3, 0)
mu, par, blk = L.sweep_masks(4, 1)
bsize = lat.flow_define(mu, par, blk, 1, 0.3, 1e-3)
w0 = L.init_weights(1, bsize, 9, seed=1, scale=0.3)
w0[32 * 9 + 32: 32 * 9 + 64] = 0.05
fa = L.flow_action_function()
w = torch.tensor(w0, device="cuda", requires_grad=True)
loss = fa.apply(w, lat, 6.0)
loss.backward()
lat.flow_set_weights(w0)
lat.flow_gradient(6.0, wgrad=True)
ref = lat.flow_weight_gradient()
err = float(np.max(np.abs(w.grad.cpu().numpy() - ref)))
V = lat.links()
wt, losses = L.train(lat, [V], 6.0, w0, epochs=15, lr=5e-3)
print(json.dumps({"err": err, "first": losses[0], "last": losses[-1]}))
""" % (str(__import__("conftest").ROOT),)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    import json
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["err"] < 1e-5 * max(1.0, abs(out["first"]))
    assert out["last"] < out["first"]


@pytest.mark.parametrize("kind,rho", [("rank-one", 0.3), ("stout", 0.15)])
def test_trained_map_lowers_heldout_objective(kind, rho):
    """What the trained network gains over the method without it.  Without
    the network the map is the identity (w2 = b2 = 0: rho = 0), and the
    objective E_q[beta S_W(F(V)) - log J(V)] is then beta E_q[S_W(V)]: plain
    HMC.  The objective is KL(q || p~) up to a constant, with q the plain-HMC
    distribution at beta_q = 5.7 and p~ the distribution the learned chain at
    beta = 6.0 samples.  Trained on 8 plain-HMC configurations of a 6^4
    lattice (one sweep, 8 sub-layers), the map must lower it on 4 held-out
    configurations; the gain (nats per configuration) is printed.  Runs in a
    child process: torch must own the HIP runtime before the library loads."""
    import json
    import subprocess
    import sys

    code = r"""
import sys, json, numpy as np, torch
torch.cuda.init()
sys.path.insert(0, %r)
import askja_amd as A
from askja_amd import learned as L
lat = A.Lattice(3, [6, 6, 6, 6], A.FP64)
lat.haar(5, 3)
L.plain_chain(lat, 5.7, 0.02, 50, 8, 5, 0, 50, measure=False)
configs = []
for k in range(12):
    L.plain_chain(lat, 5.7, 0.02, 50, 8, 5, 1000 + 3 * k, 3, measure=False)
    configs.append(lat.links())
mu, par, blk = L.sweep_masks(4, 1)
bsize = lat.flow_define(mu, par, blk, 1, %r, 1e-3, kind=%r)
w0 = L.init_weights(1, bsize, 9, seed=1)
def objective(w, cs):
    lat.flow_set_weights(w)
    out = []
    for V in cs:
        lat.set_links(V)
        s, lj = lat.flow_gradient(6.0)
        out.append(6.0 * s - lj)
    return np.array(out)
wt, losses = L.train(lat, configs[:8], 6.0, w0, epochs=10, lr=2e-3)
u, t = objective(w0, configs[8:]), objective(wt, configs[8:])
print(json.dumps({"untrained": u.tolist(), "gain": (u - t).tolist(), "train": [losses[0], losses[-1]]}))
""" % (str(__import__("conftest").ROOT), rho, kind)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    gain = np.asarray(out["gain"])
    print(f"held-out objective gain of the trained {kind} map: {gain.mean():.3f} nats per configuration "
          f"(per configuration {np.round(gain, 3).tolist()}; untrained objective ~{np.mean(out['untrained']):.0f})")
    assert out["train"][-1] < out["train"][0]
    assert gain.mean() > 0.0
