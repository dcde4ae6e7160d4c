"""The learned map at its workload size: SU(3), D = 4, 16^4, beta 6.0, four
sweep layers (32 masked sub-layers, four MLP weight blocks), rho_max 0.3.

Against the batched restatement of oracle/flow_oracle.py (itself pinned to
the per-link restatement in test_flow_oracle.py): U = F(V) of every link and
log J; the hand-written backward against finite differences of the oracle's
Phi = beta S_W(F(V)) - log J(V) (one sweep, eps = 1e-6); the rho = 0 map
against the plain-HMC force of the C oracle; and, on the GPU alone, that
field-transformation HMC at this size is reversible and its energy error
scales as dt^2."""
import numpy as np
import pytest
import scipy.linalg

import askja_amd as A
from askja_amd import learned as L

pytestmark = pytest.mark.gpu

EXT, BETA, RHO, DELTA = [16, 16, 16, 16], 6.0, 0.3, 1e-3
NF = 9
BSIZE = 32 * NF + 65


@pytest.fixture(scope="module")
def fo():
    from oracle import flow_oracle
    return flow_oracle


def _lattice(layers, seed, scale=0.3):
    lat = A.Lattice(3, EXT, A.FP64)
    lat.haar(seed, 0)
    for t in range(3):  # a few trajectories away from the Haar start
        lat.hmc_trajectory(BETA, 0.02, 20, 8, seed, t)
    mu, par, blk = L.sweep_masks(4, layers)
    assert lat.flow_define(mu, par, blk, layers, RHO, DELTA) == BSIZE
    rng = np.random.default_rng(seed)
    w = (rng.standard_normal(layers * BSIZE) * scale).astype(np.float32)
    lat.flow_set_weights(w)
    return lat, list(zip(mu, par, blk)), w, rng


def test_c4_forward_matches_oracle(fo):
    lat, subs, w, _ = _lattice(4, 1)
    V = lat.links()
    lj = lat.flow_apply()
    U = lat.flow_output()
    Uo, ljo = fo.forward_batched(EXT, V, subs, w, BSIZE, NF, RHO, DELTA)
    print(f"C4 forward: max |U - U_oracle| = {np.max(np.abs(U - Uo)):.2e}, log J = {lj:.6f} (oracle {ljo:.6f})")
    assert np.max(np.abs(U - Uo)) < 1e-11
    assert abs(lj - ljo) < 1e-10 * abs(ljo) and abs(ljo) > 1.0
    assert np.max(np.abs(U - V)) > 1e-3
    s, lj2 = lat.flow_gradient(BETA)
    assert abs(s - fo.energy_mag_batched(EXT, Uo)) < 1e-8 * s
    assert lj2 == lj


def test_c4_link_gradient_vs_finite_differences(fo):
    lat, subs, w, rng = _lattice(1, 2)
    V = lat.links()
    lat.flow_gradient(BETA)
    g = lat.flow_link_gradient()
    import oracle.oracle as O
    gens = O.generators(3)
    eps = 1e-6
    for x, d in ((int(rng.integers(V.shape[0])), 0), (int(rng.integers(V.shape[0])), 3)):
        X = np.einsum("a,aij->ij", rng.standard_normal(8), gens)
        vals = []
        for sgn in (1, -1):
            W = V.copy()
            W[x, d] = scipy.linalg.expm(1j * sgn * eps * X) @ V[x, d]
            U, lj = fo.forward_batched(EXT, W, subs, w, BSIZE, NF, RHO, DELTA)
            vals.append(BETA * fo.energy_mag_batched(EXT, U) - lj)
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = 2 * np.trace(g[x, d] @ X).real
        # Phi ~ 1e6: the central difference carries ~1e-10 * Phi / eps of rounding
        assert abs(fd - an) < 2e-3 + 1e-6 * abs(an), (x, d, fd, an)


def test_c4_identity_map_is_plain_force(oracle):
    lat, subs, w, _ = _lattice(4, 3, scale=0.0)
    V = lat.links()
    s, lj = lat.flow_gradient(BETA)
    assert lj == 0.0
    g = lat.flow_link_gradient()
    o = oracle.Oracle(3, EXT)
    o.field("u")[...] = V
    o.field("pi")[...] = 0
    o.updateE(-1.0)
    assert np.max(np.abs(g - BETA / 6.0 * o.field("Pi"))) < 1e-11
    assert abs(s - o.energy_mag()) < 1e-8 * s


def test_c4_fthmc_energy_error_scales_as_dt_squared():
    lat, subs, w, _ = _lattice(4, 4)
    V0 = lat.links()
    dh = []
    for dt, nd in ((0.04, 5), (0.02, 10)):
        lat.set_links(V0)
        r = lat.flow_hmc_trajectory(BETA, dt, nd, 8, 7, 11, reunit=False)
        dh.append(abs(r.dh))
    print(f"C4 FTHMC |dH| at dt 0.04 / 0.02 (length 0.2): {dh}")
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_c4_fthmc_reversible():
    """Leapfrog in V with the flow force, forward 5 steps of 0.02, momenta
    negated, 5 steps back: V returns to its start (host-side kicks with the
    library's flow gradient, the library's drift)."""
    lat, subs, w, rng = _lattice(4, 5)
    V0 = lat.links()
    lat.mom_refresh(9, 0, BETA)
    cf = 2 * 3 / BETA

    def kick(eps):
        lat.flow_gradient(BETA)
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
