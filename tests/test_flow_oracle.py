"""Pins the CPU restatement of the learned field transformation
(oracle/flow_oracle.py) by mathematics: its log-Jacobian against a
brute-force dense Jacobian of every updated link, gauge equivariance, and the
rho = 0 identity.  (The reference has no counterpart to compare with.)"""
import numpy as np
import pytest
import scipy.linalg

from conftest import random_su


@pytest.fixture(scope="module")
def fo():
    from oracle import flow_oracle
    return flow_oracle


def _setup(oracle, ext, seed, nblocks=1):
    o = oracle.Oracle(3, ext)
    rng = np.random.default_rng(seed)
    V = random_su(oracle, rng, (o.V, o.dim), 3)
    nf = 3 * (o.dim - 1) * (o.dim - 2) // 2  # plaquette and two rectangles per plane
    bsize = 32 * nf + 65
    w = (rng.standard_normal(nblocks * bsize) * 0.5).astype(np.float32)
    return o, rng, V, nf, bsize, w


def test_logj_matches_dense_jacobian(oracle, fo):
    o, rng, V, nf, bsize, w = _setup(oracle, [4, 2, 2, 2], 1)
    gens = oracle.generators(3)
    mu, par = 1, 0
    Up, lj = fo.sub_layer(o, V, mu, par, w, nf, 0.3, 1e-3)
    eps = 1e-5
    total = 0.0
    for x in range(o.V):
        if fo.parity(o.ext, x) != par:
            continue
        J = np.zeros((8, 8))
        for a in range(8):
            outs = []
            for sgn in (1, -1):
                W = V.copy()
                W[x, mu] = scipy.linalg.expm(1j * sgn * eps * gens[a]) @ V[x, mu]
                outs.append(fo.sub_layer(o, W, mu, par, w, nf, 0.3, 1e-3)[0][x, mu])
            L = scipy.linalg.logm(outs[0] @ np.linalg.inv(outs[1])) / (2j * eps)
            for c in range(8):
                J[c, a] = 2 * np.trace(gens[c] @ L).real
        total += np.log(np.linalg.det(J))
    assert abs(total - lj) < 1e-6 * max(1.0, abs(lj)), (total, lj)
    assert abs(lj) > 1e-3  # the test is not vacuous


def test_gauge_equivariance(oracle, fo):
    o, rng, V, nf, bsize, w = _setup(oracle, [4, 2, 2, 4], 2, nblocks=2)
    subs = [(0, 0, 0), (2, 1, 1), (3, 0, 0)]
    U, lj = fo.forward(o, V, subs, w, bsize, nf, 0.3, 1e-3)
    g = random_su(oracle, rng, (o.V,), 3)
    Vg = np.empty_like(V)
    for x in range(o.V):
        for d in range(o.dim):
            Vg[x, d] = g[x] @ V[x, d] @ g[o.getplus(x, d)].conj().T
    Ug, ljg = fo.forward(o, Vg, subs, w, bsize, nf, 0.3, 1e-3)
    for x in range(o.V):
        for d in range(o.dim):
            assert np.allclose(Ug[x, d], g[x] @ U[x, d] @ g[o.getplus(x, d)].conj().T, atol=1e-12)
    assert abs(ljg - lj) < 1e-11


def test_zero_coefficients_are_identity(oracle, fo):
    o, rng, V, nf, bsize, w = _setup(oracle, [2, 2, 2, 2], 3)
    w[:] = 0
    U, lj = fo.forward(o, V, [(0, 0, 0), (1, 1, 0)], w, bsize, nf, 0.3, 1e-3)
    assert np.array_equal(U, V) and lj == 0.0


def test_features_closed_form_and_invariance(oracle, fo):
    """Plaquette and rectangle features: for diagonal (abelian) links the
    rectangle is the product of its two plaquettes, so ReTr R / 3 is the mean
    cosine of the summed plaquette angles; every feature is gauge invariant."""
    o = oracle.Oracle(3, [4, 4, 2, 2])
    rng = np.random.default_rng(5)
    th = rng.uniform(-np.pi, np.pi, (o.V, o.dim, 2))
    ph = np.stack([th[..., 0], th[..., 1], -th[..., 0] - th[..., 1]], -1)
    U = np.zeros((o.V, o.dim, 3, 3), complex)
    for c in range(3):
        U[:, :, c, c] = np.exp(1j * ph[..., c])
    x, mu = 5, 1

    def pang(y, a, b):  # plaquette angles at y in plane (a, b)
        return ph[y, a] + ph[o.getplus(y, a), b] - ph[o.getplus(y, b), a] - ph[y, b]

    _, _, f = fo.link_inputs(o, U, x, mu)
    planes = [(a, b) for a in range(4) for b in range(a + 1, 4) if mu not in (a, b)]
    npl = len(planes)
    assert len(f) == 3 * npl
    for j, (a, b) in enumerate(planes):
        p0 = pang(x, a, b)
        assert abs(f[j] - np.cos(p0).mean()) < 1e-13
        r1 = p0 + pang(o.getplus(x, a), a, b)
        r2 = p0 + pang(o.getplus(x, b), a, b)
        assert abs(f[npl + 2 * j] - np.cos(r1).mean()) < 1e-13
        assert abs(f[npl + 2 * j + 1] - np.cos(r2).mean()) < 1e-13
    # gauge invariance on a random SU(3) field
    V = random_su(oracle, rng, (o.V, o.dim), 3)
    g = random_su(oracle, rng, (o.V,), 3)
    Vg = np.empty_like(V)
    for y in range(o.V):
        for d in range(o.dim):
            Vg[y, d] = g[y] @ V[y, d] @ g[o.getplus(y, d)].conj().T
    assert np.allclose(fo.link_inputs(o, V, x, mu)[2], fo.link_inputs(o, Vg, x, mu)[2], atol=1e-13)


def test_batched_restatement_matches_per_link(oracle):
    """The batched restatement (used at the 16^4 size) equals the per-link
    one: same links to rounding, same log J, same action."""
    from oracle import flow_oracle as fo

    for ext, subs in (([4, 4, 2, 2], [(0, 0, 0), (1, 1, 1), (2, 0, 0), (3, 1, 1), (0, 1, 0)]),
                      ([4, 4, 4], [(0, 0, 0), (2, 1, 1), (1, 0, 0)])):
        o = oracle.Oracle(3, ext)
        rng = np.random.default_rng(3)
        V = random_su(oracle, rng, (o.V, o.dim), 3)
        nf = 3 * (o.dim - 1) * (o.dim - 2) // 2
        bsize = 32 * nf + 65
        w = (rng.standard_normal(2 * bsize) * 0.5).astype(np.float32)
        U1, l1 = fo.forward(o, V, subs, w, bsize, nf, 0.3, 1e-3)
        U2, l2 = fo.forward_batched(ext, V, subs, w, bsize, nf, 0.3, 1e-3)
        assert np.max(np.abs(U1 - U2)) < 1e-13
        assert abs(l1 - l2) < 1e-12 * max(1.0, abs(l1)) and abs(l1) > 1e-3
        assert abs(fo.energy_mag(o, U1) - fo.energy_mag_batched(ext, U2)) < 1e-10
        up, dn, par = fo.neighbours(ext)
        for x in range(o.V):
            for d in range(o.dim):
                assert up[d][x] == o.getplus(x, d) and dn[d][x] == o.getminus(x, d)



def _rsu3(rng):
    a = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    q, r = np.linalg.qr(a)
    q = q @ np.diag(np.diag(r) / np.abs(np.diag(r)))
    return q / np.linalg.det(q) ** (1 / 3)


def test_stout_logdet_vs_numerical_jacobian():
    """The stout link map U -> exp(-i rho TAH(U S)) U: the oracle's log det
    (1 + g(ad_iQ) M in su(3) coordinates, g by its power series) against the
    8 x 8 Jacobian taken by central differences in left-invariant
    coordinates (matrix log of U'(eps) U'(0)^+), for rho of both signs up to
    the invertibility bound 1/6 and six unitary staples."""
    from oracle import flow_oracle as fo

    T = fo.su3_basis()
    rng = np.random.default_rng(11)
    for trial in range(6):
        U = _rsu3(rng)
        S = sum(_rsu3(rng) for _ in range(6))
        rho = (0.165 if trial == 0 else 0.16 * rng.uniform(-1, 1))

        def F(V):
            E, ld = fo.stout_link_batched((V @ S)[None], np.array([rho]))
            return E[0] @ V, ld[0]

        U1, ld = F(U)
        eps = 1e-5
        J = np.zeros((8, 8))
        for b in range(8):
            Up, _ = F(scipy.linalg.expm(1j * eps * T[b]) @ U)
            Um, _ = F(scipy.linalg.expm(-1j * eps * T[b]) @ U)
            X = (scipy.linalg.logm(Up @ U1.conj().T) - scipy.linalg.logm(Um @ U1.conj().T)) / (2j * eps)
            J[:, b] = fo.coords(X[None], T)[0]
        sgn, lnum = np.linalg.slogdet(J)
        assert sgn > 0 and abs(lnum - ld) < 1e-8, (trial, rho, lnum, ld)
        # smearing direction: rho > 0 raises ReTr(U S)
        assert (np.trace(U1 @ S).real > np.trace(U @ S).real) == (rho > 0)


def test_stout_map_equivariance_and_identity(oracle):
    """F(g V g^+) = g F(V) g^+ with the same log J (gauge equivariance), and
    rho = 0 (zero weights) is the identity with log J = 0."""
    from oracle import flow_oracle as fo

    ext = [4, 4, 2, 2]
    o = oracle.Oracle(3, ext)
    rng = np.random.default_rng(5)
    V = random_su(oracle, rng, (o.V, o.dim), 3)
    subs = [(0, 0, 0), (1, 1, 0), (3, 0, 0)]
    nf, bsize = 9, 32 * 9 + 65
    w = (rng.standard_normal(bsize) * 0.5).astype(np.float32)
    U, lj = fo.stout_forward_batched(ext, V, subs, w, bsize, nf, 0.15)
    assert abs(lj) > 1e-3 and np.max(np.abs(U - V)) > 1e-3
    g = random_su(oracle, rng, (o.V,), 3)
    up, dn, _ = fo.neighbours(ext)
    Vg = np.einsum("xij,xdjk,xdlk->xdil", g, V, np.conj(g[up.T]))
    Ug, ljg = fo.stout_forward_batched(ext, Vg, subs, w, bsize, nf, 0.15)
    assert np.max(np.abs(Ug - np.einsum("xij,xdjk,xdlk->xdil", g, U, np.conj(g[up.T])))) < 1e-12
    assert abs(ljg - lj) < 1e-10
    U0, lj0 = fo.stout_forward_batched(ext, V, subs, np.zeros(bsize, np.float32), bsize, nf, 0.15)
    assert np.max(np.abs(U0 - V)) < 1e-14 and lj0 == 0.0
