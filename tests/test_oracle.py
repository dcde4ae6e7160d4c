"""Pins the CPU oracle (the checker) against what fixes each function
independently of the oracle's own code: library generators, closed forms,
invariants, finite differences and published known-answer vectors.

The reference ships no golden vectors or tests (its only fixture is
params_N8S10T10B20k.txt), so every oracle function is pinned here by
mathematics instead.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from conftest import random_hermitian, random_su


# ---------------------------------------------------------------- RNG streams
def test_mt19937_matches_numpy(oracle):
    # numpy's legacy RandomState seeds MT19937 with init_genrand, the same
    # recursion as GSL's mt19937 (gsl_rng_set), and randint over [0, 2^32)
    # returns the raw 32-bit outputs.
    for seed in (1, 120351616, 4357, 2**32 - 1):
        ref = np.random.RandomState(seed).randint(0, 2**32, size=2000, dtype=np.uint64)
        assert np.array_equal(oracle.mt_stream(seed, 2000).astype(np.uint64), ref)
    assert list(oracle.mt_stream(120351616, 3)) == [955617057, 4207408535, 227764342]
    # GSL maps seed 0 to 4357
    assert np.array_equal(oracle.mt_stream(0, 10), oracle.mt_stream(4357, 10))
    # gsl_rng_uniform = x / 2^32
    assert oracle.mt_first_uniform(120351616) == 955617057 / 4294967296.0


def test_polar_gaussian_statistics(oracle):
    g = oracle.gauss_stream(120351616, 200000)
    assert abs(g.mean()) < 0.01
    assert abs(g.var() - 1.0) < 0.01
    assert abs(np.mean(g**4) - 3.0) < 0.06
    from scipy import stats
    assert stats.kstest(g, "norm").pvalue > 1e-3


def test_philox_known_answers(oracle):
    # Random123 known-answer vectors for philox4x32_10 (kat_vectors)
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in cases:
        assert [int(x) for x in oracle.philox4x32_10(ctr, key)] == want


def _ulp_err(got, want):
    return abs(got - want) / math.ulp(want) if want != 0 else abs(got) / math.ulp(1.0)


def test_repro_log_and_sincos_within_2ulp(oracle):
    """The Box-Muller logarithm and sine/cosine (fixed IEEE operation
    sequences, so that the device reproduces them bit for bit) against the C
    library (math.log, math.sin, math.cos): within 2 ulp over the whole range
    the Philox uniforms take, including the end points and the quadrant
    boundaries of 2 pi u."""
    rng = np.random.default_rng(5)
    words = np.concatenate([rng.integers(0, 2**32, 20000), [0, 1, 2, 2**31, 2**32 - 2, 2**32 - 1],
                            np.arange(0, 2**32, 2**28) - 1, np.arange(0, 2**32, 2**30)]).astype(np.uint64) % 2**32
    import mpmath

    mpmath.mp.prec = 120
    worst_l = worst_s = 0.0
    for k, w in enumerate(words):
        u = (float(w) + 0.5) * 2.3283064365386963e-10
        worst_l = max(worst_l, _ulp_err(oracle.repro_log(u), math.log(u)))
        if k % 8:
            continue
        s, c = oracle.repro_sincos2pi(u)
        # true values of the exact angle 2 pi u (u is exact), to 120 bits; the
        # error is counted in ulp of the value, or of 2^-53 where the value is
        # below that scale (near the zeros)
        ang = 2 * mpmath.pi * mpmath.mpf(u)
        for got, want in ((s, float(mpmath.sin(ang))), (c, float(mpmath.cos(ang)))):
            worst_s = max(worst_s, abs(got - want) / max(math.ulp(want), 2.0**-53))
    assert worst_l <= 2.0, worst_l
    assert worst_s <= 2.0, worst_s
    # exact special values
    assert oracle.repro_log(1.0) == 0.0
    assert oracle.repro_log(0.5) == -math.log(2.0) or _ulp_err(oracle.repro_log(0.5), -math.log(2.0)) <= 1
    assert oracle.repro_sincos2pi(0.0) == (0.0, 1.0)
    assert oracle.repro_sincos2pi(0.25) == (1.0, -0.0) or oracle.repro_sincos2pi(0.25) == (1.0, 0.0)


def test_philox_normals_statistics(oracle):
    z = np.concatenate([oracle.philox_normals(7, 3, 1, link, 64) for link in range(2000)])
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
    # distinct links / counters / purposes give distinct streams
    a = oracle.philox_normals(7, 3, 1, 5, 8)
    assert not np.allclose(a, oracle.philox_normals(7, 4, 1, 5, 8))
    assert not np.allclose(a, oracle.philox_normals(7, 3, 2, 5, 8))
    assert not np.allclose(a, oracle.philox_normals(7, 3, 1, 6, 8))


# ---------------------------------------------------------------- SU(N) algebra
@pytest.mark.parametrize("n", [2, 3, 4, 8])
def test_generators_closed_form(oracle, n):
    g = oracle.generators(n)
    assert g.shape == (n * n - 1, n, n)
    for a in range(n * n - 1):
        assert np.allclose(g[a], g[a].conj().T, atol=0)
        assert abs(np.trace(g[a])) < 1e-15
        for b in range(n * n - 1):
            assert abs(np.trace(g[a] @ g[b]) - (0.5 if a == b else 0.0)) < 1e-14
    # ordering: (0,1) symmetric, (0,1) antisymmetric, ..., Cartan last
    assert g[0][0, 1] == 0.5 and g[0][1, 0] == 0.5
    assert g[1][0, 1] == -0.5j and g[1][1, 0] == 0.5j
    last = np.diag([1.0] * (n - 1) + [-(n - 1.0)]) / np.sqrt(2.0 * n * (n - 1))
    assert np.allclose(g[-1], last, atol=1e-15)


@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_matmul_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    assert np.allclose(oracle.matmul(a, b), a @ b, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [2, 3, 8])
@pytest.mark.parametrize("level", [0, 1, 2, 5, 8, 11])
def test_taylor_exp_vs_expm(oracle, n, level):
    rng = np.random.default_rng(100 * n + level)
    h = random_hermitian(rng, (), n, 0.05)
    got = oracle.mexpi(h, level)
    L = 8 if level > 8 or level < 0 else level
    want = scipy.linalg.expm(1j * h)
    nrm = np.linalg.norm(h, 2)
    bound = nrm ** (L + 1) / math.factorial(L + 1) * np.exp(nrm) * 1.01 + 1e-14
    assert np.linalg.norm(got - want, 2) <= bound
    # the truncated series itself (exact Horner value up to rounding)
    series = sum(np.linalg.matrix_power(1j * h, k) / math.factorial(k) for k in range(L + 1))
    assert np.allclose(got, series, rtol=0, atol=1e-14)
    if level == 0:
        assert np.array_equal(got, np.eye(n))
    if level == 1:
        assert np.array_equal(got, np.eye(n) + 1j * h)


def test_taylor_exp_nilpotent_exact(oracle):
    # strictly upper triangular: the series terminates, T_L = expm for L >= n-1
    n = 4
    a = np.triu(np.arange(1, n * n + 1).reshape(n, n).astype(float), 1) * 0.125
    got = oracle.mexpi(-1j * a, 8)  # exp(i * (-i a)) = exp(a)
    assert np.allclose(got, scipy.linalg.expm(a), rtol=0, atol=1e-14)


@pytest.mark.parametrize("n", [2, 3, 6])
def test_tah_invariants(oracle, n):
    rng = np.random.default_rng(n)
    w = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    f = oracle.tah(w)
    assert np.allclose(f, f.conj().T, atol=1e-15)
    assert abs(np.trace(f)) < 1e-14
    # -i/2 [(W - W^+) - Tr(W - W^+)/N]
    d = w - w.conj().T
    assert np.allclose(f, -0.5j * (d - np.trace(d) / n * np.eye(n)), atol=1e-14)
    # Hermitian W and multiples of i*1 project to zero
    assert np.allclose(oracle.tah(w + w.conj().T), 0, atol=1e-15)
    assert np.allclose(oracle.tah(1j * np.eye(n)), 0, atol=1e-15)
    # small-angle: TAH(exp(iH)) = H - Tr H/N + O(H^2)
    h = random_hermitian(rng, (), n, 1e-4)
    h0 = h - np.trace(h) / n * np.eye(n)
    assert np.allclose(oracle.tah(scipy.linalg.expm(1j * h)), h0, atol=1e-9)


@pytest.mark.parametrize("n", [2, 3, 5])
def test_project_sun(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    u = oracle.project_sun(a)
    assert np.allclose(u @ u.conj().T, np.eye(n), atol=1e-14)
    assert abs(np.linalg.det(u) - 1) < 1e-13
    assert np.allclose(oracle.project_sun(u), u, atol=1e-14)  # idempotent
    assert abs(oracle.det(a) - np.linalg.det(a)) < 1e-12 * abs(np.linalg.det(a))


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("ext", [[4, 3, 2], [10, 10, 1, 1], [2, 5], [3]])
def test_geometry_brute_force(oracle, ext):
    o = oracle.Oracle(2, ext)
    V = int(np.prod(ext))
    coords = list(np.ndindex(*ext))  # x[D-1] fastest
    for i, x in enumerate(coords):
        assert o.getind(list(x)) == i
        assert list(o.getcoord(i)) == list(x)
        for d in range(len(ext)):
            xp = list(x)
            xp[d] = (xp[d] + 1) % ext[d]
            xm = list(x)
            xm[d] = (xm[d] - 1) % ext[d]
            assert o.getplus(i, d) == coords.index(tuple(xp))
            assert o.getminus(i, d) == coords.index(tuple(xm))
    assert o.V == V


# ---------------------------------------------------------------- physics
def _random_state(oracle, n, ext, seed, beta=4.0):
    o = oracle.Oracle(n, ext, beta=beta, level=8)
    rng = np.random.default_rng(seed)
    u = random_su(oracle, rng, (o.V, o.dim), n)
    o.field("u")[...] = u
    o.field("U")[...] = u
    return o, rng


def test_cold_start(oracle):
    o = oracle.Oracle(3, [4, 4, 2], beta=5.0)
    o.initializeU()
    assert o.energy_mag() == 0.0
    assert o.uviol() == 0.0
    assert o.wilson_origin() == 1.0
    o.field("pi")[...] = 0
    o.updateE(0.1)
    assert np.all(o.field("Pi") == 0)


@pytest.mark.parametrize("n,ext", [(2, [3, 4, 2]), (3, [4, 2, 3, 2])])
def test_force_is_gradient_of_action(oracle, n, ext):
    """d/dt S(exp(itX) U_l) at t=0 = (1/N) Tr(X F_l), F from the kick with pi=0,
    eps=-1 -- the relation that makes H = beta(S + sum Tr Pi^2/2N) conserved."""
    o, rng = _random_state(oracle, n, ext, 11)
    o.field("pi")[...] = 0
    o.updateE(-1.0)
    F = o.field("Pi").copy()
    u0 = o.field("u").copy()
    for trial in range(4):
        site, d = int(rng.integers(o.V)), int(rng.integers(o.dim))
        X = random_hermitian(rng, (), n)
        X -= np.trace(X) / n * np.eye(n)
        h = 1e-5
        vals = []
        for t in (h, -h):
            o.field("u")[...] = u0
            o.field("u")[site, d] = scipy.linalg.expm(1j * t * X) @ u0[site, d]
            vals.append(o.energy_mag())
        fd = (vals[0] - vals[1]) / (2 * h)
        an = np.trace(X @ F[site, d]).real / n
        assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (fd, an)
    o.field("u")[...] = u0
    # F is Hermitian and traceless
    assert np.allclose(F, np.conj(np.swapaxes(F, -1, -2)), atol=1e-14)
    assert np.allclose(np.trace(F, axis1=-2, axis2=-1), 0, atol=1e-13)


def test_gauge_invariance_and_covariance(oracle):
    n, ext = 3, [3, 2, 4]
    o, rng = _random_state(oracle, n, ext, 5)
    o.field("pi")[...] = 0
    s0 = o.energy_mag()
    p0 = o.polyakov_field().copy()
    o.updateE(-1.0)
    F0 = o.field("Pi").copy()
    g = random_su(oracle, rng, (o.V,), n)
    u = o.field("u")
    for i in range(o.V):
        for d in range(o.dim):
            u[i, d] = g[i] @ u[i, d] @ g[o.getplus(i, d)].conj().T
    assert abs(o.energy_mag() - s0) < 1e-11
    assert np.allclose(o.polyakov_field(), p0, atol=1e-13)
    o.updateE(-1.0)
    F1 = o.field("Pi")
    for i in range(o.V):
        for d in range(o.dim):
            assert np.allclose(F1[i, d], g[i] @ F0[i, d] @ g[i].conj().T, atol=1e-12)


def test_initializeE_normalisation(oracle):
    """sum Tr pi^2 = (N/beta) sum xi^2 with the deviates consumed site -> dir -> a."""
    n, ext, beta, seed = 3, [2, 3, 2], 5.9, 120351616
    o = oracle.Oracle(n, ext, beta=beta)
    o.initializeE(seed)
    xi = oracle.gauss_stream(seed, o.V * o.dim * (n * n - 1))
    assert abs(o.simple_energy_el("pi") - n / beta * np.sum(xi**2)) < 1e-11 * np.sum(xi**2)
    pi = o.field("pi")
    gens = oracle.generators(n)
    # first link = sqrt(2N/beta) sum_a xi_a T^a
    want = np.sqrt(2 * n / beta) * np.einsum("a,aij->ij", xi[: n * n - 1], gens)
    assert np.allclose(pi[0, 0], want, atol=1e-15)
    want_last = np.sqrt(2 * n / beta) * np.einsum("a,aij->ij", xi[-(n * n - 1):], gens)
    assert np.allclose(pi[-1, -1], want_last, atol=1e-15)


def test_momentum_philox_normalisation(oracle):
    n, ext, beta = 4, [2, 2, 3], 3.0
    o = oracle.Oracle(n, ext, beta=beta)
    o.mom_philox(9, 2)
    pi = o.field("pi")
    gens = oracle.generators(n)
    for i in (0, 5, o.V - 1):
        for d in range(o.dim):
            xi = oracle.philox_normals(9, 2, 1, d * o.V + i, n * n - 1)
            want = np.sqrt(2 * n / beta) * np.einsum("a,aij->ij", xi, gens)
            assert np.allclose(pi[i, d], want, atol=1e-14)


def test_energy_conservation_scaling_and_reversibility(oracle):
    n, ext, beta = 3, [4, 4, 4], 5.5
    dh = []
    for dt, nd in ((0.04, 10), (0.02, 20)):
        o, _ = _random_state(oracle, n, ext, 3, beta)
        o.mom_philox(1, 0)
        h0 = o.hamiltonian("pi")
        o.md_trajectory(dt, nd)
        o.copyDown()
        dh.append(abs(o.hamiltonian("Pi") - h0))
    ratio = dh[0] / dh[1]
    assert 3.0 < ratio < 5.5, ratio  # leapfrog: dH ~ dt^2
    # forward, flip momenta, forward again: back to the start
    o, _ = _random_state(oracle, n, ext, 4, beta)
    u0 = o.field("u").copy()
    o.mom_philox(2, 0)
    p0 = o.field("pi").copy()
    o.md_trajectory(0.02, 15)
    o.copyDown()
    o.field("pi")[...] = -o.field("pi")
    o.md_trajectory(0.02, 15)
    o.copyDown()
    assert np.max(np.abs(o.field("u") - u0)) < 1e-12
    assert np.max(np.abs(o.field("pi") + p0)) < 1e-12


def test_haar_start(oracle):
    n, ext = 3, [4, 4, 4, 4]
    o = oracle.Oracle(n, ext)
    o.haar(5, 0)
    u = o.field("u").reshape(-1, n, n)
    eye = np.eye(n)
    assert np.max(np.abs(u @ np.conj(np.swapaxes(u, -1, -2)) - eye)) < 5e-14
    assert np.max(np.abs(np.linalg.det(u) - 1)) < 1e-13
    tr = np.trace(u, axis1=-2, axis2=-1)
    # Haar on SU(N), N >= 2: E|Tr U|^2 = 1, E Tr U = 0
    assert abs(np.mean(np.abs(tr) ** 2) - 1) < 0.12
    assert abs(np.mean(tr)) < 0.08
    assert o.uviol() < 1e-25


def test_two_dim_single_plaquette(oracle):
    """Plaquettes of U(1)-embedded diagonal links: ReTr P/N = mean_k cos(theta_k)
    with theta the summed link phases (abelian closed form)."""
    n, ext = 2, [3, 3]
    o = oracle.Oracle(n, ext)
    rng = np.random.default_rng(1)
    ph = rng.uniform(-np.pi, np.pi, (o.V, 2))
    u = o.field("u")
    for i in range(o.V):
        for d in range(2):
            u[i, d] = np.diag([np.exp(1j * ph[i, d]), np.exp(-1j * ph[i, d])])
    want = 0.0
    for i in range(o.V):
        th = ph[i, 0] + ph[o.getplus(i, 0), 1] - ph[o.getplus(i, 1), 0] - ph[i, 1]
        want += 1 - np.cos(th)
    assert abs(o.energy_mag() - want) < 1e-12


def test_uviol_gauss_energy_el(oracle):
    n, ext = 3, [3, 3, 2]
    o, rng = _random_state(oracle, n, ext, 8)
    assert o.uviol() < 1e-24
    u = o.field("u")
    u[0, 0] = u[0, 0] * 1.001
    # tr2(U U^+ - 1) for one scaled link: N (1.001^2 - 1)^2
    assert abs(o.uviol() - n * (1.001**2 - 1) ** 2) < 1e-12
    # E_L = 1/4 sum tr2(pi + Pi); with pi = Pi it is sum Tr pi^2
    o.mom_philox(3, 0)
    o.field("Pi")[...] = o.field("pi")
    assert abs(o.energy_el() - o.simple_energy_el("pi")) < 1e-10
    # Gauss law of the sum of P - U^+ P U: zero for P = 0, tr2 of the sum otherwise
    g = 0
    Pi = o.field("Pi")
    for i in range(o.V):
        for d in range(o.dim):
            m = o.getminus(i, d)
            g = g + Pi[i, d] - u[m, d].conj().T @ Pi[m, d] @ u[m, d]
    assert abs(o.gauss_law() - np.trace(g @ g).real) < 1e-10


def test_polyakov_measure_definitions(oracle):
    n, ext = 2, [3, 4, 2]
    o, _ = _random_state(oracle, n, ext, 2)
    pf = o.polyakov_field()
    u = o.field("u")
    # brute force: product along x0 for every spatial site
    want = []
    for s in range(o.V // ext[0]):
        w = np.eye(n, dtype=complex)
        pos = s
        for t in range(ext[0]):
            w = w @ u[pos, 0]
            pos = o.getplus(pos, 0)
        want.append(np.trace(w) / n)
    assert np.allclose(pf, want, atol=1e-14)
    assert abs(o.wilson_origin() - abs(want[0])) < 1e-14
    # measure4d's PolyakovLoop: mean over x1 (x2.. = 0) of |Tr P|/N
    stride = o.V // ext[0] // ext[1]
    want_m = np.mean([abs(want[x1 * stride]) for x1 in range(ext[1])])
    assert abs(o.polyakov_measure() - want_m) < 1e-13
