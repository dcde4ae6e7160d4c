"""HIP product path (libaskja.so through the C ABI) against the CPU oracle on
the same seeded inputs.

Bar: FP64_EXACT links and momenta bit-identical to the oracle (the
reference's operation order); FP64 (FMA, factored staples) within 1e-13
relative per kick/drift and 1e-10 absolute per trajectory; FP32 within 2e-5
relative per kick/drift.  Reductions in FP64 within 1e-12 relative.
"""
import numpy as np
import pytest

import askja_amd as A
from conftest import random_hermitian, random_su

pytestmark = pytest.mark.gpu

# (N, extents): register path N <= 3 (incl. odd extents, size-1 directions
# and a volume that is not a multiple of the 64-link tile), workgroup path
# N >= 4 for every size class.  For 4 <= N <= 8 the small lattices run the
# workgroup-per-link kernel and the last three cases (at least 2560 links) the
# row-per-thread kernel; the last one has enough workgroups (1821) for the
# FP64 row kernel's two-wavefront variant.
CASES = [
    (2, [4, 3, 2, 2]),
    (3, [4, 4, 4, 4]),
    (3, [3, 5, 1, 2]),
    (4, [3, 2, 2, 2]),
    (8, [2, 2, 3, 2]),
    (16, [2, 3, 2]),
    (32, [2, 2]),
    (64, [2, 2]),
    (4, [8, 8, 8, 8]),
    (8, [8, 8, 8, 4]),
    (7, [8, 8, 8, 8]),
]
IDS = [f"N{n}-{'x'.join(map(str, e))}" for n, e in CASES]


def _state(oracle, n, ext, seed, beta=5.0, level=8):
    rng = np.random.default_rng(seed)
    o = oracle.Oracle(n, ext, beta=beta, level=level)
    u = random_su(oracle, rng, (o.V, o.dim), n)
    p = random_hermitian(rng, (o.V, o.dim), n, 0.7)
    o.field("u")[...] = u
    o.field("U")[...] = u
    o.field("pi")[...] = p
    o.field("Pi")[...] = p
    return o, u, p


def _gpu(n, ext, prec, u=None, p=None):
    lat = A.Lattice(n, ext, prec)
    if u is not None:
        lat.set_links(u)
    if p is not None:
        lat.set_momenta(p)
    return lat


@pytest.mark.parametrize("n,ext", CASES, ids=IDS)
def test_layout_roundtrip(oracle, n, ext):
    o, u, p = _state(oracle, n, ext, 1)
    lat = _gpu(n, ext, A.FP64, u, p)
    assert np.array_equal(lat.links(), u)
    assert np.array_equal(lat.momenta(), p)
    assert np.array_equal(lat.momenta(prev=True), p)
    lat32 = _gpu(n, ext, A.FP32, u, p)
    assert np.array_equal(lat32.links(), u.astype(np.complex64).astype(np.complex128))


@pytest.mark.parametrize("n,ext", CASES, ids=IDS)
def test_kick_exact_bitwise(oracle, n, ext):
    o, u, p = _state(oracle, n, ext, 2)
    lat = _gpu(n, ext, A.FP64_EXACT, u, p)
    o.updateE(0.037)
    lat.updateE(0.037)
    assert np.array_equal(lat.momenta(), o.field("Pi"))
    assert np.array_equal(lat.momenta(prev=True), p)


@pytest.mark.parametrize("level", [0, 1, 2, 6, 8, 11])
@pytest.mark.parametrize("n,ext", CASES, ids=IDS)
def test_drift_exact_bitwise(oracle, n, ext, level):
    o, u, p = _state(oracle, n, ext, 3, level=level)
    lat = _gpu(n, ext, A.FP64_EXACT, u, p)
    o.updateUandcopyDown(0.05)
    lat.updateUandcopyDown(0.05, level)
    assert np.array_equal(lat.links(), o.field("U"))


@pytest.mark.parametrize("n,ext", CASES, ids=IDS)
def test_kick_drift_fast_and_fp32(oracle, n, ext):
    o, u, p = _state(oracle, n, ext, 4)
    o.updateE(0.05)
    o.updateUandcopyDown(0.05)
    for prec, tol in ((A.FP64, 1e-13), (A.FP32, 2e-5)):
        lat = _gpu(n, ext, prec, u, p)
        lat.updateE(0.05)
        lat.updateUandcopyDown(0.05, 8)
        scale_p = np.max(np.abs(o.field("pi")))
        assert np.max(np.abs(lat.momenta() - o.field("pi"))) <= tol * scale_p * n
        assert np.max(np.abs(lat.links() - o.field("u"))) <= tol * n


@pytest.mark.parametrize("n,ext", [(9, [2, 2, 2]), (12, [2, 3, 2]), (24, [2, 2]), (40, [2, 2])],
                         ids=["N9", "N12", "N24", "N40"])
def test_mfma_kick_drift_padded_sizes(oracle, n, ext):
    """N >= 9 runs the link products on MFMA (FP64 16x16x4 / FP32 16x16x4
    tiles) with the matrix zero-padded to a multiple of 16; the padding must
    not leak into the result."""
    o, u, p = _state(oracle, n, ext, 12)
    o.updateE(0.05)
    o.updateUandcopyDown(0.05)
    for prec, tol in ((A.FP64, 1e-13), (A.FP32, 2e-5)):
        lat = _gpu(n, ext, prec, u, p)
        lat.kick_drift(0.05, 0.05, 8)
        scale_p = np.max(np.abs(o.field("pi")))
        assert np.max(np.abs(lat.momenta() - o.field("pi"))) <= tol * scale_p * n, prec
        assert np.max(np.abs(lat.links() - o.field("u"))) <= tol * n, prec
        # the standalone drift (its own MFMA kernel for N >= 9)
        lat = _gpu(n, ext, prec, u, p)
        lat.updateE(0.05)
        lat.updateUandcopyDown(0.05, 8)
        assert np.max(np.abs(lat.links() - o.field("u"))) <= tol * n, prec


@pytest.mark.parametrize("prec", [A.FP64_EXACT, A.FP64, A.FP32])
@pytest.mark.parametrize("n,ext", CASES[:5], ids=IDS[:5])
def test_fused_kick_drift_equals_separate(oracle, n, ext, prec):
    o, u, p = _state(oracle, n, ext, 5)
    a = _gpu(n, ext, prec, u, p)
    b = _gpu(n, ext, prec, u, p)
    a.updateE(0.03)
    a.updateUandcopyDown(0.04, 6)
    b.kick_drift(0.03, 0.04, 6)
    if prec == A.FP64_EXACT:
        assert np.array_equal(a.links(), b.links())
        assert np.array_equal(a.momenta(), b.momenta())
    else:
        tol = 1e-14 if prec == A.FP64 else 1e-6
        assert np.max(np.abs(a.links() - b.links())) <= tol * n
        assert np.max(np.abs(a.momenta() - b.momenta())) <= tol * n * 4


@pytest.mark.parametrize("n,ext", CASES[:6], ids=IDS[:6])
def test_md_trajectory_exact_and_fast(oracle, n, ext):
    o, u, p = _state(oracle, n, ext, 6, beta=4.0)
    h0 = o.hamiltonian("pi")
    o.md_trajectory(0.04, 6)
    o.copyDown()
    lat = _gpu(n, ext, A.FP64_EXACT, u, p)
    e = lat.md_trajectory(0.04, 6, 8)
    assert np.array_equal(lat.links(), o.field("u"))
    assert np.array_equal(lat.momenta(), o.field("pi"))
    beta, n2 = 4.0, 2.0 * n
    assert abs(beta * (e[0] + e[1] / n2) - h0) < 1e-12 * abs(h0)
    assert abs(beta * (e[2] + e[3] / n2) - o.hamiltonian("pi")) < 1e-12 * abs(h0)
    fast = _gpu(n, ext, A.FP64, u, p)
    fast.md_trajectory(0.04, 6, 8)
    assert np.max(np.abs(fast.links() - o.field("u"))) < 1e-10


@pytest.mark.parametrize("n,ext", CASES, ids=IDS)
def test_reductions(oracle, n, ext):
    o, u, p = _state(oracle, n, ext, 7)
    o.updateE(0.02)
    for prec in (A.FP64_EXACT, A.FP64, A.FP32):
        tol = 1e-12 if prec != A.FP32 else 2e-5
        lat = _gpu(n, ext, prec, u, p)
        lat.updateE(0.02)
        t, s = lat.plaq_split()
        ot, os_ = o.plaq_exp()
        assert abs(t - ot) <= tol * max(1.0, abs(ot)) * 10
        assert abs(s - os_) <= tol * max(1.0, abs(os_)) * 10
        assert abs(lat.simpleenergyEl() - o.simple_energy_el("Pi")) <= tol * o.simple_energy_el("Pi") * 10
        assert abs(lat.EnergyEl() - o.energy_el()) <= tol * o.energy_el() * 10
        g, og = lat.GaussLaw(), o.gauss_law()
        assert abs(g - og) <= tol * 1e3 * max(1.0, abs(og))
        pl = lat.wilson_lines(0)
        assert np.max(np.abs(pl - o.polyakov_field())) <= tol * 10
        if prec == A.FP64_EXACT:
            assert abs(lat.WilsonLoop() - o.wilson_origin()) < 1e-14
    # unitarity violation of perturbed links
    o.field("u")[...] = u * 1.0003
    lat = _gpu(n, ext, A.FP64, u * 1.0003, p)
    assert abs(lat.UViol() - o.uviol()) < 1e-9 * o.uviol()


@pytest.mark.parametrize("n,ext", CASES, ids=IDS)
def test_wilson_lines_all_directions(oracle, n, ext):
    o, u, p = _state(oracle, n, ext, 8)
    lat = _gpu(n, ext, A.FP64, u, p)
    V = o.V
    for d in range(len(ext)):
        got = lat.wilson_lines(d)
        # brute force: lines with x_d = 0, lexicographic order of the start
        starts = [i for i, x in enumerate(np.ndindex(*ext)) if x[d] == 0]
        assert len(got) == len(starts) == V // ext[d]
        for q, s in enumerate(starts):
            w = np.eye(n, dtype=complex)
            pos = s
            for _ in range(ext[d]):
                w = w @ u[pos, d]
                pos = o.getplus(pos, d)
            assert abs(got[q] - np.trace(w) / n) < 1e-12


@pytest.mark.parametrize("n,ext", [(3, [2, 3, 2, 2]), (8, [2, 2, 2, 2]), (2, [3, 1, 2])])
def test_mom_gsl_bitwise(oracle, n, ext):
    beta, seed = 5.9, 120351616
    o = oracle.Oracle(n, ext, beta=beta)
    o.initializeE(seed)
    for prec in (A.FP64, A.FP64_EXACT):
        lat = _gpu(n, ext, prec)
        lat.mom_gsl(seed, beta)
        assert np.array_equal(lat.momenta(), o.field("pi"))


@pytest.mark.parametrize("n,ext", CASES, ids=IDS)
def test_mom_refresh_and_haar_vs_oracle(oracle, n, ext):
    beta = 4.5
    o = oracle.Oracle(n, ext, beta=beta)
    o.mom_philox(11, 3)
    o.haar(12, 1)
    lat = _gpu(n, ext, A.FP64)
    lat.mom_refresh(11, 3, beta)
    lat.haar(12, 1)
    # the Box-Muller steps are IEEE operations in a fixed order on both
    # sides: the device's Philox momenta equal the oracle's bit for bit
    assert np.array_equal(lat.momenta(), o.field("pi"))
    ex = _gpu(n, ext, A.FP64_EXACT)
    ex.mom_refresh(11, 3, beta)
    assert np.array_equal(ex.momenta(), o.field("pi"))
    assert np.max(np.abs(lat.links() - o.field("u"))) < 1e-12
    # reunitarisation of perturbed links
    rng = np.random.default_rng(0)
    u = o.field("u") + 1e-3 * (rng.standard_normal(o.field("u").shape) + 1j * rng.standard_normal(o.field("u").shape))
    o.field("u")[...] = u
    o.reunitarize()
    lat.set_links(u)
    lat.reunitarize()
    assert np.max(np.abs(lat.links() - o.field("u"))) < 1e-12
    assert lat.UViol() < 1e-24 * o.V * n


def test_snapshot_restore(oracle):
    n, ext = 3, [4, 4, 2, 2]
    o, u, p = _state(oracle, n, ext, 9)
    lat = _gpu(n, ext, A.FP64, u, p)
    lat.takeconfiguration()
    lat.kick_drift(0.1, 0.1, 8)
    assert not np.array_equal(lat.links(), u)
    lat.loadconfiguration()
    assert np.array_equal(lat.links(), u)


@pytest.mark.parametrize("prec,tol", [(A.FP64, 1e-11), (A.FP32, 5e-5)])
def test_reversibility(oracle, prec, tol):
    n, ext = 3, [4, 4, 4, 4]
    lat = _gpu(n, ext, prec)
    lat.haar(3, 0)
    u0 = lat.links()
    lat.mom_refresh(5, 0, 5.9)
    p0 = lat.momenta()
    lat.md_trajectory(0.02, 25, 8)
    lat.mom_negate()
    lat.md_trajectory(0.02, 25, 8)
    assert np.max(np.abs(lat.links() - u0)) < tol
    assert np.max(np.abs(lat.momenta() + p0)) < tol * 10


def test_hmc_chain_vs_oracle(oracle):
    n, ext, beta, dt, nd, seed = 3, [4, 4, 4, 4], 5.7, 0.05, 8, 17
    o = oracle.Oracle(n, ext, beta=beta, level=8)
    o.haar(21, 0)
    rec = o.hmc(dt, nd, 6, seed, first_counter=0, reunit=True)
    lat = _gpu(n, ext, A.FP64)
    lat.haar(21, 0)
    for t in range(6):
        r = lat.hmc_trajectory(beta, dt, nd, 8, seed, t, reunit=True)
        assert abs(r.dh - rec[t, 0]) < 1e-8, (t, r.dh, rec[t, 0])
        assert r.accepted == int(rec[t, 1])
        plaq = 1.0 - r.action / lat.n_plaquettes()
        assert abs(plaq - rec[t, 2]) < 1e-12
    assert np.max(np.abs(lat.links() - o.field("u"))) < 1e-10
    pl = lat.polyakov()
    assert abs(pl - complex(rec[-1, 3], rec[-1, 4])) < 1e-12


def test_hmc_chain_exact_bitwise(oracle):
    """Six trajectories of the exact mode from the oracle's Haar links: the
    Philox momenta, the MD (the reference's operation order) and the
    Metropolis decisions reproduce the oracle's chain, and the final links
    are bit-identical (no reunitarisation, which rounds differently)."""
    n, ext, beta, dt, nd, seed = 3, [4, 4, 4, 4], 5.7, 0.05, 8, 17
    o = oracle.Oracle(n, ext, beta=beta, level=8)
    o.haar(21, 0)
    u0 = o.field("u").copy()
    rec = o.hmc(dt, nd, 6, seed, first_counter=0, reunit=False)
    lat = _gpu(n, ext, A.FP64_EXACT, u0)
    for t in range(6):
        r = lat.hmc_trajectory(beta, dt, nd, 8, seed, t, reunit=False)
        assert r.accepted == int(rec[t, 1]), t
        assert abs(r.dh - rec[t, 0]) < 1e-8, (t, r.dh, rec[t, 0])
    assert rec[:, 1].sum() >= 1
    assert np.array_equal(lat.links(), o.field("u"))


def test_nan_trajectory_rejected_links_restored():
    """A trajectory whose energies overflow (dt = 1000: the links reach inf
    and NaN within a few steps) is rejected, as the reference's accept tests
    both fail for a NaN H (sun.cpp:1258, 1288, 1304-1310), and the chain
    state is the one before it, bit for bit; the next sane trajectory runs
    normally."""
    for prec in (A.FP64, A.FP32):
        lat = _gpu(3, [4, 4, 4, 4], prec)
        lat.haar(3, 0)
        u0 = lat.links()
        r = lat.hmc_trajectory(5.9, 1000.0, 10, 6, 5, 0, reunit=False)
        assert not np.isfinite(r.dh) and not r.accepted and r.accept_case == 0
        assert np.array_equal(lat.links(), u0)
        r = lat.hmc_trajectory(5.9, 0.02, 10, 6, 5, 1, reunit=False)
        assert np.isfinite(r.dh) and abs(r.dh) < 1.0


def test_errors_are_reported(oracle):
    with pytest.raises(A.AskjaError):
        A.Lattice(65, [2, 2])
    lat = _gpu(3, [2, 2, 2, 2], A.FP64)
    buf = np.zeros(64)
    rc = A.lib().askja_wilson_lines(lat._h, 7, buf.ctypes.data_as(A._DP))
    assert rc == -1  # ASKJA_EINVAL
    assert b"out of range" in A.lib().askja_last_error(lat._h)
    # sticky until cleared
    with pytest.raises(A.AskjaError, match="ESTICKY"):
        lat.updateE(0.1)
    A.lib().askja_clear_error(lat._h)
    lat.updateE(0.1)


@pytest.mark.parametrize("n,ext", [(12, [2, 2, 2]), (16, [2, 3, 2]), (24, [2, 2]), (64, [2, 2])],
                         ids=["N12", "N16", "N24", "N64"])
def test_bf16x3_kick_drift(oracle, n, ext):
    """Split-precision BF16x3 MFMA link products (FP32 accumulation): about
    16 significant bits per product, so the kick and drift agree with the FP64
    oracle to ~1e-4 relative, and with the FP32 VALU path as closely."""
    o, u, p = _state(oracle, n, ext, 10)
    o.updateE(0.05)
    o.updateUandcopyDown(0.05)
    for prec in (A.BF16X3, A.FP32):
        lat = _gpu(n, ext, prec, u, p)
        lat.updateE(0.05)
        lat.updateUandcopyDown(0.05, 8)
        scale_p = np.max(np.abs(o.field("pi")))
        dp = np.max(np.abs(lat.momenta() - o.field("pi"))) / scale_p
        du = np.max(np.abs(lat.links() - o.field("u")))
        assert dp < 2e-4 * n and du < 2e-4 * n, (prec, dp, du)
    # the fused kick+drift on MFMA equals kick then drift to rounding
    a = _gpu(n, ext, A.BF16X3, u, p)
    b = _gpu(n, ext, A.BF16X3, u, p)
    a.updateE(0.03)
    a.updateUandcopyDown(0.04, 6)
    b.kick_drift(0.03, 0.04, 6)
    assert np.max(np.abs(a.links() - b.links())) < 2e-4 * n
    # HMC with BF16x3 MD stays exact through the FP64 Metropolis test
    c = _gpu(n, ext, A.BF16X3)
    c.haar(1, 0)
    r = c.hmc_trajectory(n * n / 2.0, 0.01, 5, 8, 2, 0)
    assert np.isfinite(r.dh) and abs(r.dh) < 5.0


@pytest.mark.parametrize("n,ext,prec", [(3, [4, 4, 4, 4], A.FP32), (5, [2, 2, 2, 2], A.FP32),
                                        (12, [2, 2, 2], A.BF16X3), (16, [2, 3, 2], A.FP32)],
                         ids=["N3-fp32", "N5-fp32", "N12-bf16x3", "N16-fp32"])
def test_reduced_precision_metropolis_energy_is_fp64(n, ext, prec):
    """In the FP32 and BF16x3 modes the MD runs in reduced precision, but the
    action in H_old / H_new is the FP64-arithmetic action of the stored links
    (not the MD's reduced-precision ReTr W), so the Metropolis test is exact
    for the stored chain state."""
    beta = 2.0 * n
    lat = _gpu(n, ext, prec)
    lat.haar(3, 0)
    for t in range(3):
        r = lat.hmc_trajectory(beta, 0.01, 5, 8, 2, t)
        ref = _gpu(n, ext, A.FP64, lat.links())  # the stored FP32 links, exactly
        s_ref = ref.EnergyMag()
        assert abs(r.action - s_ref) < 1e-12 * lat.n_plaquettes(), (t, r.action, s_ref)
    e = lat.md_trajectory(0.01, 3, 8)
    ref = _gpu(n, ext, A.FP64, lat.links())
    assert abs(e[2] - ref.EnergyMag()) < 1e-12 * lat.n_plaquettes()


def test_example_params_regime_dim10(oracle):
    """The reference's own example (params_N8S10T10B20k.txt): NC 8, DIM 10,
    10 x 10 x 1^8, beta 20000, LEVEL 8.  Eight size-1 directions make every
    neighbour step in them a self-reference; the exact mode must still be
    bit-identical to the oracle, and the fast mode within the FP64 bar."""
    n, ext = 8, [10, 10] + [1] * 8
    o, u, p = _state(oracle, n, ext, 21, beta=20000.0, level=8)
    lat = _gpu(n, ext, A.FP64_EXACT, u, p)
    for _ in range(3):
        o.updateE(0.01)
        o.updateUandcopyDown(0.01)
        lat.updateE(0.01)
        lat.updateUandcopyDown(0.01, 8)
    assert np.array_equal(lat.momenta(), o.field("pi"))
    assert np.array_equal(lat.links(), o.field("u"))
    fast = _gpu(n, ext, A.FP64, u, p)
    for _ in range(3):
        fast.kick_drift(0.01, 0.01, 8)
    assert np.max(np.abs(fast.links() - o.field("u"))) < 1e-12


@pytest.mark.parametrize("n,ext", [(3, [4, 3, 2, 2]), (2, [3, 4, 2]), (4, [2, 3, 3, 2])])
def test_measure_definitions_vs_oracle(oracle, n, ext):
    """measure4d's PolyakovLoop and WilsonLoop(evs, d) from the GPU lines
    against the oracle's restatement (incl. its loop-variable quirk)."""
    o, u, p = _state(oracle, n, ext, 31)
    lat = _gpu(n, ext, A.FP64, u, p)
    assert abs(lat.polyakov_measure() - o.polyakov_measure()) < 1e-13
    for d in range(1, len(ext)):
        assert abs(lat.wilson_measure(d) - o.wilson_measure(d)) < 1e-13


def test_async_download_into_pinned_buffers(oracle):
    """askja_links_download_async (conversion on the context's stream, copy on
    a second stream into page-locked memory) gives the synchronous download's
    field, also when a second download is queued behind MD work."""
    import ctypes as C

    n, ext = 3, [4, 4, 2, 2]
    lat = _gpu(n, ext, A.FP64)
    lat.haar(5, 0)
    nd = A.lib().askja_field_doubles(lat._h)
    bufs = []
    for _ in range(2):
        p = C.POINTER(C.c_double)()
        assert A.lib().askja_host_alloc(nd, C.byref(p)) == 0
        bufs.append(p)
    try:
        want0 = lat.links()
        assert A.lib().askja_links_download_async(lat._h, bufs[0]) == 0
        lat.mom_refresh(1, 0, 5.0)
        lat.kick_drift(0.02, 0.02, 8)  # queued after the first conversion
        assert A.lib().askja_links_download_async(lat._h, bufs[1]) == 0
        assert A.lib().askja_download_wait(lat._h) == 0
        got0 = np.ctypeslib.as_array(bufs[0], (nd,)).view(np.complex128).reshape(want0.shape)
        got1 = np.ctypeslib.as_array(bufs[1], (nd,)).view(np.complex128).reshape(want0.shape)
        assert np.array_equal(got0, want0)
        assert np.array_equal(got1, lat.links())
        assert not np.array_equal(got0, got1)
    finally:
        for p in bufs:
            A.lib().askja_host_free(p)
