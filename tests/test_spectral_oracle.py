"""The CPU oracle of the two-dimensional spectral map (oracle/spectral_oracle.py)
pinned by what the mathematics fixes: its log J (Weyl's formula through the
eigenvalues) against a central-difference Jacobian of the link map and
against the closed form the HIP kernels evaluate; the map stays in SU(N), is
the identity at t = 0 and commutes with conjugation; the continuous branch of
phi from LU pivots plus the truncated series (the kernels' rule) equals the
eigenvalue sum; and every sub-layer's inputs avoid its updated links."""
import numpy as np
import pytest

from oracle import spectral_oracle as S


@pytest.mark.parametrize("n", [2, 3, 4, 5])
@pytest.mark.parametrize("t", [0.45, -0.3, 0.1])
def test_logj_weyl_closed_form_and_numerical_jacobian(n, t):
    rng = np.random.default_rng(10 * n + int(100 * t) + 100)
    for _ in range(2):
        W = S.haar(n, rng)
        _, lj = S.spectral_map(W, t)
        assert abs(lj - S.logj_closed_form(W, t)) < 1e-11 * max(1.0, abs(lj))
        assert abs(lj - S.numerical_logdet(W, t)) < 1e-7 * max(1.0, abs(lj))
        assert abs(lj) > 1e-3


def test_logj_catches_a_dropped_term():
    """The mean term of the angle map's determinant is not negligible: without
    it the oracle's log J misses the numerical Jacobian."""
    rng = np.random.default_rng(3)
    W = S.haar(3, rng)
    t = 0.4
    lam = np.linalg.eigvals(W)
    mp = (1 - t * t) / np.abs(1 - t * lam) ** 2
    _, lj = S.spectral_map(W, t)
    assert abs((lj - np.log(np.mean(1 / mp))) - S.numerical_logdet(W, t)) > 1e-3


@pytest.mark.parametrize("n", [2, 3, 8, 64])
def test_map_in_su_n_identity_and_equivariance(n):
    rng = np.random.default_rng(n)
    W, g = S.haar(n, rng), S.haar(n, rng)
    Wp, _ = S.spectral_map(W, 0.35)
    assert np.max(np.abs(Wp @ Wp.conj().T - np.eye(n))) < 1e-12
    assert abs(np.linalg.det(Wp) - 1) < 1e-12
    W0, lj0 = S.spectral_map(W, 0.0)
    assert np.max(np.abs(W0 - W)) < 1e-12 and abs(lj0) < 1e-12
    Wg, _ = S.spectral_map(g @ W @ g.conj().T, 0.35)
    assert np.max(np.abs(Wg - g @ Wp @ g.conj().T)) < 1e-11


def _lu_pivots(B):
    """Pivots of LU without pivoting (the kernels' factorisation)."""
    A = B.astype(complex).copy()
    n = A.shape[0]
    for k in range(n - 1):
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k] / A[k, k], A[k, k + 1:])
    return np.diag(A)


@pytest.mark.parametrize("n", [3, 16, 64])
def test_phase_branch_from_pivots_and_series(n):
    rng = np.random.default_rng(100 + n)
    for t in (0.5, -0.5, 0.27):
        for _ in range(6):
            W = S.haar(n, rng)
            lam = np.linalg.eigvals(W)
            phi = float(np.sum(np.angle(1 - t * lam)))
            W2 = W @ W
            tr = [np.trace(W), np.sum(W * W.T), np.sum(W2 * W.T), np.sum(W2 * W2.T)]
            est = -sum(t ** (k + 1) * tr[k].imag / (k + 1) for k in range(4))
            assert abs(est - phi) < n * abs(t) ** 5 / (5 * (1 - abs(t)))
            piv = _lu_pivots(np.eye(n) - t * W)
            argsum = float(np.sum(np.angle(piv)))
            got = argsum + 2 * np.pi * np.rint((est - argsum) / (2 * np.pi))
            assert abs(got - phi) < 1e-9
            assert abs(np.sum(np.log(np.abs(piv))) - np.sum(np.log(np.abs(1 - t * lam)))) < 1e-10


def test_moebius_angle_map_is_monotone():
    th = np.linspace(-np.pi, np.pi, 2001)
    for t in (0.5, -0.5):
        m = np.unwrap(np.angle((np.exp(1j * th) - t) / (1 - t * np.exp(1j * th))))
        assert np.all(np.diff(m) > 0)
        assert abs((m[-1] - m[0]) - 2 * np.pi) < 1e-9


@pytest.mark.parametrize("ext", [(8, 12), (4, 8, 4), (4, 4, 4, 8)], ids=["D2", "D3", "D4"])
@pytest.mark.parametrize("off", [0, 1, 2, 3])
def test_sublayer_inputs_avoid_updated_links(ext, off):
    for mu in range(len(ext)):
        nu = S.partner(ext, mu)
        upd = {(x, mu) for x in S.active_sites(ext, mu, off)}
        assert len(upd) == int(np.prod(ext)) // 4
        for x, _ in upd:
            staple = {(S.shift(ext, x, mu), nu), (S.shift(ext, x, nu), mu), (x, nu)}
            feats = set()
            y0 = S.shift(ext, x, nu)
            for k in (-1, 0, 1):
                y = S.shift(ext, y0, mu, k)
                feats |= {(y, mu), (S.shift(ext, y, mu), nu), (S.shift(ext, y, nu), mu), (y, nu)}
            assert not (staple | feats) & upd


def test_forward_log_j_sums_sublayers():
    rng = np.random.default_rng(4)
    ext, n = (8, 8), 3
    V = S.random_field(ext, n, rng)
    w = (rng.standard_normal((1, 32 * 3 + 65)) * 0.5).astype(np.float32)
    subs = [(0, 0, 0), (1, 2, 0)]
    U, lj = S.forward(V, ext, subs, w, 0.4)
    U1 = V.copy()
    l1, _ = S.sublayer(U1, ext, 0, 0, w[0], 0.4)
    l2, _ = S.sublayer(U1, ext, 1, 2, w[0], 0.4)
    assert np.array_equal(U, U1) and abs(lj - l1 - l2) < 1e-12
    changed = np.any(np.abs(U - V) > 1e-12, axis=(2, 3))
    assert changed.sum() == 2 * (ext[0] * ext[1] // 4)
