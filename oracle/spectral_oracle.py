"""CPU restatement (numpy, FP64) of the spectral field transformation
U = F(V) for SU(N), any N >= 2, in any dimension D >= 2.

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels of
askja_amd/csrc/flow_spectral.hip.  The reference has no counterpart (the learned
proposal is an addition to its method, whose action is EnergyMag,
sun.cpp:507-537); the definitions are these.

Extents (L0, ..., L_{D-1}), sites lexicographic (x_{D-1} fastest), periodic.
P_mu,nu(y) = U_mu(y) U_nu(y+mu) U_mu(y+nu)^+ U_nu(y)^+ is the plaquette at y.
Sub-layer (mu, off), nu = (mu + 1) mod D (= 1 - mu in D = 2): the links
U_mu(x) with x_nu = off (mod 4) are updated, every other link is frozen.
For an updated link:
    S  = U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+     (frozen: x+nu is in row off+1)
    W  = U_mu(x) S                               (the plaquette through the link)
    f  = ReTr P_mu,nu(x+nu+k mu) / N, k = -1, 0, 1  (frozen plaquettes, row off+1)
    t  = t_max tanh(b2 + w2 . tanh(W1 f + b1))    (|t| < 1)
    W' = the spectral map of W: each eigenvalue lambda of W goes to
         e^{2 i phi / N} (lambda - t) / (1 - t lambda),
         phi = sum_a arg(1 - t lambda_a)  (each term in (-pi/2, pi/2), so that
         det W' = 1 and the map is continuous)
    U'_mu(x) = W' S^+.
The map is the identity at t = 0 and a diffeomorphism for |t| < 1 (the
Moebius angle map is monotone).  Its Jacobian against the Haar measure, from
Weyl's integration formula (eigenvectors fixed, angles theta -> theta'):
    log J = sum_a log m'_a + log(mean_a 1/m'_a)
            + sum_{a<b} [log|lambda'_a - lambda'_b|^2 - log|lambda_a - lambda_b|^2],
    m'_a = (1 - t^2) / |1 - t lambda_a|^2
(the mean term is the determinant of the angle map restricted to
sum dtheta = 0).  This module evaluates everything through numpy's
eigen-decomposition; the HIP kernels use the equivalent closed form without
eigenvalues, log J = N^2 log(1-t^2) - 2N log|det(1 - tW)|
+ log((1 + t^2 - (2t/N) ReTr W) / (1 - t^2)), and that equivalence is what
the tests check.

Weight block layout (float32): W1[32][3], b1[32], w2[32], b2.  Fields are
complex128 arrays [V, D, N, N] in the reference's U-*.conf order.
"""
import numpy as np

HID = 32
NF = 3


def coords(ext, i):
    return tuple(int(c) for c in np.unravel_index(i, ext))


def site(ext, x):
    return int(np.ravel_multi_index(tuple(c % e for c, e in zip(x, ext)), ext))


def shift(ext, i, d, k=1):
    x = list(coords(ext, i))
    x[d] += k
    return site(ext, x)


def partner(ext, mu):
    """The direction nu of sub-layers of direction mu."""
    return (mu + 1) % len(ext)


def plaquette(U, ext, y, mu=0, nu=1):
    """P_mu,nu(y) = U_mu(y) U_nu(y+mu) U_mu(y+nu)^+ U_nu(y)^+."""
    return U[y, mu] @ U[shift(ext, y, mu), nu] @ U[shift(ext, y, nu), mu].conj().T @ U[y, nu].conj().T


def wilson_action(U, ext):
    """S_W = sum_y sum_{mu < nu} (1 - ReTr P_mu,nu(y) / N)  (EnergyMag, sun.cpp:507-537)."""
    n, D = U.shape[-1], len(ext)
    return float(sum(1.0 - np.trace(plaquette(U, ext, y, m, v)).real / n
                     for y in range(U.shape[0]) for m in range(D) for v in range(m + 1, D)))


def active_sites(ext, mu, off):
    nu = partner(ext, mu)
    return [i for i in range(int(np.prod(ext))) if coords(ext, i)[nu] % 4 == off]


def mlp(w, f):
    w = np.asarray(w, np.float64)
    W1 = w[: HID * NF].reshape(HID, NF)
    b1 = w[HID * NF: HID * NF + HID]
    w2 = w[HID * NF + HID: HID * NF + 2 * HID]
    b2 = w[HID * NF + 2 * HID]
    return b2 + w2 @ np.tanh(W1 @ np.asarray(f) + b1)


def link_inputs(U, ext, x, mu):
    """(S, features) of the updated link (x, mu)."""
    nu = partner(ext, mu)
    n = U.shape[-1]
    S = U[shift(ext, x, mu), nu] @ U[shift(ext, x, nu), mu].conj().T @ U[x, nu].conj().T
    y = shift(ext, x, nu)
    f = [np.trace(plaquette(U, ext, shift(ext, y, mu, k), mu, nu)).real / n for k in (-1, 0, 1)]
    return S, f


def spectral_map(W, t):
    """(W', log J) of the spectral map, through the eigen-decomposition of W."""
    n = W.shape[0]
    lam, vec = np.linalg.eig(W)
    lam = lam / np.abs(lam)  # unitary: eigenvalues on the circle
    phi = float(np.sum(np.angle(1.0 - t * lam)))
    lp = np.exp(2j * phi / n) * (lam - t) / (1.0 - t * lam)
    Wp = vec @ np.diag(lp) @ np.linalg.inv(vec)
    mp = (1.0 - t * t) / np.abs(1.0 - t * lam) ** 2
    lj = float(np.sum(np.log(mp)) + np.log(np.mean(1.0 / mp)))
    iu = np.triu_indices(n, 1)
    lj += float(np.sum(np.log(np.abs(lp[iu[0]] - lp[iu[1]]) ** 2) - np.log(np.abs(lam[iu[0]] - lam[iu[1]]) ** 2)))
    return Wp, lj


def logj_closed_form(W, t):
    """The HIP kernels' form of log J (no eigenvalues)."""
    n = W.shape[0]
    _, ld = np.linalg.slogdet(np.eye(n) - t * W)
    m = 1.0 + t * t - 2.0 * t / n * np.trace(W).real
    return n * n * np.log(1.0 - t * t) - 2.0 * n * ld + np.log(m / (1.0 - t * t))


def sublayer(U, ext, mu, off, w, t_max):
    """Apply one sub-layer in place; returns (log J of the sub-layer, t per
    updated link)."""
    lj, ts = 0.0, {}
    upd = {}
    for x in active_sites(ext, mu, off):
        S, f = link_inputs(U, ext, x, mu)
        t = t_max * np.tanh(mlp(w, f))
        Wp, l = spectral_map(U[x, mu] @ S, t)
        upd[x] = Wp @ S.conj().T
        lj += l
        ts[x] = t
    for x, m in upd.items():
        U[x, mu] = m
    return lj, ts


def forward(V, ext, subs, weights, t_max):
    """U = F(V) and log J.  subs: list of (mu, off, block); weights: [nblocks,
    bsize] float32."""
    U = np.array(V, np.complex128, copy=True)
    lj = 0.0
    for mu, off, b in subs:
        l, _ = sublayer(U, ext, mu, off, weights[b], t_max)
        lj += l
    return U, lj


def phi_action(V, ext, subs, weights, t_max, beta):
    """Phi(V) = beta S_W(F(V)) - log J(V), the FTHMC action in V."""
    U, lj = forward(V, ext, subs, weights, t_max)
    return beta * wilson_action(U, ext) - lj


def su_generators(n):
    """Hermitian generators T^a, Tr T^a T^b = delta_ab / 2 (n^2 - 1 of them)."""
    T = []
    for i in range(n):
        for j in range(i + 1, n):
            m = np.zeros((n, n), complex)
            m[i, j] = m[j, i] = 0.5
            T.append(m)
            m = np.zeros((n, n), complex)
            m[i, j], m[j, i] = -0.5j, 0.5j
            T.append(m)
    for k in range(1, n):
        m = np.zeros((n, n), complex)
        m[np.arange(k), np.arange(k)] = 1.0
        m[k, k] = -k
        T.append(m / np.sqrt(2.0 * k * (k + 1)))
    return T


def expm_herm(X):
    """exp(i X) for Hermitian X."""
    w, v = np.linalg.eigh(X)
    return (v * np.exp(1j * w)) @ v.conj().T


def lie_gradient(f, U, x, d, eps=1e-6):
    """Central-difference Lie gradient g = sum_a D^a f T^a of f(U) at link
    (x, d), D^a f = d/ds f(e^{i s T^a} U_(x,d)) at s = 0."""
    n = U.shape[-1]
    g = np.zeros((n, n), complex)
    for T in su_generators(n):
        Up, Um = U.copy(), U.copy()
        Up[x, d] = expm_herm(eps * T) @ U[x, d]
        Um[x, d] = expm_herm(-eps * T) @ U[x, d]
        g += (f(Up) - f(Um)) / (2 * eps) * T
    return g


def numerical_logdet(W, t, eps=1e-6):
    """log |det| of the Jacobian of W -> W' (spectral_map) in left-invariant
    coordinates x_a = 2 Tr(T^a X), W = e^{iX} W0, by central differences."""
    n = W.shape[0]
    T = su_generators(n)
    Wp0, _ = spectral_map(W, t)

    def coords_of(Wp):
        # left-invariant coordinates of Wp Wp0^{-1} = e^{iY}: Y from the log
        w, v = np.linalg.eig(Wp @ Wp0.conj().T)
        Y = (v * np.angle(w)) @ np.linalg.inv(v)
        return np.array([2.0 * np.trace(Ta @ Y).real for Ta in T])

    J = np.zeros((len(T), len(T)))
    for b, Tb in enumerate(T):
        p, _ = spectral_map(expm_herm(eps * Tb) @ W, t)
        m, _ = spectral_map(expm_herm(-eps * Tb) @ W, t)
        J[:, b] = (coords_of(p) - coords_of(m)) / (2 * eps)
    return float(np.linalg.slogdet(J)[1])


def haar(n, rng):
    """Haar-random SU(n) (QR of a Ginibre matrix, phases fixed, det 1)."""
    z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    q = q * (np.diag(r) / np.abs(np.diag(r)))
    return q / np.linalg.det(q) ** (1.0 / n)


def random_field(ext, n, rng, spread=None):
    """[V, D, n, n] field: Haar links, or e^{i spread H} near the identity."""
    V, D = int(np.prod(ext)), len(ext)
    U = np.zeros((V, D, n, n), complex)
    for i in range(V):
        for d in range(D):
            if spread is None:
                U[i, d] = haar(n, rng)
            else:
                h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
                h = (h + h.conj().T) / 2
                h -= np.trace(h) / n * np.eye(n)
                U[i, d] = expm_herm(spread * h)
    return U
