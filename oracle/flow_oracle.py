"""CPU restatement (numpy, FP64) of the learned field transformation F(V).

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in
askja_amd/csrc/flow.hip.  The reference has no counterpart (the learned
proposal is an addition to its method); the definitions are these:

  for every link l = (x, mu) of the sub-layer's mask (direction mu, sites of
  one parity), with every other link frozen:
    S     = sum_{nu != mu} [U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
                            + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)]
    W     = U_l S
    P_j   = plaquettes at x in the planes (nu < la) without mu
    R_j   = the two rectangles at x in plane j: 2 x 1 (two steps along nu)
            and 1 x 2 (two steps along la), same orientation as P_j
    Lam   = TAH(sum_j P_j), Lhat = Lam / sqrt(2 Tr Lam^2 + delta^2)
    rho   = rho_max tanh(b2 + w2 . tanh(W1 f + b1)), features
            f = (ReTr P_j / 3 for every plane, then ReTr R1_j / 3, ReTr R2_j / 3
            plane by plane): nf = 3 x (number of planes)
    theta = rho Im Tr(Lhat W)
    U'_l  = expm(i theta Lhat) U_l,   log J += log(1 + rho ReTr(Lhat^2 W))

Weight block layout: W1[32][nf], b1[32], w2[32], b2.  Fields are the AoS
arrays [V, D, 3, 3] complex128 of the reference's U-*.conf order; geometry
from the Oracle (structs.h:169-218 restated).
"""
import numpy as np
import scipy.linalg

HID = 32


def tah(c):
    d = c - c.conj().T
    d = d - np.trace(d) / 3 * np.eye(3)
    return -0.5j * d


def parity(ext, i):
    return sum(np.unravel_index(i, ext)) % 2


def link_inputs(o, U, x, mu):
    """Frozen inputs of link (x, mu): staple sum S, plaquette sum C and the
    feature vector (plaquettes, then the two rectangles of every plane)."""
    D = o.dim
    xp = o.getplus(x, mu)
    S = np.zeros((3, 3), complex)
    C = np.zeros((3, 3), complex)
    f, fr = [], []
    for nu in range(D):
        if nu == mu:
            continue
        xpn, xmn = o.getplus(x, nu), o.getminus(x, nu)
        xpmn = o.getminus(xp, nu)
        S += U[xp, nu] @ U[xpn, mu].conj().T @ U[x, nu].conj().T
        S += U[xpmn, nu].conj().T @ U[xmn, mu].conj().T @ U[xmn, nu]
        for la in range(nu + 1, D):
            if la == mu:
                continue
            xla = o.getplus(x, la)
            P = U[x, nu] @ U[xpn, la] @ U[xla, nu].conj().T @ U[x, la].conj().T
            f.append(np.trace(P).real / 3)
            C += P
            x2n, xnl, x2l = o.getplus(xpn, nu), o.getplus(xpn, la), o.getplus(xla, la)
            R1 = (U[x, nu] @ U[xpn, nu] @ U[x2n, la] @ U[xnl, nu].conj().T @ U[xla, nu].conj().T
                  @ U[x, la].conj().T)
            R2 = (U[x, nu] @ U[xpn, la] @ U[xnl, la] @ U[x2l, nu].conj().T @ U[xla, la].conj().T
                  @ U[x, la].conj().T)
            fr += [np.trace(R1).real / 3, np.trace(R2).real / 3]
    return S, C, np.array(f + fr)


def sub_layer(o, U, mu, par, w, nf, rho_max, delta):
    """One masked sub-layer applied to U (copied); returns (U', log J)."""
    V = o.V
    out = U.copy()
    logj = 0.0
    W1 = w[: HID * nf].reshape(HID, nf).astype(np.float64)
    b1 = w[HID * nf: HID * nf + HID].astype(np.float64)
    w2 = w[HID * nf + HID: HID * nf + 2 * HID].astype(np.float64)
    b2 = float(w[HID * nf + 2 * HID])
    for x in range(V):
        if parity(o.ext, x) != par:
            continue
        S, C, f = link_inputs(o, U, x, mu)
        Wl = U[x, mu] @ S
        Lam = tah(C)
        n = np.sqrt(2 * np.trace(Lam @ Lam).real + delta * delta)
        Lh = Lam / n
        z = b2 + w2 @ np.tanh(W1 @ f + b1)
        rho = rho_max * np.tanh(z)
        theta = rho * np.trace(Lh @ Wl).imag
        K = np.trace(Lh @ Lh @ Wl).real
        out[x, mu] = scipy.linalg.expm(1j * theta * Lh) @ U[x, mu]
        logj += np.log1p(rho * K)
    return out, logj


def forward(o, V, subs, weights, bsize, nf, rho_max, delta):
    """subs: list of (mu, parity, block).  Returns (U = F(V), log J)."""
    U = V.copy()
    lj = 0.0
    for mu, par, blk in subs:
        U, l = sub_layer(o, U, mu, par, weights[blk * bsize:(blk + 1) * bsize], nf, rho_max, delta)
        lj += l
    return U, lj


def energy_mag(o, U):
    """EnergyMag (sun.cpp:507-537) of an AoS field."""
    s = 0.0
    for x in range(o.V):
        for j in range(o.dim):
            for k in range(j + 1, o.dim):
                P = U[x, j] @ U[o.getplus(x, j), k] @ U[o.getplus(x, k), j].conj().T @ U[x, k].conj().T
                s += 1 - np.trace(P).real / 3
    return s


def phi(o, V, subs, weights, bsize, nf, rho_max, delta, beta):
    """Phi = beta S_W(F(V)) - log J(V)."""
    U, lj = forward(o, V, subs, weights, bsize, nf, rho_max, delta)
    return beta * energy_mag(o, U) - lj


# -- batched restatement (whole lattice at once) --------------------------------
# The same map, evaluated for every masked link of a sub-layer at once with
# batched 3x3 products, so that lattices of the 16^4 size are checked in
# seconds; the exponential of the Hermitian i theta Lhat is taken from the
# eigendecomposition (agrees with expm to rounding).  Pinned against the
# per-link restatement above on small lattices.

def neighbours(ext):
    """(up[d][x], down[d][x], parity[x]) for lexicographic sites (x[D-1]
    fastest), periodic +-1 steps (structs.h:169-218)."""
    ext = list(ext)
    V = int(np.prod(ext))
    c = np.array(np.unravel_index(np.arange(V), ext))
    up, dn = [], []
    for d in range(len(ext)):
        cu = c.copy()
        cu[d] = (cu[d] + 1) % ext[d]
        up.append(np.ravel_multi_index(cu, ext))
        cd = c.copy()
        cd[d] = (cd[d] - 1) % ext[d]
        dn.append(np.ravel_multi_index(cd, ext))
    return np.array(up), np.array(dn), c.sum(axis=0) % 2


def _dag(a):
    return np.conj(np.swapaxes(a, -1, -2))


def _tr(a):
    return np.trace(a, axis1=-2, axis2=-1)


def _tah_b(c):
    d = c - _dag(c)
    d = d - (_tr(d) / 3)[..., None, None] * np.eye(3)
    return -0.5j * d


def link_inputs_batched(up, dn, U, xs, mu):
    """S, C and the features of the links (xs, mu), as link_inputs."""
    D = U.shape[1]
    xp = up[mu][xs]
    S = np.zeros((len(xs), 3, 3), complex)
    C = np.zeros_like(S)
    f, fr = [], []
    for nu in range(D):
        if nu == mu:
            continue
        xpn, xmn = up[nu][xs], dn[nu][xs]
        xpmn = dn[nu][xp]
        S += U[xp, nu] @ _dag(U[xpn, mu]) @ _dag(U[xs, nu])
        S += _dag(U[xpmn, nu]) @ _dag(U[xmn, mu]) @ U[xmn, nu]
        for la in range(nu + 1, D):
            if la == mu:
                continue
            xla = up[la][xs]
            P = U[xs, nu] @ U[xpn, la] @ _dag(U[xla, nu]) @ _dag(U[xs, la])
            f.append(_tr(P).real / 3)
            C += P
            x2n, xnl, x2l = up[nu][xpn], up[la][xpn], up[la][xla]
            R1 = U[xs, nu] @ U[xpn, nu] @ U[x2n, la] @ _dag(U[xnl, nu]) @ _dag(U[xla, nu]) @ _dag(U[xs, la])
            R2 = U[xs, nu] @ U[xpn, la] @ U[xnl, la] @ _dag(U[x2l, nu]) @ _dag(U[xla, la]) @ _dag(U[xs, la])
            fr += [_tr(R1).real / 3, _tr(R2).real / 3]
    return S, C, np.stack(f + fr, axis=1)


def _expi_herm(h, theta):
    """exp(i theta h) for Hermitian h, batched."""
    w, v = np.linalg.eigh(h)
    return (v * np.exp(1j * theta[:, None] * w)[:, None, :]) @ _dag(v)


def sub_layer_batched(up, dn, par_of, U, mu, par, w, nf, rho_max, delta):
    xs = np.nonzero(par_of == par)[0]
    W1 = w[: HID * nf].reshape(HID, nf).astype(np.float64)
    b1 = w[HID * nf: HID * nf + HID].astype(np.float64)
    w2 = w[HID * nf + HID: HID * nf + 2 * HID].astype(np.float64)
    b2 = float(w[HID * nf + 2 * HID])
    S, C, f = link_inputs_batched(up, dn, U, xs, mu)
    Wl = U[xs, mu] @ S
    Lam = _tah_b(C)
    n = np.sqrt(2 * _tr(Lam @ Lam).real + delta * delta)
    Lh = Lam / n[:, None, None]
    z = b2 + np.tanh(f @ W1.T + b1) @ w2
    rho = rho_max * np.tanh(z)
    theta = rho * _tr(Lh @ Wl).imag
    K = _tr(Lh @ Lh @ Wl).real
    out = U.copy()
    out[xs, mu] = _expi_herm(Lh, theta) @ U[xs, mu]
    return out, float(np.sum(np.log1p(rho * K)))


def forward_batched(ext, V, subs, weights, bsize, nf, rho_max, delta):
    """forward() for the whole lattice at once: (U = F(V), log J)."""
    up, dn, par_of = neighbours(ext)
    U = V.copy()
    lj = 0.0
    for mu, par, blk in subs:
        U, l = sub_layer_batched(up, dn, par_of, U, mu, par, weights[blk * bsize:(blk + 1) * bsize], nf, rho_max,
                                 delta)
        lj += l
    return U, lj


def energy_mag_batched(ext, U):
    """EnergyMag (sun.cpp:507-537), batched over sites."""
    up, _, _ = neighbours(ext)
    D = U.shape[1]
    s = 0.0
    for j in range(D):
        for k in range(j + 1, D):
            P = U[:, j] @ U[up[j], k] @ _dag(U[up[k], j]) @ _dag(U[:, k])
            s += float(np.sum(1 - _tr(P).real / 3))
    return s


# -- stout sub-layer ----------------------------------------------------------------
# For every masked link l = (x, mu), with S, the features f and rho as above
# (rho = rho_max tanh(MLP(f)), rho_max < 1/(2(D-1))):
#   W  = U_l S,  Q = -rho TAH(W)  (Hermitian traceless),  U'_l = exp(i Q) U_l.
# Jacobian in left-invariant coordinates x_a = 2 Tr(T^a X) (Tr T^a T^b = d/2):
#   J = Ad_{exp(iQ)} (1 + g(A) M),  g(z) = (1 - e^{-z}) / z,
#   A_ab = f_abc q_c (the matrix of ad_{iQ}; [T^a, T^b] = i f_abc T^c),
#   M_ab = -2 rho ReTr(T^a T^b W)   (dQ for dU = i T^b U),
# so log det J = log det(1 + g(A) M) (det Ad = 1).  Here g(A) is summed as a
# power series of the 8 x 8 matrix A and the determinant taken by slogdet:
# an independent route from the device's (eigenbasis of Q).

def su3_basis():
    """An orthonormal basis T^a of su(3): the Gell-Mann matrices / 2."""
    l = np.zeros((8, 3, 3), complex)
    l[0][0, 1] = l[0][1, 0] = 1
    l[1][0, 1], l[1][1, 0] = -1j, 1j
    l[2][0, 0], l[2][1, 1] = 1, -1
    l[3][0, 2] = l[3][2, 0] = 1
    l[4][0, 2], l[4][2, 0] = -1j, 1j
    l[5][1, 2] = l[5][2, 1] = 1
    l[6][1, 2], l[6][2, 1] = -1j, 1j
    l[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return l / 2


def structure_constants(T):
    """f_abc = -2i Tr([T^a, T^b] T^c), real and totally antisymmetric."""
    comm = np.einsum("aij,bjk->abik", T, T) - np.einsum("bij,ajk->abik", T, T)
    f = -2j * np.einsum("abij,cji->abc", comm, T)
    assert np.max(np.abs(f.imag)) < 1e-14
    return f.real


def coords(X, T):
    """x_a = 2 Re Tr(T^a X) of (batched) Hermitian traceless X."""
    return 2 * np.einsum("aij,nji->na", T, X).real


def stout_link_batched(W, rho, nterms=40):
    """(exp(iQ), log det J) per link for W = U S, Q = -rho TAH(W)."""
    T = su3_basis()
    f = structure_constants(T)
    Q = -rho[:, None, None] * _tah_b(W)
    q = coords(Q, T)
    A = np.einsum("abc,nc->nab", f, q)
    TT = np.einsum("aij,bjk->abik", T, T)
    M = -2 * rho[:, None, None] * np.einsum("abij,nji->nab", TT, W).real
    gA = np.zeros_like(A)
    P = np.broadcast_to(np.eye(8), A.shape).copy()
    fact = 1.0
    for k in range(nterms):
        fact /= (k + 1)
        gA += ((-1) ** k) * fact * P
        P = P @ A
    J = np.eye(8) + gA @ M
    sign, ld = np.linalg.slogdet(J)
    assert np.all(sign > 0)
    w, v = np.linalg.eigh(Q)
    E = (v * np.exp(1j * w)[:, None, :]) @ _dag(v)
    return E, ld


def stout_sub_layer_batched(up, dn, par_of, U, mu, par, w, nf, rho_max):
    xs = np.nonzero(par_of == par)[0]
    W1 = w[: HID * nf].reshape(HID, nf).astype(np.float64)
    b1 = w[HID * nf: HID * nf + HID].astype(np.float64)
    w2 = w[HID * nf + HID: HID * nf + 2 * HID].astype(np.float64)
    b2 = float(w[HID * nf + 2 * HID])
    S, C, feat = link_inputs_batched(up, dn, U, xs, mu)
    Wl = U[xs, mu] @ S
    rho = rho_max * np.tanh(b2 + np.tanh(feat @ W1.T + b1) @ w2)
    E, ld = stout_link_batched(Wl, rho)
    out = U.copy()
    out[xs, mu] = E @ U[xs, mu]
    return out, float(np.sum(ld))


def stout_forward_batched(ext, V, subs, weights, bsize, nf, rho_max):
    """The stout map for the whole lattice: (U = F(V), log J)."""
    up, dn, par_of = neighbours(ext)
    U = V.copy()
    lj = 0.0
    for mu, par, blk in subs:
        U, l = stout_sub_layer_batched(up, dn, par_of, U, mu, par, weights[blk * bsize:(blk + 1) * bsize], nf,
                                       rho_max)
        lj += l
    return U, lj


def stout_phi(ext, V, subs, weights, bsize, nf, rho_max, beta):
    """Phi = beta S_W(F(V)) - log J(V) of the stout map."""
    U, lj = stout_forward_batched(ext, V, subs, weights, bsize, nf, rho_max)
    return beta * energy_mag_batched(ext, U) - lj
