"""Exact plaquette expectation of 2D SU(N) lattice gauge theory.

TEST INFRASTRUCTURE ONLY (the checker for the GPU chains at D = 2).

In two dimensions with the Wilson action beta * sum_p (1 - ReTr U_p / N), the
plaquettes of a large periodic lattice are independent (up to corrections
exponentially small in the volume), each distributed as the single-plaquette
integral
    Z(x) = int_SU(N) dU exp(x ReTr U),   x = beta / N,
and  <ReTr U_p / N> = (1/N) d ln Z / dx.  For SU(N) (Bars & Green 1980;
Gross & Witten 1980 for the large-N limit)
    Z(x) = sum_k det[ I_{k+i-j}(x) ]_{i,j = 1..N},
the k-th term being the U(N) integral weighted by det U^k.  The terms fall
off like exp(-c k^2); the sum is cut where they drop below 10^-(dps/2).
"""
import mpmath as mp


def _det_term(n, x, k, cache):
    M = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            m = k + i - j
            if m not in cache:
                cache[m] = mp.besseli(m, x)
            M[i, j] = cache[m]
    return mp.det(M)


def log_z(n, x, dps=60):
    with mp.workdps(dps):
        x = mp.mpf(x)
        cache = {}
        z = _det_term(n, x, 0, cache)
        k = 1
        while True:
            t = _det_term(n, x, k, cache) + _det_term(n, x, -k, cache)
            z += t
            if abs(t) < abs(z) * mp.mpf(10) ** (-dps // 2) or k > 64:
                break
            k += 1
        return mp.log(z)


def plaquette(n, beta, dps=60):
    """<ReTr U_p / N> at coupling beta, by a central difference of log Z in
    x = beta/N with step h = 10^-(dps/5): truncation ~ h^2, rounding
    ~ 10^-(dps - c)/h, where c ~ 2x/ln 10 digits are lost to the condition
    of the Toeplitz matrix (symbol exp(x cos t)); dps must exceed about
    c + 2 dps/5 + 16 for a double-precision result."""
    with mp.workdps(dps):
        x = mp.mpf(beta) / n
        h = mp.mpf(10) ** (-(dps // 5))
        lp, lm = log_z(n, x + h, dps), log_z(n, x - h, dps)
        if mp.im(lp) != 0 or mp.im(lm) != 0:
            raise ArithmeticError("working precision too low for the Bessel determinants")
        return float((lp - lm) / (2 * h) / n)
