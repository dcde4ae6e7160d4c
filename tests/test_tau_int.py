"""askja_tau_int (host code in libaskja.so, no GPU): the integrated
autocorrelation time with Wolff's automatic window, pinned by the closed form
of a first-order autoregressive series, x_t = phi x_{t-1} + e_t, whose
normalised autocorrelation is rho(t) = phi^t and so tau = 1 + 2 sum rho =
(1 + phi) / (1 - phi); by white noise (tau = 1); and against an FFT-based
estimate of the same autocorrelation."""
import math

import numpy as np
import pytest

import askja_amd as A


def ar1(phi, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0] / math.sqrt(1 - phi * phi)
    for i in range(1, n):
        x[i] = phi * x[i - 1] + e[i]
    return x


@pytest.mark.parametrize("phi", [0.3, 0.6, 0.85])
def test_ar1_closed_form(phi):
    n = 200_000
    tau, w = A.tau_int(ar1(phi, n, 7))
    exact = (1 + phi) / (1 - phi)
    # statistical error of tau: tau sqrt(2 (2W + 1) / n) (Wolff 2004, eq. 42)
    err = exact * math.sqrt(2 * (2 * w + 1) / n)
    assert abs(tau - exact) < 4 * err + 0.02 * exact, (tau, exact, w)
    assert w > 0


def test_white_noise():
    tau, w = A.tau_int(np.random.default_rng(3).standard_normal(100_000))
    assert abs(tau - 1.0) < 0.05, tau


def test_constant_and_short_series():
    assert A.tau_int(np.full(50, 2.5)) == (1.0, 0)
    with pytest.raises(A.AskjaError):
        A.tau_int(np.arange(7.0))
    with pytest.raises(A.AskjaError):
        A.tau_int(np.arange(20.0), S=0.0)


def test_matches_fft_autocorrelation():
    """The same estimator with the autocorrelation from an FFT (a different
    summation) picks the same window and agrees to rounding."""
    x = ar1(0.7, 5000, 11)
    n = len(x)
    d = x - x.mean()
    var = d @ d / n
    f = np.fft.rfft(d, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / (n - np.arange(n)) / var
    tint, want_w = 0.5, None
    for W in range(1, n // 2):
        tint += acf[W]
        tw = 1.5 / math.log((2 * tint + 1) / (2 * tint - 1)) if tint > 0.5 else 1e-9
        if math.exp(-W / tw) - tw / math.sqrt(W * n) < 0:
            want_w = W
            break
    tau, w = A.tau_int(x)
    assert w == want_w
    assert abs(tau - 2 * tint) < 1e-10 * tau
