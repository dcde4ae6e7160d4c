import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def random_hermitian(rng, shape, n, scale=1.0):
    a = rng.standard_normal(shape + (n, n)) + 1j * rng.standard_normal(shape + (n, n))
    h = 0.5 * (a + np.conj(np.swapaxes(a, -1, -2)))
    return scale * h


def random_su(oracle_mod, rng, shape, n):
    """Haar-ish SU(n) matrices via the oracle's projection of Ginibre draws."""
    a = rng.standard_normal(shape + (n, n)) + 1j * rng.standard_normal(shape + (n, n))
    flat = a.reshape(-1, n, n)
    out = np.empty_like(flat)
    for k in range(flat.shape[0]):
        out[k] = oracle_mod.project_sun(flat[k])
    return out.reshape(shape + (n, n))
