"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py`` as the checker.  The product path
(askja_amd, libaskja.so) never imports it.

The oracle restates the reference HMC (sun.cpp, sunmat.h, structs.h,
measurements4d.cpp) in plain C; see askja_oracle.c for the function-by-function
citations.  Fields are numpy views in the reference's AoS layout:
``[site (lexicographic, x[D-1] fastest)][dir][row][col]`` complex128.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        P, D, I, L_, U64, U32 = C.c_void_p, C.c_double, C.c_int, C.c_long, C.c_uint64, C.c_uint32
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        sig = {
            "oracle_create": (P, [I, I, ip, D, I]),
            "oracle_destroy": (None, [P]),
            "oracle_volume": (L_, [P]),
            "oracle_set_beta": (None, [P, D]),
            "oracle_set_level": (None, [P, I]),
            "oracle_field": (C.POINTER(C.c_double), [P, I]),
            "oracle_getplus": (L_, [P, L_, I]),
            "oracle_getminus": (L_, [P, L_, I]),
            "oracle_getind": (L_, [P, ip]),
            "oracle_getcoord": (None, [P, L_, ip]),
            "oracle_generators": (None, [I, dp]),
            "oracle_mt_stream": (None, [C.c_ulong, I, up]),
            "oracle_gauss_stream": (None, [C.c_ulong, I, dp]),
            "oracle_mt_first_uniform": (D, [C.c_ulong]),
            "oracle_initializeE": (None, [P, C.c_ulong]),
            "oracle_initializeU": (None, [P]),
            "oracle_updateE": (None, [P, D]),
            "oracle_updateUandcopyDown": (None, [P, D]),
            "oracle_copyDown": (None, [P]),
            "oracle_takeconfiguration": (None, [P]),
            "oracle_loadconfiguration": (None, [P]),
            "oracle_matmul": (None, [I, dp, dp, dp]),
            "oracle_tah": (None, [I, dp, dp]),
            "oracle_mexpi": (None, [I, dp, I, dp]),
            "oracle_energy_mag": (D, [P]),
            "oracle_plaq_exp": (None, [P, C.POINTER(D), C.POINTER(D)]),
            "oracle_simple_energy_el": (D, [P, I]),
            "oracle_energy_el": (D, [P]),
            "oracle_total_energy": (D, [P]),
            "oracle_gauss_law": (D, [P]),
            "oracle_uviol": (D, [P]),
            "oracle_wilson_origin": (D, [P]),
            "oracle_polyakov_measure": (D, [P]),
            "oracle_wilson_measure": (D, [P, I]),
            "oracle_polyakov_field": (None, [P, dp]),
            "oracle_philox4x32_10": (None, [up, up, up]),
            "oracle_philox_normals": (None, [U64, U64, U32, U32, I, dp]),
            "oracle_repro_log": (D, [D]),
            "oracle_repro_sincos2pi": (None, [D, C.POINTER(D), C.POINTER(D)]),
            "oracle_mom_philox": (None, [P, U64, U64]),
            "oracle_project_sun": (None, [I, dp]),
            "oracle_det": (None, [I, dp, dp]),
            "oracle_reunitarize": (None, [P]),
            "oracle_haar": (None, [P, U64, U64]),
            "oracle_hamiltonian": (D, [P, I]),
            "oracle_md_trajectory": (None, [P, D, I]),
            "oracle_hmc": (None, [P, D, I, I, U64, U64, I, dp]),
            "oracle_num_threads": (I, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


FIELDS = {"u": 0, "U": 1, "pi": 2, "Pi": 3, "uconf": 4}


class Oracle:
    """One lattice of the CPU oracle (the reference's globals u, U, pi, Pi, uconf)."""

    def __init__(self, nc, extents, beta=1.0, level=8):
        self.nc = int(nc)
        self.ext = [int(x) for x in extents]
        self.dim = len(self.ext)
        self._h = lib().oracle_create(self.nc, self.dim, np.asarray(self.ext, np.int32), float(beta), int(level))
        if not self._h:
            raise RuntimeError("oracle_create failed")
        self.V = int(lib().oracle_volume(self._h))
        self.beta = float(beta)
        self.level = int(level)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    # -- fields -----------------------------------------------------------
    def field(self, name):
        """numpy complex128 view [V, D, N, N] of one of u, U, pi, Pi, uconf."""
        ptr = lib().oracle_field(self._h, FIELDS[name])
        n = self.V * self.dim * self.nc * self.nc * 2
        arr = np.ctypeslib.as_array(ptr, shape=(n,))
        return arr.view(np.complex128).reshape(self.V, self.dim, self.nc, self.nc)

    def set_beta(self, beta):
        self.beta = float(beta)
        lib().oracle_set_beta(self._h, self.beta)

    def set_level(self, level):
        self.level = int(level)
        lib().oracle_set_level(self._h, self.level)

    # -- geometry ---------------------------------------------------------
    def getplus(self, i, d):
        return lib().oracle_getplus(self._h, i, d)

    def getminus(self, i, d):
        return lib().oracle_getminus(self._h, i, d)

    def getind(self, x):
        return lib().oracle_getind(self._h, np.asarray(x, np.int32))

    def getcoord(self, i):
        x = np.zeros(self.dim, np.int32)
        lib().oracle_getcoord(self._h, i, x)
        return x

    # -- reference operations ---------------------------------------------
    def initializeU(self):
        lib().oracle_initializeU(self._h)

    def initializeE(self, seed):
        lib().oracle_initializeE(self._h, seed)

    def updateE(self, eps):
        lib().oracle_updateE(self._h, eps)

    def updateUandcopyDown(self, eps):
        lib().oracle_updateUandcopyDown(self._h, eps)

    def copyDown(self):
        lib().oracle_copyDown(self._h)

    def take(self):
        lib().oracle_takeconfiguration(self._h)

    def load(self):
        lib().oracle_loadconfiguration(self._h)

    def energy_mag(self):
        return lib().oracle_energy_mag(self._h)

    def plaq_exp(self):
        t, s = C.c_double(), C.c_double()
        lib().oracle_plaq_exp(self._h, C.byref(t), C.byref(s))
        return t.value, s.value

    def simple_energy_el(self, which="pi"):
        return lib().oracle_simple_energy_el(self._h, FIELDS[which])

    def energy_el(self):
        return lib().oracle_energy_el(self._h)

    def total_energy(self):
        return lib().oracle_total_energy(self._h)

    def gauss_law(self):
        return lib().oracle_gauss_law(self._h)

    def uviol(self):
        return lib().oracle_uviol(self._h)

    def wilson_origin(self):
        return lib().oracle_wilson_origin(self._h)

    def polyakov_measure(self):
        return lib().oracle_polyakov_measure(self._h)

    def wilson_measure(self, d):
        return lib().oracle_wilson_measure(self._h, d)

    def polyakov_field(self):
        out = np.zeros(2 * (self.V // self.ext[0]), np.float64)
        lib().oracle_polyakov_field(self._h, out)
        return out.view(np.complex128)

    def mom_philox(self, seed, counter):
        lib().oracle_mom_philox(self._h, seed, counter)

    def haar(self, seed, stream=0):
        lib().oracle_haar(self._h, seed, stream)

    def reunitarize(self):
        lib().oracle_reunitarize(self._h)

    def hamiltonian(self, which="pi"):
        return lib().oracle_hamiltonian(self._h, FIELDS[which])

    def md_trajectory(self, dt, ndrift):
        lib().oracle_md_trajectory(self._h, dt, ndrift)

    def hmc(self, dt, ndrift, ntraj, seed, first_counter=0, reunit=True):
        rec = np.zeros(10 * ntraj, np.float64)
        lib().oracle_hmc(self._h, dt, ndrift, ntraj, seed, first_counter, 1 if reunit else 0, rec)
        return rec.reshape(ntraj, 10)


# -- stand-alone helpers -------------------------------------------------------

def generators(nc):
    g = np.zeros((nc * nc - 1) * nc * nc * 2, np.float64)
    lib().oracle_generators(nc, g)
    return g.view(np.complex128).reshape(nc * nc - 1, nc, nc)


def _cm(a):
    return np.ascontiguousarray(np.asarray(a, np.complex128)).view(np.float64).ravel()


def matmul(a, b):
    n = a.shape[0]
    out = np.zeros(2 * n * n)
    lib().oracle_matmul(n, _cm(a), _cm(b), out)
    return out.view(np.complex128).reshape(n, n)


def tah(w):
    n = w.shape[0]
    out = np.zeros(2 * n * n)
    lib().oracle_tah(n, _cm(w), out)
    return out.view(np.complex128).reshape(n, n)


def mexpi(a, level):
    n = a.shape[0]
    out = np.zeros(2 * n * n)
    lib().oracle_mexpi(n, _cm(a), level, out)
    return out.view(np.complex128).reshape(n, n)


def project_sun(a):
    n = a.shape[0]
    buf = _cm(a).copy()
    lib().oracle_project_sun(n, buf)
    return buf.view(np.complex128).reshape(n, n)


def det(a):
    n = a.shape[0]
    out = np.zeros(2)
    lib().oracle_det(n, _cm(a), out)
    return complex(out[0], out[1])


def mt_stream(seed, n):
    out = np.zeros(n, np.uint32)
    lib().oracle_mt_stream(seed, n, out)
    return out


def gauss_stream(seed, n):
    out = np.zeros(n, np.float64)
    lib().oracle_gauss_stream(seed, n, out)
    return out


def mt_first_uniform(seed):
    return lib().oracle_mt_first_uniform(seed)


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def philox_normals(seed, counter, purpose, link, n):
    out = np.zeros(n, np.float64)
    lib().oracle_philox_normals(seed, counter, purpose, link, n, out)
    return out


def num_threads():
    return lib().oracle_num_threads()


def repro_log(x):
    return lib().oracle_repro_log(float(x))


def repro_sincos2pi(u):
    s, c = C.c_double(), C.c_double()
    lib().oracle_repro_sincos2pi(float(u), C.byref(s), C.byref(c))
    return s.value, c.value
