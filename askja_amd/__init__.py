"""askja_amd -- MI355X-native SU(N) Hybrid Monte Carlo engine.

The product is the C ABI library ``libaskja.so`` (declared in
``include/askja.h``) built from the HIP kernels in ``askja_amd/csrc``; the
reference's own driver interface (the ``askja <params>`` program) is rebuilt
in C++ in ``askja_amd/driver``.  This module is a thin ctypes binding of the
C ABI for Python callers (tests, benchmark, training code).

``Lattice`` mirrors the reference's in-process functions (sun.cpp):

=====================================  =====================================
reference (sun.cpp)                    Lattice
=====================================  =====================================
initializeU            (194-207)       ``cold()``
initializeE(seed)      (209-252)       ``mom_gsl(seed, beta)``; ``mom_refresh``
updateE(eps)           (672-723)       ``updateE(eps)``
updateUandcopyDown(eps)(774-790)       ``updateUandcopyDown(eps, level)``
takeconfiguration      (725-732)       ``takeconfiguration()``
loadconfiguration      (734-742)       ``loadconfiguration()``
EnergyMag              (507-537)       ``EnergyMag()``
simpleenergyEl(Pi)     (540-548)       ``simpleenergyEl()``
EnergyEl               (560-567)       ``EnergyEl()``
GaussLaw               (415-438)       ``GaussLaw()``
UViol                  (441-471)       ``UViol()``
WilsonLoop             (307-345)       ``WilsonLoop()``
=====================================  =====================================

There is no CPU fallback: without the built library every constructor raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ASKJA_LIB") or os.path.join(_HERE, "libaskja.so")

FP64, FP64_EXACT, FP32, BF16X3 = 0, 1, 2, 3
_ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EUNSUP", -5: "ESTICKY", -6: "EIO"}

# every exported entry point: name -> (restype, argtypes)
_P, _D, _I, _L, _U64 = C.c_void_p, C.c_double, C.c_int, C.c_long, C.c_uint64
_DP = C.POINTER(C.c_double)


class HmcParams(C.Structure):
    _fields_ = [("beta", C.c_double), ("dt", C.c_double), ("ndrift", C.c_int), ("level", C.c_int),
                ("reunit", C.c_int), ("seed", C.c_uint64)]


class HmcResult(C.Structure):
    _fields_ = [("h_old", C.c_double), ("h_new", C.c_double), ("dh", C.c_double), ("uniform", C.c_double),
                ("action", C.c_double), ("accepted", C.c_int), ("accept_case", C.c_int)]


SIGNATURES = {
    "askja_create": (_I, [_I, _I, C.POINTER(C.c_int), _I, C.POINTER(_P)]),
    "askja_destroy": (_I, [_P]),
    "askja_last_error": (C.c_char_p, [_P]),
    "askja_clear_error": (_I, [_P]),
    "askja_info": (_I, [_P, C.POINTER(C.c_long), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "askja_field_doubles": (C.c_size_t, [_P]),
    "askja_links_upload": (_I, [_P, _DP]),
    "askja_links_download": (_I, [_P, _DP]),
    "askja_links_download_async": (_I, [_P, _DP]),
    "askja_download_wait": (_I, [_P]),
    "askja_host_alloc": (_I, [C.c_size_t, C.POINTER(_DP)]),
    "askja_host_free": (_I, [_DP]),
    "askja_links_cold": (_I, [_P]),
    "askja_links_haar": (_I, [_P, _U64, _U64]),
    "askja_reunitarize": (_I, [_P]),
    "askja_snapshot": (_I, [_P, _I]),
    "askja_mom_upload": (_I, [_P, _DP]),
    "askja_mom_download": (_I, [_P, _DP]),
    "askja_mom_download_prev": (_I, [_P, _DP]),
    "askja_mom_gsl": (_I, [_P, C.c_ulong, _D]),
    "askja_mom_refresh": (_I, [_P, _U64, _U64, _D]),
    "askja_mom_negate": (_I, [_P]),
    "askja_kick": (_I, [_P, _D]),
    "askja_drift": (_I, [_P, _D, _I]),
    "askja_kick_drift": (_I, [_P, _D, _D, _I]),
    "askja_md_trajectory": (_I, [_P, _D, _I, _I, _DP]),
    "askja_action": (_I, [_P, _DP, _DP]),
    "askja_kinetic": (_I, [_P, _DP]),
    "askja_energy_el": (_I, [_P, _DP]),
    "askja_gauss_law": (_I, [_P, _DP]),
    "askja_uviol": (_I, [_P, _DP]),
    "askja_wilson_lines": (_I, [_P, _I, _DP]),
    "askja_line_matrix": (_I, [_P, _I, C.POINTER(C.c_int), _DP]),
    "askja_slice_traces": (_I, [_P, _I, _DP]),
    "askja_line_phases": (_I, [_P, _I, _I, C.POINTER(C.c_int64), _DP]),
    "askja_hmc_trajectory": (_I, [_P, C.POINTER(HmcParams), _U64, C.POINTER(HmcResult)]),
    "askja_hmc_trajectory_noalloc": (_I, [_P, C.POINTER(HmcParams), _U64, C.POINTER(HmcResult)]),
    "askja_uniform": (_I, [_P, _U64, _U64, _DP]),
    "askja_tau_int": (_I, [_DP, _L, _D, _DP, C.POINTER(C.c_long)]),
    "askja_flow_define": (_I, [_P, _I, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _I, _D, _D]),
    "askja_flow_define_stout": (_I, [_P, _I, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _I, _D]),
    "askja_flow_define_spectral": (_I, [_P, _I, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _I, _D]),
    "askja_flow_block_size": (_I, [_P, C.POINTER(C.c_int)]),
    "askja_flow_set_weights": (_I, [_P, _P, _I]),
    "askja_flow_apply": (_I, [_P, _DP]),
    "askja_flow_download": (_I, [_P, _DP]),
    "askja_flow_gradient": (_I, [_P, _D, _I, _DP, _DP]),
    "askja_flow_gradient_download": (_I, [_P, _DP]),
    "askja_flow_weight_gradient": (_I, [_P, _DP]),
    "askja_flow_hmc_trajectory": (_I, [_P, C.POINTER(HmcParams), _U64, C.POINTER(HmcResult)]),
    "askja_flow_hmc_trajectory_nested": (_I, [_P, C.POINTER(HmcParams), _I, _D, _U64, C.POINTER(HmcResult)]),
    "askja_flow_wilson_lines": (_I, [_P, _I, _DP]),
    "askja_sync": (_I, [_P]),
    "askja_timer_start": (_I, [_P]),
    "askja_timer_stop": (_I, [_P, _DP]),
    "askja_device_fields": (_I, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
}

_LIB = None


class AskjaError(RuntimeError):
    pass


def lib():
    """Load libaskja.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise AskjaError(f"{LIB_PATH} is not built: run `make -C askja_amd/csrc` "
                             "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dptr(a):
    return a.ctypes.data_as(_DP)


def tau_int(x, S=1.5):
    """Integrated autocorrelation time tau = 1 + 2 sum rho(t) of a series and
    Wolff's automatic window W (askja_tau_int; host code, no GPU)."""
    x = np.ascontiguousarray(x, np.float64)
    tau, w = C.c_double(), C.c_long()
    rc = lib().askja_tau_int(_dptr(x), len(x), S, C.byref(tau), C.byref(w))
    if rc:
        raise AskjaError(f"tau_int: {_ERRORS.get(rc, rc)} (n = {len(x)}, S = {S})")
    return tau.value, w.value


class Lattice:
    """One device-resident SU(N) lattice (an ``askja_ctx``).

    ``extents`` = (N0, ..., N_{D-1}); ``precision`` one of FP64, FP64_EXACT,
    FP32.  Host fields are complex128 arrays [V, D, N, N] in the reference's
    U-*.conf order (lexicographic sites, x[D-1] fastest).
    """

    def __init__(self, nc, extents, precision=FP64):
        self.nc = int(nc)
        self.ext = tuple(int(e) for e in extents)
        self.dim = len(self.ext)
        self.V = int(np.prod(self.ext))
        self.precision = int(precision)
        h = _P()
        ext = (C.c_int * self.dim)(*self.ext)
        rc = lib().askja_create(self.nc, self.dim, ext, self.precision, C.byref(h))
        if rc != 0:
            raise AskjaError(f"askja_create(nc={nc}, extents={extents}, precision={precision}) failed: "
                             f"{_ERRORS.get(rc, rc)}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().askja_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            msg = lib().askja_last_error(self._h)
            raise AskjaError(f"{what}: {_ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    # -- fields ---------------------------------------------------------------
    def _empty(self):
        return np.zeros((self.V, self.dim, self.nc, self.nc), np.complex128)

    def links(self):
        out = self._empty()
        self._ck(lib().askja_links_download(self._h, _dptr(out.view(np.float64))), "links_download")
        return out

    def set_links(self, u):
        u = np.ascontiguousarray(np.asarray(u, np.complex128).reshape(self.V, self.dim, self.nc, self.nc))
        self._ck(lib().askja_links_upload(self._h, _dptr(u.view(np.float64))), "links_upload")

    def momenta(self, prev=False):
        out = self._empty()
        f = lib().askja_mom_download_prev if prev else lib().askja_mom_download
        self._ck(f(self._h, _dptr(out.view(np.float64))), "mom_download")
        return out

    def set_momenta(self, p):
        p = np.ascontiguousarray(np.asarray(p, np.complex128).reshape(self.V, self.dim, self.nc, self.nc))
        self._ck(lib().askja_mom_upload(self._h, _dptr(p.view(np.float64))), "mom_upload")

    # -- set-up -----------------------------------------------------------------
    def cold(self):
        self._ck(lib().askja_links_cold(self._h), "links_cold")

    def haar(self, seed, stream=0):
        self._ck(lib().askja_links_haar(self._h, seed, stream), "links_haar")

    def reunitarize(self):
        self._ck(lib().askja_reunitarize(self._h), "reunitarize")

    def takeconfiguration(self):
        self._ck(lib().askja_snapshot(self._h, 0), "snapshot save")

    def loadconfiguration(self):
        self._ck(lib().askja_snapshot(self._h, 1), "snapshot restore")

    def mom_gsl(self, seed, beta):
        self._ck(lib().askja_mom_gsl(self._h, seed, beta), "mom_gsl")

    def mom_refresh(self, seed, counter, beta):
        self._ck(lib().askja_mom_refresh(self._h, seed, counter, beta), "mom_refresh")

    def mom_negate(self):
        self._ck(lib().askja_mom_negate(self._h), "mom_negate")

    # -- molecular dynamics --------------------------------------------------------
    def updateE(self, eps):
        self._ck(lib().askja_kick(self._h, eps), "kick")

    def updateUandcopyDown(self, eps, level):
        self._ck(lib().askja_drift(self._h, eps, level), "drift")

    def kick_drift(self, eps_kick, eps_drift, level):
        self._ck(lib().askja_kick_drift(self._h, eps_kick, eps_drift, level), "kick_drift")

    def md_trajectory(self, dt, ndrift, level):
        e = np.zeros(4)
        self._ck(lib().askja_md_trajectory(self._h, dt, ndrift, level, _dptr(e)), "md_trajectory")
        return e

    # -- reductions ---------------------------------------------------------------
    def plaq_split(self):
        t, s = C.c_double(), C.c_double()
        self._ck(lib().askja_action(self._h, C.byref(t), C.byref(s)), "action")
        return t.value, s.value

    def EnergyMag(self):
        t, s = self.plaq_split()
        return t + s

    def _scalar(self, f, what):
        v = C.c_double()
        self._ck(f(self._h, C.byref(v)), what)
        return v.value

    def simpleenergyEl(self):
        return self._scalar(lib().askja_kinetic, "kinetic")

    def EnergyEl(self):
        return self._scalar(lib().askja_energy_el, "energy_el")

    def GaussLaw(self):
        return self._scalar(lib().askja_gauss_law, "gauss_law")

    def UViol(self):
        return self._scalar(lib().askja_uviol, "uviol")

    def wilson_lines(self, d):
        out = np.zeros(2 * (self.V // self.ext[d]))
        self._ck(lib().askja_wilson_lines(self._h, d, _dptr(out)), "wilson_lines")
        return out.view(np.complex128)

    def line_matrix(self, d, x):
        out = np.zeros(2 * self.nc * self.nc)
        xx = (C.c_int * self.dim)(*[int(v) for v in x])
        self._ck(lib().askja_line_matrix(self._h, d, xx, _dptr(out)), "line_matrix")
        return out.view(np.complex128).reshape(self.nc, self.nc)

    def slice_traces(self, d):
        """scalarLoopcorr (measurements4d.cpp:357-408) for links of direction d:
        mean over each time slice x0 of ReTr U_d / N."""
        out = np.zeros(self.ext[0])
        self._ck(lib().askja_slice_traces(self._h, d, _dptr(out)), "slice_traces")
        return out

    def line_phases(self, d, starts):
        """Sorted, winding-recentred eigen-phases (getevs,
        measurements4d.cpp:185-309) of the lines in direction d through the
        lexicographic sites `starts`; shape (len(starts), N)."""
        st = np.ascontiguousarray(starts, dtype=np.int64)
        out = np.zeros(len(st) * self.nc)
        self._ck(lib().askja_line_phases(self._h, d, len(st), st.ctypes.data_as(C.POINTER(C.c_int64)), _dptr(out)),
                 "line_phases")
        return out.reshape(len(st), self.nc)

    def WilsonLoop(self):
        """|Tr prod_t U_0(t, 0)|/N at the spatial origin (sun.cpp:307-345)."""
        return abs(self.wilson_lines(0)[0])

    def polyakov(self):
        """Volume average of Tr P/N over all spatial sites (complex)."""
        return complex(self.wilson_lines(0).mean())

    def polyakov_measure(self):
        """measure4d's PolyakovLoop (measurements4d.cpp:410-466): mean over x1
        (x2.. = 0) of |Tr prod_t U_0|/N."""
        lines = self.wilson_lines(0)
        stride = self.V // (self.ext[0] * self.ext[1])
        return float(np.mean(np.abs(lines[:: stride][: self.ext[1]])))

    def wilson_measure(self, d):
        """measure4d's WilsonLoop(evs, d) (measurements4d.cpp:311-355): mean over
        x0 of |Tr prod U_d|/N along the line through (x0, 0, ..., 0); the trace
        of a closed line does not depend on where it starts, so the
        reference's stale loop variable (the start x_d) does not matter."""
        lines = self.wilson_lines(d)
        stride = self.V // (self.ext[d] * self.ext[0])
        return float(np.mean(np.abs(lines[:: stride][: self.ext[0]])))

    # -- HMC --------------------------------------------------------------------------
    def uniform(self, seed, counter):
        """The Metropolis uniform of trajectory `counter` (askja_uniform)."""
        v = C.c_double()
        self._ck(lib().askja_uniform(self._h, seed, counter, C.byref(v)), "uniform")
        return v.value

    def hmc_trajectory(self, beta, dt, ndrift, level, seed, counter, reunit=True):
        p = HmcParams(beta, dt, ndrift, level, 1 if reunit else 0, seed)
        r = HmcResult()
        self._ck(lib().askja_hmc_trajectory(self._h, C.byref(p), counter, C.byref(r)), "hmc_trajectory")
        return r

    # -- learned field transformation ------------------------------------------
    def flow_define(self, mu, parity, block, nblocks, rho_max=0.3, delta=1e-3, kind="rank-one"):
        """kind "rank-one" (askja_flow_define: rotation along the frozen
        plaquettes' direction), "stout" (askja_flow_define_stout: smearing
        step along the staple force, rho_max < 1/(2(D-1)), delta unused) or
        "spectral" (askja_flow_define_spectral, D = 2, any N: `parity` holds
        the row offsets 0..3, rho_max is t_max <= 1/2, delta unused)."""
        n = len(mu)
        arr = lambda v: (C.c_int * n)(*[int(x) for x in v])
        if kind == "spectral":
            self._ck(lib().askja_flow_define_spectral(self._h, n, arr(mu), arr(parity), arr(block), int(nblocks),
                                                      rho_max), "flow_define_spectral")
        elif kind == "stout":
            self._ck(lib().askja_flow_define_stout(self._h, n, arr(mu), arr(parity), arr(block), int(nblocks), rho_max),
                     "flow_define_stout")
        elif kind == "rank-one":
            self._ck(lib().askja_flow_define(self._h, n, arr(mu), arr(parity), arr(block), int(nblocks), rho_max,
                                             delta), "flow_define")
        else:
            raise ValueError(f"unknown flow kind {kind!r}")
        self.flow_kind = kind
        b = C.c_int()
        self._ck(lib().askja_flow_block_size(self._h, C.byref(b)), "flow_block_size")
        self.flow_bsize, self.flow_nblocks = b.value, int(nblocks)
        return self.flow_bsize

    def flow_set_weights(self, w):
        """w: numpy float32 array (host) or an object with data_ptr() (device)."""
        if hasattr(w, "data_ptr"):
            self._ck(lib().askja_flow_set_weights(self._h, C.c_void_p(w.data_ptr()), 1), "flow_set_weights")
        else:
            w = np.ascontiguousarray(np.asarray(w, np.float32))
            assert w.size == self.flow_bsize * self.flow_nblocks
            self._ck(lib().askja_flow_set_weights(self._h, w.ctypes.data_as(C.c_void_p), 0), "flow_set_weights")

    def flow_apply(self):
        lj = C.c_double()
        self._ck(lib().askja_flow_apply(self._h, C.byref(lj)), "flow_apply")
        return lj.value

    def flow_output(self):
        out = self._empty()
        self._ck(lib().askja_flow_download(self._h, _dptr(out.view(np.float64))), "flow_download")
        return out

    def flow_gradient(self, beta, wgrad=False):
        s, lj = C.c_double(), C.c_double()
        self._ck(lib().askja_flow_gradient(self._h, beta, 1 if wgrad else 0, C.byref(s), C.byref(lj)),
                 "flow_gradient")
        return s.value, lj.value

    def flow_link_gradient(self):
        out = self._empty()
        self._ck(lib().askja_flow_gradient_download(self._h, _dptr(out.view(np.float64))), "flow_gradient_download")
        return out

    def flow_weight_gradient(self):
        out = np.zeros(self.flow_bsize * self.flow_nblocks)
        self._ck(lib().askja_flow_weight_gradient(self._h, _dptr(out)), "flow_weight_gradient")
        return out

    def flow_hmc_trajectory(self, beta, dt, ndrift, level, seed, counter, reunit=True, inner=None, alpha=1.0):
        """One FTHMC trajectory of ndrift steps of dt.  inner: the two-time-scale
        form (askja_flow_hmc_trajectory_nested), the flow's correction to the
        force of alpha beta S_W kicking every `inner` fine steps."""
        p = HmcParams(beta, dt, ndrift, level, 1 if reunit else 0, seed)
        r = HmcResult()
        if inner is None:
            self._ck(lib().askja_flow_hmc_trajectory(self._h, C.byref(p), counter, C.byref(r)), "flow_hmc_trajectory")
        else:
            self._ck(lib().askja_flow_hmc_trajectory_nested(self._h, C.byref(p), int(inner), float(alpha), counter,
                                                            C.byref(r)), "flow_hmc_trajectory_nested")
        return r

    def flow_polyakov(self):
        out = np.zeros(2 * (self.V // self.ext[0]))
        self._ck(lib().askja_flow_wilson_lines(self._h, 0, _dptr(out)), "flow_wilson_lines")
        return complex(out.view(np.complex128).mean())

    def sync(self):
        self._ck(lib().askja_sync(self._h), "sync")

    def timer_start(self):
        self._ck(lib().askja_timer_start(self._h), "timer_start")

    def timer_stop(self):
        return self._scalar(lib().askja_timer_stop, "timer_stop")

    def n_plaquettes(self):
        return self.V * self.dim * (self.dim - 1) // 2
