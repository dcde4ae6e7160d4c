"""Benchmark of the HMC hot path on one MI355X.

Default workload (the configuration the headline metric is quoted on):
learned-proposal HMC against plain FP64 HMC for SU(3) in D = 4 on a 16^4
lattice, beta = 6.0, DT = 0.02, 50 MD steps per trajectory (length 1),
LEVEL 8.  The learned proposal is field-transformation HMC with the stout
map U = F(V) of one sweep layer (8 masked sub-layers: every direction and
parity, MLP of hidden width 32 over plaquette and rectangle traces), weights
trained on 500 plain-HMC configurations at beta 5.7
(askja_amd/data/flow_c4_weights_1sweep.json, written by ``python -m
askja_amd.learned --layers 1 --save-weights``), exact Metropolis with log J.
Plain HMC decorrelates |P| in about 1.5 trajectories here, which caps what
any proposal can gain per trajectory, so the map's cost decides the value:
one sweep costs a quarter of the four-sweep map
(askja_amd/data/flow_c4_weights.json, ``--weights`` selects it).  Data are synthetic: a Haar-random start from a fixed seed,
thermalised by plain HMC (untimed) before either chain.

One step = one learned-proposal trajectory: momentum refresh, 50 MD steps
of 0.02 in two time scales (Sexton and Weingarten's nested leapfrog): the
plain Wilson force on every step (the fused kick+drift), the flow's
correction Phi - beta S_W every 5 steps (the forward and hand-written
backward of every sub-layer, 11 flow gradients per trajectory instead of
51; ``--flow-inner 1`` runs the one-scale trajectory), H with log J,
Metropolis, the Polyakov loop of U.
--warmup trajectories of each chain run untimed, then exactly --steps
learned trajectories are timed; the plain chain is timed over its own
series.

metric value = (independent configurations/s of the learned chain) /
(independent configurations/s of the plain chain), each = trajectories/s /
tau with tau the integrated autocorrelation time of |volume-averaged
Polyakov loop| (Wolff's automatic window, S = 1.5; clamped to >= 1).  Each
tau comes from a series of at least 100 tau trajectories (extended beyond
--steps as needed, up to a cap; without it the value is null with the
reason).

``--workload c2`` runs the earlier plain-HMC line instead: SU(3) 8x32^3,
beta 5.9, DT 0.02, LEVEL 6, 100 MD steps per trajectory.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Trajectory length (the workload fixes DT, not the number of MD steps per
# trajectory), measured on this chain with long runs, independent
# configurations/s = traj/s / tau_int(|P|): 25 steps 0.37 (tau about 400),
# 50 steps 1.52 (tau 50.1 +- 8.5), 100 steps 1.89 (tau 20.4 +- 2.9), 150 steps
# 1.99 (tau 13.0 +- 1.8).  100 steps: near the best, and tau settles within
# a few minutes of GPU time.
NC, EXT, BETA, DT, LEVEL, NDRIFT = 3, [8, 32, 32, 32], 5.9, 0.02, 6, 100
SEED = 120351616
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # MI355X dense


# HBM bytes per fused kick+drift launch of this workload, measured with two
# rocprofv3 --pmc passes (FETCH_SIZE scaled by 2.0, WRITE_SIZE by 1.0, from a
# 1 GiB copy's readings): reads 772.7 MB + writes 226.8 MB (FP64),
# 264.3 MB + 113.5 MB (FP32)
PMC_TRAFFIC_BYTES = {"fp64": 999.5e6, "fp32": 377.8e6}
# the same for the 16^4 lattice's FP64 kick+drift (the plain chain of the
# learned-vs-plain workload): 154.6 MB reads + 56.7 MB writes
PMC_TRAFFIC_BYTES_C4 = 211.3e6


def _finite(o):
    """The JSON line with every non-finite number written as null (JSON has
    no NaN)."""
    if isinstance(o, dict):
        return {k: _finite(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_finite(v) for v in o]
    if isinstance(o, (float, np.floating)):
        return float(o) if math.isfinite(o) else None
    if isinstance(o, np.integer):
        return int(o)
    if isinstance(o, np.bool_):
        return bool(o)
    return o


def tau_int_wolff(x, S=1.5):
    """Integrated autocorrelation time tau = 1 + 2 sum rho(t) (1 for
    independent samples) with Wolff's automatic window (Comput. Phys.
    Commun. 156 (2004) 143), from the library's host code askja_tau_int.
    Returns (tau, window); NaN below 8 samples."""
    import askja_amd as A

    if len(x) < 8:
        return float("nan"), 0
    return A.tau_int(x, S)


def cpu_baseline(steps_sample, ndrift=NDRIFT):
    """The oracle (FP64 C restatement of the reference, OpenMP over sites in
    the kick and drift like sun.cpp:679,777) on the host cores, same lattice:
    `steps_sample` MD steps timed and extrapolated to one trajectory
    (ndrift steps + the two Hamiltonian evaluations)."""
    from oracle import oracle as O

    o = O.Oracle(NC, EXT, beta=BETA, level=LEVEL)
    o.haar(SEED, 2)
    o.mom_philox(SEED, 0)
    t0 = time.perf_counter()
    h = o.hamiltonian("pi")
    t_h = time.perf_counter() - t0
    o.updateE(DT / 2)
    t0 = time.perf_counter()
    for _ in range(steps_sample):
        o.updateUandcopyDown(DT)
        o.updateE(DT)
    t_step = (time.perf_counter() - t0) / steps_sample
    t_traj = ndrift * t_step + 2 * t_h
    assert math.isfinite(h)
    return {"value": 1.0 / t_traj, "unit": "traj/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{steps_sample} FP64 MD steps (kick+drift) of the oracle on 8x32^3 SU(3), "
                      f"{t_step * 1e3:.0f} ms/step, extrapolated to {ndrift} steps + 2 H evaluations"}


def copy_probe(gib=1, reps=5):
    """A plain device-to-device copy of a known size (the runtime's copy
    kernel), for calibrating the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = gib << 30
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 0, C.c_size_t(n)) == 0
    for _ in range(reps):
        t0 = time.perf_counter()
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        dt = time.perf_counter() - t0
        print(json.dumps({"copy_bytes": n, "ms": dt * 1e3, "GBps_read_plus_write": 2 * n / dt / 1e9}))
    hip.hipFree(a)
    hip.hipFree(b)


STUDY_CONFIGS = {
    # name: (N, extents, beta, dt, LEVEL, MD steps per trajectory (length 1),
    #        precisions, thermalising trajectories)
    "C2": (3, [8, 32, 32, 32], 5.9, 0.02, 6, 50, ("fp64", "fp32"), 100),
    "C3": (16, [6, 24, 24, 24], 4 * 16 * 16 / 10.0, 0.01, 8, 100, ("fp64", "fp32", "bf16x3"), 30),
    "C5": (64, [128, 128], 64.0 * 64.0, 0.005, 8, 200, ("fp64", "fp32", "bf16x3"), 20),
}


def precision_study(names):
    """Delta-H and plaquette error of FP32 and BF16x3 against FP64 on one
    thermalised configuration (Haar start, thermalised by exact HMC with the
    fastest MD precision) with the same momenta: one trajectory of length 1
    in each precision."""
    import askja_amd as A

    pm = {"fp64": A.FP64, "fp32": A.FP32, "bf16x3": A.BF16X3}
    for name in names:
        n, ext, beta, dt, level, nd, precs, therm = STUDY_CONFIGS[name]
        ref = A.Lattice(n, ext, A.BF16X3 if n > 8 else A.FP32)
        ref.haar(SEED, 5)
        for t in range(therm):
            ref.hmc_trajectory(beta, dt, nd, level, SEED, t, reunit=True)
        print(f"[study] {name} thermalised ({therm} trajectories)", file=sys.stderr, flush=True)
        u0 = ref.links()
        rows = {}
        for pn in precs:
            lat = A.Lattice(n, ext, pm[pn])
            lat.set_links(u0)
            lat.reunitarize()
            lat.mom_refresh(SEED, 1000, beta)
            lat.sync()
            t0 = time.perf_counter()
            e = lat.md_trajectory(dt, nd, level)
            T = time.perf_counter() - t0
            h0 = beta * (e[0] + e[1] / (2 * n))
            h1 = beta * (e[2] + e[3] / (2 * n))
            rows[pn] = {"dH": h1 - h0, "H_old": h0, "seconds": T,
                        "plaquette_start": 1.0 - e[0] / lat.n_plaquettes(),
                        "plaquette_end": 1.0 - e[2] / lat.n_plaquettes()}
            lat.close()
        f64 = rows["fp64"]
        for pn, v in rows.items():
            v["dH_err_vs_fp64"] = v["dH"] - f64["dH"]
            v["H_old_err_vs_fp64"] = v["H_old"] - f64["H_old"]
            v["plaquette_err_vs_fp64"] = v["plaquette_end"] - f64["plaquette_end"]
        print(json.dumps({"study": "precision", "config": name, "nc": n, "extents": ext, "beta": beta, "dt": dt,
                          "level": level, "md_steps": nd, "therm": therm, "results": rows}), flush=True)
        ref.close()


def c3_scan(ntraj, precs=("bf16x3", "fp32")):
    """beta = 4 N^2/10 x {0.8, 0.9, 1.0, 1.1, 1.2} for SU(16), 6x24^3: mean
    plaquette and acceptance over ntraj trajectories from a Haar start."""
    import askja_amd as A

    pm = {"fp32": A.FP32, "bf16x3": A.BF16X3, "fp64": A.FP64}
    n, ext, beta0, dt, level, nd, _, _ = STUDY_CONFIGS["C3"]
    for pn in precs:
        lat = A.Lattice(n, ext, pm[pn])
        for f in (0.8, 0.9, 1.0, 1.1, 1.2):
            beta = beta0 * f
            lat.haar(SEED, 6)
            plaq, acc = [], 0
            t0 = time.perf_counter()
            for t in range(ntraj):
                r = lat.hmc_trajectory(beta, dt, nd, level, SEED, t, reunit=False)
                acc += r.accepted
                plaq.append(1.0 - r.action / lat.n_plaquettes())
            lat.sync()
            T = time.perf_counter() - t0
            print(json.dumps({"study": "c3_scan", "precision": pn, "beta": beta, "ntraj": ntraj,
                              "plaquette_last_half": float(np.mean(plaq[ntraj // 2:])), "acceptance": acc / ntraj,
                              "traj_per_s": ntraj / T}), flush=True)
        lat.close()


def c5_exact(ntraj, therm=20, precs=("bf16x3", "fp64")):
    """The 2D SU(64) workload (128^2, beta = N^2 = 4096, DT 0.005, LEVEL 8,
    200 MD steps: trajectory length 1) in FP64 and in BF16x3 with the FP64
    Metropolis test, Haar start: mean plaquette over ntraj trajectories
    after therm, against the exact single-plaquette value of 2D SU(64)
    (tests/golden/exact_2d_plaquette.json)."""
    import askja_amd as A

    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                       "exact_2d_plaquette.json")))
    want = [p["plaquette"] for p in gold["points"] if p["nc"] == 64 and p["beta"] == 4096.0][0]
    pm = {"bf16x3": A.BF16X3, "fp32": A.FP32, "fp64": A.FP64}
    for pn in precs:
        lat = A.Lattice(64, [128, 128], pm[pn])
        lat.haar(SEED, 7)
        plaq, acc, dh = [], 0, []
        t0 = time.perf_counter()
        for t in range(therm + ntraj):
            if t == therm:
                lat.sync()
                t0 = time.perf_counter()
            r = lat.hmc_trajectory(4096.0, 0.005, 200, 8, SEED, t, reunit=True)
            if t >= therm:
                acc += r.accepted
                dh.append(r.dh)
                plaq.append(1.0 - r.action / lat.n_plaquettes())
            print(f"[c5] {pn} trajectory {t + 1}/{therm + ntraj} plaquette {1.0 - r.action / lat.n_plaquettes():.6f}",
                  file=sys.stderr, flush=True)
        lat.sync()
        T = time.perf_counter() - t0
        tau, _ = tau_int_wolff(plaq) if len(plaq) >= 8 else (1.0, 0)
        m = float(np.mean(plaq))
        e = float(np.std(plaq) * math.sqrt(max(tau, 1.0) / len(plaq)))
        print(json.dumps({"study": "c5_exact_2d", "precision": pn, "ntraj": ntraj, "therm": therm,
                          "plaquette": m, "err": e, "exact": want, "z": (m - want) / e if e > 0 else None,
                          "acceptance": acc / ntraj, "mean_abs_dH": float(np.mean(np.abs(dh))),
                          "traj_per_s": ntraj / T}), flush=True)
        lat.close()


def kernel_study(reps=None):
    """Per-config time of one fused kick+drift launch (HIP events on the
    library's stream) and of one reunitarisation, every precision; flops per
    link (4(D-1) + 1 + the drift's products) 8 N^3 (drift_products)."""
    import askja_amd as A

    cfgs = [("C2", 3, [8, 32, 32, 32], 6), ("C4", 3, [16, 16, 16, 16], 8), ("C1", 8, [10, 10, 10, 10], 8),
            ("C3", 16, [6, 24, 24, 24], 8), ("C5", 64, [128, 128], 8)]
    pm = (("fp64", A.FP64), ("fp64-exact", A.FP64_EXACT), ("fp32", A.FP32), ("bf16x3", A.BF16X3))
    for name, n, ext, level in cfgs:
        for pn, prec in pm:
            if pn == "bf16x3" and n <= 8:
                continue  # N <= 8 runs BF16x3 as FP32
            lat = A.Lattice(n, ext, prec)
            lat.haar(SEED, 2)
            lat.mom_refresh(SEED, 0, 5.9)
            lat.kick_drift(0.01, 0.01, level)
            lat.sync()
            r = reps or (20 if n <= 8 else 3)
            lat.timer_start()
            for _ in range(r):
                lat.kick_drift(0.01, 0.01, level)
            ms = lat.timer_stop() / r
            lat.timer_start()
            lat.reunitarize()
            ms_re = lat.timer_stop()
            flops = (4 * (lat.dim - 1) + 1 + drift_products(n, level, prec)) * 8 * n ** 3 * lat.V * lat.dim
            tf = flops / ms / 1e9
            # the peak of the unit the arithmetic runs on: FP64 / FP32 vector or
            # matrix (the same figures on MI355X), or for BF16x3 the dense BF16
            # MFMA peak against the three BF16 products per FP32 product
            peak, work = (PEAK_TFLOPS["bf16"], 3.0) if (pn == "bf16x3" and n > 8) else \
                (PEAK_TFLOPS["fp32"], 1.0) if pn in ("fp32", "bf16x3") else (PEAK_TFLOPS["fp64"], 1.0)
            print(json.dumps({"study": "kernels", "config": name, "nc": n, "extents": ext, "precision": pn,
                              "ms_kick_drift": round(ms, 4), "tflops": round(tf, 2),
                              "roofline_frac": round(work * tf / peak, 3), "ms_reunitarize": round(ms_re, 3)}),
                  flush=True)
            lat.close()


def example_study(reps=200, ntraj=5):
    """The reference's example parameters (params_N8S10T10B20k.txt: NC 8,
    DIM 10, N0 = N1 = 10, the other eight extents 1, BETA 20000, DT 0.01,
    LEVEL 8, UPDATE2 1000): time of one kick+drift launch (HIP events on the
    library's stream) and of one HMC trajectory of 1000 drifts, per
    precision.  Haar start, counter-based momenta."""
    import askja_amd as A

    ext = [10, 10] + [1] * 8
    for name, prec in (("fp64", A.FP64), ("fp64-exact", A.FP64_EXACT), ("fp32", A.FP32)):
        lat = A.Lattice(8, ext, prec)
        lat.haar(SEED, 1)
        lat.mom_refresh(SEED, 0, 20000.0)
        lat.kick_drift(0.01, 0.01, 8)
        lat.sync()
        lat.timer_start()
        for _ in range(reps):
            lat.kick_drift(0.01, 0.01, 8)
        us = lat.timer_stop() / reps * 1e3
        lat.sync()
        t0 = time.perf_counter()
        acc = 0
        for k in range(ntraj):
            acc += lat.hmc_trajectory(20000.0, 0.01, 1000, 8, SEED, k).accepted
        lat.sync()
        T = (time.perf_counter() - t0) / ntraj
        print(json.dumps({"study": "example", "nc": 8, "extents": ext, "precision": name,
                          "us_kick_drift": round(us, 2), "s_per_trajectory": round(T, 4), "acceptance": acc / ntraj}),
              flush=True)
        lat.close()


def run_chain(lat, args, tag):
    """Haar start, --warmup untimed trajectories, then --steps timed ones with
    the Polyakov loop after each; returns (traj/s, tau, window, value,
    acceptances, dH, plaquettes) and the timed seconds."""
    lat.haar(SEED, 2)
    t_w = time.perf_counter()
    for t in range(args.warmup):
        lat.hmc_trajectory(BETA, DT, args.md_steps, LEVEL, SEED, t, reunit=True)
        if (t + 1) % 100 == 0:
            print(f"[{tag}] warmup {t + 1}/{args.warmup} ({time.perf_counter() - t_w:.1f} s)", file=sys.stderr,
                  flush=True)
    lat.sync()
    poly, dh, acc, plaq = [], [], [], []
    t0 = time.perf_counter()
    last = t0
    for k in range(args.steps):
        r = lat.hmc_trajectory(BETA, DT, args.md_steps, LEVEL, SEED, args.warmup + k, reunit=True)
        poly.append(abs(lat.polyakov()))
        dh.append(r.dh)
        acc.append(r.accepted)
        plaq.append(1.0 - r.action / lat.n_plaquettes())
        now = time.perf_counter()
        if now - last > 30:
            print(f"[{tag}] timed {k + 1}/{args.steps}", file=sys.stderr, flush=True)
            last = now
    lat.sync()
    T = time.perf_counter() - t0
    traj_s = args.steps / T
    tau, window = tau_int_wolff(poly)
    return (traj_s, tau, window, traj_s / tau, acc, dh, plaq), T


# ------------------------------------------------------------------ C4 line
C4_NC, C4_EXT, C4_BETA, C4_DT, C4_NDRIFT, C4_LEVEL = 3, [16, 16, 16, 16], 6.0, 0.02, 50, 8
C4_THERM = 300            # plain trajectories from the Haar start, untimed
# The learned chain then relaxes in V (untimed): a configuration thermalised
# for the plain chain is atypical for the learned chain's target density of
# V, and from it every dt = 0.02 trajectory has dH of about +7.7 (measured:
# no trajectory accepted); 20 trajectories of dt = 0.005 x 200 steps (90 %
# accepted) bring it to where dt = 0.02 is accepted 70 % of the time.
C4_RELAX = (20, 0.005, 200)
# the learned chain's integrator: two time scales, the plain force on every
# MD step of C4_DT, the flow's correction Phi - beta S_W every C4_FLOW_INNER
# steps (1: the one-scale trajectory).  Scanned at this workload with the
# trained one-sweep map: 2 and 5 keep the acceptance at 0.83 and 0.75, 10
# (a coarse step of 0.2) is unstable.
C4_FLOW_INNER = 5
# The learned chain's MD steps per trajectory of the same length (1.0): 55
# steps of 0.01818 with the correction every 5 (12 flow gradients) keep its
# acceptance above 0.7; 50 steps of 0.02 (11 gradients) gave 0.68 over 1500
# trajectories.
C4_LEARNED_NDRIFT = 55
C4_WEIGHTS = os.path.join(ROOT, "askja_amd", "data", "flow_c4_weights_1sweep.json")
TAU_SPAN = 100            # a tau estimate needs a series of >= 100 tau
# The learned series is 1500 trajectories (about 40 s at 40 trajectories/s,
# untimed beyond the first --steps): at 300 its tau_int carried a 27 % error
# (2.08 +- 0.56), at 1500 about 12 %.
MIN_SERIES = {"plain": 3000, "learned": 1500}
MAX_SERIES = {"plain": 20000, "learned": 2000}


def drift_products(nc, level, precision):
    """Complex N x N products of one drift: the Taylor series to LEVEL by
    Horner (LEVEL - 1) then U (1); the MFMA kernels (N >= 9, not the exact
    mode) evaluate it for LEVEL >= 4 as even/odd blocks by Horner in R^2:
    1 + ceil((LEVEL + 1) / 2) - 1 products, then U."""
    import askja_amd as A

    if nc >= 9 and precision != A.FP64_EXACT and level >= 4:
        return (level + 2) // 2 + 1
    return max(level, 1)


def kick_drift_roofline(lat, nc, level, dt, launches):
    """The fused kick+drift launch timed with HIP events on the library's
    stream.  Algorithmic bytes of one MD step: six field passes of 2 N^2
    reals per link (kick: read U, read pi, write pi; drift: read pi, read U,
    write U).  The fused kernel's own compulsory traffic is less (U read
    once, momenta compact Hermitian: read + write link, read + write N^2
    momentum reals) and is reported beside it.  Flops: (4(D-1) + 1 + LEVEL)
    8 N^3 per link (LEVEL: the drift's products, fewer on the MFMA kernels, see
    drift_products), over the average launch."""
    import askja_amd as A

    nl = lat.V * lat.dim
    esz = 8 if lat.precision in (A.FP64, A.FP64_EXACT) else 4
    bytes_per_link = 6 * (2 * nc * nc * esz)
    compulsory_per_link = 2 * (2 * nc * nc * esz) + 2 * (nc * nc * esz)
    flops_per_link = (4 * (lat.dim - 1) + 1 + drift_products(nc, level, lat.precision)) * 8 * nc ** 3
    lat.mom_refresh(SEED, 10**9, 5.9)
    for _ in range(5):
        lat.kick_drift(dt, dt, level)
    lat.timer_start()
    for _ in range(launches):
        lat.kick_drift(dt, dt, level)
    ms = lat.timer_stop() / launches
    achieved = bytes_per_link * nl / (ms * 1e-3) / 1e9
    comp = compulsory_per_link * nl / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "kernel": "fused kick+drift", "ms_per_launch": ms, "bytes_per_link": bytes_per_link,
            "bytes_basis": "one MD step = 6 field passes of 2 N^2 reals per link",
            "compulsory_bytes_per_link": compulsory_per_link, "achieved_compulsory": comp,
            "frac_compulsory": comp / HBM_PEAK_GBS,
            "links_per_launch": nl, "tflops": flops_per_link * nl / (ms * 1e-3) / 1e12}


def series_summary(poly, plaq, acc, dh, seconds, ntimed):
    tau_raw, window = tau_int_wolff(poly)
    tau = max(tau_raw, 1.0) if math.isfinite(tau_raw) else float("nan")
    n = len(poly)
    err = tau_raw * math.sqrt(2.0 * (2 * window + 1) / n) if math.isfinite(tau_raw) else float("nan")

    def mean_err(x):
        x = np.asarray(x, np.float64)
        t, _ = tau_int_wolff(x)
        t = max(t, 1.0) if math.isfinite(t) else 1.0
        return float(x.mean()), float(x.std() * math.sqrt(t / len(x)))

    return {"series": n, "timed": ntimed, "seconds_timed": seconds, "traj_per_s": ntimed / seconds,
            "tau_int_polyakov_raw": tau_raw, "tau_int_polyakov": tau, "tau_int_err": err, "tau_window": window,
            "tau_span_ok": bool(math.isfinite(tau) and window >= 1 and n >= TAU_SPAN * tau),
            "indep_configs_per_s": (ntimed / seconds) / tau, "acceptance": float(np.mean(acc)),
            "abs_polyakov": mean_err(poly), "plaquette": mean_err(plaq), "mean_dH": float(np.mean(dh)),
            "exp_minus_dH": float(np.mean(np.exp(-np.asarray(dh))))}


def learned_dt(args):
    """MD step of the learned chain: its --learned-md-steps steps span the
    plain chain's trajectory length C4_NDRIFT x C4_DT."""
    return C4_DT * C4_NDRIFT / args.learned_md_steps


def c4_chain(lat, kind, first, args, log):
    """--warmup untimed trajectories, then a series whose first `ntimed`
    trajectories are timed; the series is extended (untimed) until it spans
    TAU_SPAN tau or reaches the cap."""
    inner = args.flow_inner if args.flow_inner > 1 else None
    step = (lambda c: lat.hmc_trajectory(C4_BETA, C4_DT, C4_NDRIFT, C4_LEVEL, SEED, c, reunit=True)) \
        if kind == "plain" else \
        (lambda c: lat.flow_hmc_trajectory(C4_BETA, learned_dt(args), args.learned_md_steps, C4_LEVEL, SEED, c,
                                           reunit=True, inner=inner))
    measure = lat.polyakov if kind == "plain" else lat.flow_polyakov
    if kind == "learned":
        n, dt, nd = C4_RELAX
        for k in range(n):
            lat.flow_hmc_trajectory(C4_BETA, dt, nd, C4_LEVEL, SEED, first + k, reunit=True)
        first += n
        log(f"[bench] learned: relaxed ({n} trajectories of {nd} x {dt})")
    for k in range(args.warmup):
        step(first + k)
    first += args.warmup
    # the learned chain's timed region is exactly --steps trajectories; the
    # plain chain (about 100 times cheaper per trajectory) is timed over its
    # whole minimum series
    ntimed = args.steps if kind == "learned" else max(args.steps, MIN_SERIES["plain"])
    poly, plaq, acc, dh = [], [], [], []

    last = [time.perf_counter()]

    def run(n, c0):
        for k in range(n):
            r = step(c0 + k)
            acc.append(r.accepted)
            dh.append(r.dh)
            plaq.append(1.0 - r.action / lat.n_plaquettes())
            poly.append(abs(measure()))
            now = time.perf_counter()
            if now - last[0] > 30:
                log(f"[bench] {kind}: {len(poly)} trajectories")
                last[0] = now

    lat.sync()
    t0 = time.perf_counter()
    run(ntimed, first)
    lat.sync()
    T = time.perf_counter() - t0
    log(f"[bench] {kind}: {ntimed} timed trajectories in {T:.1f} s")
    want = max(ntimed, args.learned_min_series if kind == "learned" else MIN_SERIES[kind])
    while True:
        if len(poly) < want:
            run(want - len(poly), first + len(poly))
        tau, _ = tau_int_wolff(poly)
        need = int(math.ceil(TAU_SPAN * max(tau, 1.0))) if math.isfinite(tau) else MAX_SERIES[kind]
        if len(poly) >= need or len(poly) >= MAX_SERIES[kind]:
            break
        want = min(need, MAX_SERIES[kind])
        log(f"[bench] {kind}: tau {tau:.2f}, extending the series to {want}")
    return series_summary(poly, plaq, acc, dh, T, ntimed)


def flow_roofline(ms_grad, sublayers):
    """FP64 rate of one stout flow gradient against the FP64 peak (operation
    counts per updated link of the forward, log det, log-det gradient and
    backward kernels; 32,768 updated links per sub-layer at 16^4)."""
    links = int(np.prod(C4_EXT)) // 2  # one direction, one parity
    # the backward reads W, rho and the features its forward stored, so it no
    # longer repeats the forward's 15.6 k
    flops = (15.6e3 + 6e3 + 15e3 + 43.4e3) * links * sublayers
    tf = flops / (ms_grad * 1e-3) / 1e12
    return {"bound": "fp64", "achieved": tf, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
            "frac": tf / PEAK_TFLOPS["fp64"], "flops_per_gradient": flops,
            "note": "latency-bound: the kernels need 290-512 registers (one wavefront per SIMD) and a sub-layer's "
                    "32,768 links fill the wavefront slots that allows"}


def flow_kernel_times(lat, reps=3):
    """HIP-event time of one flow gradient (forward + backward of every
    sub-layer, what each FTHMC MD step runs) and of one forward."""
    lat.flow_gradient(C4_BETA)
    lat.timer_start()
    for _ in range(reps):
        lat.flow_gradient(C4_BETA)
    ms_grad = lat.timer_stop() / reps
    lat.timer_start()
    for _ in range(reps):
        lat.flow_apply()
    ms_fwd = lat.timer_stop() / reps
    return ms_grad, ms_fwd


def flow_study(reps=5):
    """HIP-event time of one flow gradient (forward and backward of every
    sub-layer: one FTHMC MD step's force) and of one forward, at 16^4 for
    the rank-one and the stout sub-layers, one and four sweep layers."""
    import askja_amd as A
    from askja_amd import learned as L

    lat = A.Lattice(C4_NC, C4_EXT, A.FP64)
    lat.haar(SEED, 3)
    for t in range(20):
        lat.hmc_trajectory(C4_BETA, C4_DT, C4_NDRIFT, C4_LEVEL, SEED, t, reunit=True)
    rng = np.random.default_rng(1)
    for kind, rho in (("rank-one", 0.3), ("stout", 0.15)):
        for layers in (1, 4):
            mu, par, blk = L.sweep_masks(4, layers)
            bsize = lat.flow_define(mu, par, blk, layers, rho, 1e-3, kind=kind)
            lat.flow_set_weights((rng.standard_normal(layers * bsize) * 0.3).astype(np.float32))
            ms_grad, ms_fwd = flow_kernel_times(lat, reps)
            print(json.dumps({"study": "flow", "kind": kind, "layers": layers, "sublayers": len(mu),
                              "ms_gradient": ms_grad, "ms_forward": ms_fwd,
                              "us_gradient_per_sublayer": 1e3 * ms_grad / len(mu)}), flush=True)
    lat.close()


def fthmc_study(args):
    """The bench's learned chain alone, for profiling: 20 plain trajectories
    from a Haar start, the trained map of --weights, then --steps
    learned-proposal trajectories (the correction every --flow-inner MD
    steps; no relaxation, so the acceptance is the first trajectories')."""
    import askja_amd as A
    from askja_amd import learned as L

    w, recipe = L.load_weights(args.weights)
    lat = A.Lattice(C4_NC, C4_EXT, A.FP64)
    lat.haar(SEED, 3)
    for t in range(20):
        lat.hmc_trajectory(C4_BETA, C4_DT, C4_NDRIFT, C4_LEVEL, SEED, t, reunit=True)
    mu, par, blk = L.sweep_masks(4, recipe["layers"])
    lat.flow_define(mu, par, blk, recipe["layers"], recipe["rho_max"], recipe["delta"], kind=recipe.get("kind"))
    lat.flow_set_weights(w)
    inner = args.flow_inner if args.flow_inner > 1 else None
    lat.sync()
    t0 = time.perf_counter()
    acc = sum(lat.flow_hmc_trajectory(C4_BETA, C4_DT, C4_NDRIFT, C4_LEVEL, SEED, 100 + k, reunit=True,
                                      inner=inner).accepted for k in range(args.steps))
    lat.sync()
    T = time.perf_counter() - t0
    print(json.dumps({"study": "fthmc", "flow_inner": args.flow_inner, "trajectories": args.steps,
                      "traj_per_s": args.steps / T, "acceptance": acc / args.steps}), flush=True)
    lat.close()


def flow2d_study(reps=3):
    """HIP-event time of the two-dimensional spectral map at the C5 size
    (SU(64), 128^2, one sweep = 8 sub-layers of 4096 updated links, random
    weights): one forward, one flow gradient (one FTHMC MD step's force:
    forward keeping the inputs and inverses, the Wilson force of F(V), the
    backward), and the plain FP64 kick+drift of the same lattice."""
    import askja_amd as A
    from askja_amd import learned as L

    lat = A.Lattice(64, [128, 128], A.FP64)
    lat.haar(SEED, 3)
    mu, off, blk = L.spectral_masks(1)
    lat.flow_define(mu, off, blk, 1, 0.4, kind="spectral")
    lat.flow_set_weights((np.random.default_rng(1).standard_normal(161) * 0.3).astype(np.float32))
    lat.flow_gradient(4096.0)  # warm-up
    lat.sync()
    fwd, grad, kd = [], [], []
    for r in range(reps):
        lat.timer_start()
        lat.flow_apply()
        fwd.append(lat.timer_stop())
        lat.timer_start()
        lat.flow_gradient(4096.0)
        grad.append(lat.timer_stop())
        lat.timer_start()
        lat.kick_drift(0.005, 0.005, 8)
        kd.append(lat.timer_stop())
    print(json.dumps({"study": "flow2d", "nc": 64, "extents": [128, 128], "sublayers": len(mu),
                      "ms_forward": min(fwd), "ms_gradient": min(grad), "ms_plain_kick_drift": min(kd),
                      "gradient_over_kick_drift": min(grad) / min(kd)}), flush=True)
    lat.close()


def cpu_baseline_c4(steps_sample):
    """The oracle (FP64 C restatement of the reference, OpenMP over sites in
    the kick and drift like sun.cpp:679,777) on the host cores, plain HMC at
    16^4: `steps_sample` MD steps timed and extrapolated to one trajectory
    of 50 steps + the two Hamiltonian evaluations.  (The learned proposal has
    no CPU counterpart: the reference has none.)"""
    from oracle import oracle as O

    o = O.Oracle(C4_NC, C4_EXT, beta=C4_BETA, level=C4_LEVEL)
    o.haar(SEED, 2)
    o.mom_philox(SEED, 0)
    t0 = time.perf_counter()
    h = o.hamiltonian("pi")
    t_h = time.perf_counter() - t0
    o.updateE(C4_DT / 2)
    t0 = time.perf_counter()
    for _ in range(steps_sample):
        o.updateUandcopyDown(C4_DT)
        o.updateE(C4_DT)
    t_step = (time.perf_counter() - t0) / steps_sample
    t_traj = C4_NDRIFT * t_step + 2 * t_h
    assert math.isfinite(h)
    return {"value": 1.0 / t_traj, "unit": "traj/s (plain HMC)", "cores": O.num_threads(), "kind": "port",
            "sample": f"{steps_sample} FP64 MD steps (kick+drift) of the oracle on 16^4 SU(3), "
                      f"{t_step * 1e3:.1f} ms/step, extrapolated to {C4_NDRIFT} steps + 2 H evaluations"}


def main_c4(args):
    import askja_amd as A
    from askja_amd import learned as L

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    w, recipe = L.load_weights(args.weights)
    if args.layers is not None:  # the first --layers sweep layers of the trained map
        if not 1 <= args.layers <= recipe["layers"]:
            raise SystemExit(f"--layers must be in 1..{recipe['layers']}")
        w = w[: args.layers * recipe["bsize"]]
        recipe = dict(recipe, layers=args.layers, sublayers=args.layers * 8)
    lat = A.Lattice(C4_NC, C4_EXT, A.FP64)
    lat.haar(SEED, 3)
    for t in range(C4_THERM):
        lat.hmc_trajectory(C4_BETA, C4_DT, C4_NDRIFT, C4_LEVEL, SEED, t, reunit=True)
    start = lat.links()
    log(f"[bench] thermalised ({C4_THERM} plain trajectories)")
    plain = c4_chain(lat, "plain", C4_THERM, args, log)
    roof = kick_drift_roofline(lat, C4_NC, C4_LEVEL, C4_DT, args.roofline_launches)
    lat.set_links(start)
    mu, par, blk = L.sweep_masks(4, recipe["layers"])
    kind = recipe.get("kind", "rank-one")
    bsize = lat.flow_define(mu, par, blk, recipe["layers"], recipe["rho_max"], recipe["delta"], kind=kind)
    assert bsize == recipe["bsize"] and w.size == recipe["layers"] * bsize, "weights do not fit this map"
    lat.flow_set_weights(w)
    learned = c4_chain(lat, "learned", 10**7, args, log)
    ms_grad, ms_fwd = flow_kernel_times(lat)
    ok = plain["tau_span_ok"] and learned["tau_span_ok"]
    gain = learned["indep_configs_per_s"] / plain["indep_configs_per_s"]
    for obs in ("abs_polyakov", "plaquette"):
        (m, e), (m0, e0) = learned[obs], plain[obs]
        learned[obs + "_z_vs_plain"] = (m - m0) / math.hypot(e, e0)
    lat.close()
    # secondary: the earlier plain-HMC workload's kernel and trajectory rate
    c2 = A.Lattice(NC, EXT, A.FP64)
    c2.haar(SEED, 2)
    for t in range(5):
        c2.hmc_trajectory(BETA, DT, NDRIFT, LEVEL, SEED, t, reunit=True)
    c2.sync()
    t0 = time.perf_counter()
    nc2 = 20
    for t in range(nc2):
        c2.hmc_trajectory(BETA, DT, NDRIFT, LEVEL, SEED, 5 + t, reunit=True)
    c2.sync()
    c2_traj_s = nc2 / (time.perf_counter() - t0)
    roof_c2 = kick_drift_roofline(c2, NC, LEVEL, DT, args.roofline_launches)
    roof_c2["traffic"] = PMC_TRAFFIC_BYTES["fp64"]
    roof_c2["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0) of this launch, "
                                 "profiles/kick_drift_su3_8x32c_pmc.md")
    c2.close()
    roof["traffic"] = PMC_TRAFFIC_BYTES_C4
    roof["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0) of this launch, "
                              "profiles/kick_drift_su3_16x4_pmc.md")
    out = {
        "metric": "indep. configs/s (traj/s ÷ τ_int Polyakov), learned vs plain HMC; HBM GB/s",
        "value": gain if ok else None,
        "unit": "x (learned / plain independent configs/s)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * learned["seconds_timed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (Haar-random SU(3) start, seed {SEED}, {C4_THERM} untimed plain trajectories; "
                f"the learned chain then relaxes in V for {C4_RELAX[0]} untimed trajectories of "
                f"{C4_RELAX[2]} x {C4_RELAX[1]})",
        "config": {"workload": "learned-proposal (FTHMC) vs plain HMC, SU(3) D=4 16^4 beta=6.0 DT=0.02 "
                               f"{C4_NDRIFT} MD steps LEVEL={C4_LEVEL}, map: {recipe['layers']} {kind} sweep layers "
                               f"({recipe['sublayers']} sub-layers), trained on {recipe['ntrain']} plain configs "
                               f"at beta {recipe['beta_train']}",
                   "nc": C4_NC, "extents": C4_EXT, "beta": C4_BETA, "dt": C4_DT, "md_steps_per_traj": C4_NDRIFT,
                   "level": C4_LEVEL, "precision": "fp64", "weights": os.path.relpath(args.weights, ROOT),
                   "learned_integrator": (f"{args.learned_md_steps} MD steps of {learned_dt(args):.6g} on two time "
                                          f"scales: plain force every step, the flow's correction every "
                                          f"{args.flow_inner} steps ({args.flow_inner * learned_dt(args):.6g}); "
                                          f"{args.learned_md_steps // args.flow_inner + 1} flow gradients per "
                                          f"trajectory")
                   if args.flow_inner > 1 else (f"one time scale, {args.learned_md_steps} MD steps of "
                                                f"{learned_dt(args):.6g}: {args.learned_md_steps + 1} flow gradients "
                                                f"per trajectory")},
        "learned": learned,
        "plain": plain,
        "gain": gain,
        "gain_ceiling_note": (f"plain HMC decorrelates |P| in tau = {plain['tau_int_polyakov']:.2f} trajectories, "
                              "so a proposal of equal cost per trajectory could gain at most about that factor "
                              "on this observable"),
        # both chains in the regime their step sizes were chosen for (the
        # learned chain's two-time-scale integrator runs at about 0.7)
        "acceptance_ok": bool(plain["acceptance"] >= 0.7 and learned["acceptance"] >= 0.7),
        "flow_ms_per_gradient": ms_grad,
        # FP64 work of one flow gradient from the kernels' operation counts per
        # updated link (forward ~15.6 k, log det ~6 k, its gradient ~15 k,
        # backward ~43.4 k flops), 1/8 of the links per sub-layer
        "flow_roofline": flow_roofline(ms_grad, recipe["sublayers"]),
        "flow_ms_per_forward": ms_fwd,
        "plain_ms_per_kick_drift": roof["ms_per_launch"],
        "roofline": roof,
        "secondary_c2": {"workload": "plain HMC, SU(3) D=4 8x32^3 beta=5.9 DT=0.02 LEVEL=6, 100 MD steps",
                         "traj_per_s": c2_traj_s, "trajectories_timed": nc2, "roofline": roof_c2},
    }
    if not ok:
        out["value_null_reason"] = "a tau_int series did not span 100 tau within its cap"
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_c4(args.cpu_steps)
    print(json.dumps(_finite(out), allow_nan=False))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["c4", "c2"], default="c4")
    # c4: learned-proposal trajectories timed; c2: plain trajectories timed.
    # For c2, 6000 trajectories of 100 MD steps (about 160 s): Wolff's window
    # underestimates tau_int(|P|) of that chain on runs too short to show the
    # slow tail of its autocorrelation (50-step trajectories: 20.2 after
    # 3000, 50.1 +- 8.5 after 20000); with 100 steps the estimate is settled
    # from about 6000 trajectories on (20.6, 20.7, 21.2, 20.4 after 6000,
    # 8000, 10000, 12000)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--weights", default=C4_WEIGHTS)
    ap.add_argument("--learned-md-steps", type=int, default=C4_LEARNED_NDRIFT,
                    help="learned chain: MD steps per trajectory (the plain chain's trajectory length)")
    ap.add_argument("--flow-inner", type=int, default=C4_FLOW_INNER,
                    help="learned chain: plain MD steps per kick of the flow's correction (1: one time scale)")
    ap.add_argument("--layers", type=int, default=None,
                    help="c4: use only the first LAYERS sweep layers of the trained map (default: all)")
    ap.add_argument("--learned-min-series", type=int, default=MIN_SERIES["learned"],
                    help="c4: shortest learned tau series (it is extended while shorter than 100 tau)")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16x3"], default="fp64",
                    help="c2 chain: fp64 or fp32; --kernel-only: any")
    ap.add_argument("--md-steps", type=int, default=NDRIFT, help="c2: MD steps per trajectory (length = steps x DT)")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed", action="store_true",
                    help="c2: also run the chain in the FP32 mode (another --steps trajectories)")
    ap.add_argument("--roofline-launches", type=int, default=100)
    ap.add_argument("--copy-probe", action="store_true", help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--kernel-only", choices=["c2", "c4", "c3", "c5"],
                    help="only time --roofline-launches fused kick+drift launches of that lattice (for rocprofv3)")
    ap.add_argument("--study", choices=["precision", "c3scan", "c5exact", "kernels", "example", "flow", "flow2d",
                                        "fthmc"],
                    help="print a precision, beta-scan, exact-2D, per-config kernel or example-params study instead")
    ap.add_argument("--study-configs", default="C2,C3,C5")
    ap.add_argument("--study-therm", type=int, default=20, help="thermalising trajectories (c5exact)")
    ap.add_argument("--study-precisions", default="bf16x3,fp64", help="precisions (c5exact)")
    args = ap.parse_args()
    if args.learned_md_steps < 1:
        raise SystemExit("--learned-md-steps must be positive")
    if args.flow_inner < 1 or args.learned_md_steps % args.flow_inner:
        raise SystemExit(f"--flow-inner must be a positive divisor of the {args.learned_md_steps} learned MD steps")
    if args.copy_probe:
        copy_probe()
        return
    if args.kernel_only:
        import askja_amd as A

        n, ext, level, dt = {"c2": (NC, EXT, LEVEL, DT), "c4": (C4_NC, C4_EXT, C4_LEVEL, C4_DT),
                             "c3": (16, [6, 24, 24, 24], 8, 0.01), "c5": (64, [128, 128], 8, 0.005)}[args.kernel_only]
        lat = A.Lattice(n, ext, {"fp64": A.FP64, "fp32": A.FP32, "bf16x3": A.BF16X3}[args.precision])
        lat.haar(SEED, 2)
        print(json.dumps({"kernel_only": args.kernel_only, "extents": ext,
                          "roofline": kick_drift_roofline(lat, n, level, dt, args.roofline_launches)}))
        return
    if args.steps is None:
        args.steps = 100 if args.workload == "c4" else 6000
    if args.warmup is None:
        args.warmup = 10 if args.workload == "c4" else 300
    if args.study == "precision":
        precision_study(args.study_configs.split(","))
        return
    if args.study == "c3scan":
        c3_scan(max(args.steps, 2))
        return
    if args.study == "kernels":
        kernel_study()
        return
    if args.study == "example":
        example_study()
        return
    if args.study == "flow2d":
        flow2d_study()
        return
    if args.study == "flow":
        flow_study()
        return
    if args.study == "fthmc":
        fthmc_study(args)
        return
    if args.study == "c5exact":
        c5_exact(max(args.steps, 2), args.study_therm, tuple(args.study_precisions.split(",")))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is not built.",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        print("bench.py: --steps must be at least 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4":
        main_c4(args)
        return
    if args.precision == "bf16x3":
        print("bench.py: the c2 chain runs in fp64 or fp32", file=sys.stderr)
        sys.exit(2)
    main_c2(args)


def main_c2(args):
    import askja_amd as A

    prec = A.FP64 if args.precision == "fp64" else A.FP32
    lat = A.Lattice(NC, EXT, prec)
    (traj_s, tau, window, value, acc, dh, plaq), T = run_chain(lat, args, "bench")
    tau_used = max(tau, 1.0) if math.isfinite(tau) else float("nan")
    ok = math.isfinite(tau) and window >= 1 and args.steps >= TAU_SPAN * tau_used
    roof = kick_drift_roofline(lat, NC, LEVEL, DT, args.roofline_launches)
    roof["traffic"] = PMC_TRAFFIC_BYTES[args.precision]
    roof["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0) of this launch, "
                              "profiles/kick_drift_su3_8x32c_pmc.md")
    out = {
        "metric": "indep. configs/s (traj/s / tau_int Polyakov), plain HMC",
        "value": traj_s / tau_used if ok else None,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * T / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == A.FP64 else "f32",
        "data": "synthetic (Haar-random SU(3) start, seed 120351616)",
        "config": {"workload": "plain HMC, SU(3) D=4 8x32^3 beta=5.9 DT=0.02 LEVEL=6 hot start, "
                               f"{args.md_steps} MD steps/trajectory", "nc": NC, "extents": EXT, "beta": BETA, "dt": DT,
                   "level": LEVEL, "md_steps_per_traj": args.md_steps, "precision": args.precision},
        "traj_per_s": traj_s,
        "md_steps_per_s": traj_s * args.md_steps,
        "tau_int_polyakov_raw": tau,
        "tau_int_polyakov": tau_used,
        "tau_int_err": tau * math.sqrt(2.0 * (2 * window + 1) / args.steps) if math.isfinite(tau) else None,
        "tau_window": window,
        "acceptance": float(np.mean(acc)),
        "mean_plaquette": float(np.mean(plaq)),
        "mean_dH": float(np.mean(dh)),
        "exp_minus_dH": float(np.mean(np.exp(-np.asarray(dh)))),
        "roofline": roof,
    }
    if not ok:
        out["value_null_reason"] = "the timed series is shorter than 100 tau_int (or tau is undefined)"
    if prec == A.FP64 and args.mixed:
        lat32 = A.Lattice(NC, EXT, A.FP32)
        (t32, tau32, w32, v32, acc32, dh32, plaq32), _ = run_chain(lat32, args, "bench fp32")
        out["mixed_precision_fp32"] = {
            "traj_per_s": t32, "tau_int_polyakov": tau32,
            "tau_int_err": tau32 * math.sqrt(2.0 * (2 * w32 + 1) / args.steps), "tau_window": w32,
            "acceptance": float(np.mean(acc32)), "mean_plaquette": float(np.mean(plaq32)),
            "exp_minus_dH": float(np.mean(np.exp(-np.asarray(dh32))))}
        lat32.close()
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps, args.md_steps)
    print(json.dumps(_finite(out), allow_nan=False))


if __name__ == "__main__":
    main()
