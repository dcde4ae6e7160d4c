"""Learned proposal: training of the gauge-equivariant field transformation
U = F(V) (askja_amd/csrc/flow.hip) and field-transformation HMC (FTHMC) runs.

The transformation's forward, log-Jacobian and backward are the library's HIP
kernels; ``FlowAction`` wraps them in a ``torch.autograd.Function`` so that
torch's optimiser can train the MLP weights (which live in a CUDA tensor and
are handed to the kernels by device pointer).

Training objective.  With q = the distribution of plain-HMC configurations V
at a coupling beta_q and p the target at beta,
    KL(q || p~) = E_q[beta S_W(F(V)) - log J(V)] - beta_q E_q[S_W(V)] + const,
where p~(V) ~ exp(-beta S_W(F(V))) J(V) is the distribution FTHMC samples in
V.  Minimising the first term makes the V-chain at beta behave like plain HMC
at beta_q, while U = F(V) is distributed exactly as exp(-beta S_W(U)) (the
Metropolis test includes log J).
"""
import json
import math
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import askja_amd as A  # noqa: E402


def sweep_masks(dim, layers):
    """Sub-layers of `layers` full sweeps over (direction, parity); every
    sweep shares one MLP weight block."""
    mu, par, blk = [], [], []
    for layer in range(layers):
        for m in range(dim):
            for p in (0, 1):
                mu.append(m)
                par.append(p)
                blk.append(layer)
    return mu, par, blk


def spectral_masks(layers, dim=2):
    """Sub-layers of `layers` sweeps of the spectral map (every direction, the
    four row offsets; in D = 2 both directions alternate); every sweep shares
    one weight block."""
    mu, off, blk = [], [], []
    for layer in range(layers):
        for o in range(4):
            for m in range(dim):
                mu.append(m)
                off.append(o)
                blk.append(layer)
    return mu, off, blk


def init_weights(nblocks, bsize, nf, seed=0, scale=0.1):
    """Small random W1, b1; w2 and b2 zero, so the untrained map is F = id.
    Block layout W1[32][nf], b1[32], w2[32], b2 (bsize = 32 nf + 65)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(nblocks * bsize, np.float32)
    for b in range(nblocks):
        o = b * bsize
        w[o: o + 32 * nf] = rng.standard_normal(32 * nf) * scale
        w[o + 32 * nf: o + 32 * nf + 32] = rng.standard_normal(32) * scale
    return w


def save_weights(path, w, recipe):
    """The trained MLP weights (float32, as hex of their bits so the file
    round-trips exactly) with the recipe that produced them."""
    w = np.ascontiguousarray(np.asarray(w, np.float32))
    with open(path, "w") as f:
        json.dump({"recipe": recipe, "n": int(w.size), "float32_hex": w.view(np.uint32).astype(">u4").tobytes().hex()},
                  f, indent=1)


def load_weights(path):
    """(weights, recipe) written by save_weights."""
    d = json.load(open(path))
    w = np.frombuffer(bytes.fromhex(d["float32_hex"]), dtype=">u4").astype(np.uint32).view(np.float32)
    assert w.size == d["n"]
    return w.copy(), d["recipe"]


def flow_action_function():
    import torch

    class FlowAction(torch.autograd.Function):
        """Phi(w) = beta S_W(F_w(V)) - log J_w(V) for the configuration V held by
        `lat`; forward and backward run the library's HIP kernels."""

        @staticmethod
        def forward(ctx, w, lat, beta):
            torch.cuda.synchronize()
            lat.flow_set_weights(w.detach().contiguous())
            s, lj = lat.flow_gradient(beta, wgrad=True)
            g = torch.from_numpy(lat.flow_weight_gradient()).to(device=w.device, dtype=w.dtype)
            ctx.save_for_backward(g)
            return torch.tensor(beta * s - lj, dtype=torch.float64, device=w.device)

        @staticmethod
        def backward(ctx, go):
            (g,) = ctx.saved_tensors
            return go.to(g.dtype) * g, None, None

    return FlowAction


def plain_chain(lat, beta, dt, ndrift, level, seed, first, n, measure=True, plaq=None):
    """n plain HMC trajectories; returns (|P| series, acceptance, seconds);
    the plaquette series is appended to `plaq` when given."""
    series, acc = [], 0
    lat.sync()
    t0 = tl = time.perf_counter()
    for k in range(n):
        r = lat.hmc_trajectory(beta, dt, ndrift, level, seed, first + k, reunit=True)
        acc += r.accepted
        if time.perf_counter() - tl > 30:  # progress of slow chains
            tl = time.perf_counter()
            print(f"[learned] plain trajectory {k + 1}/{n} ({tl - t0:.0f} s)", file=sys.stderr, flush=True)
        if measure:
            series.append(abs(lat.polyakov()))
            if plaq is not None:
                plaq.append(1.0 - r.action / lat.n_plaquettes())
    lat.sync()
    return np.array(series), acc / max(n, 1), time.perf_counter() - t0


def flow_relax(lat, beta, level, seed, first, n=20, dt=0.005, ndrift=200):
    """FTHMC trajectories of a small step that bring V, started from a
    configuration thermalised for plain HMC (atypical for the learned chain's
    density of V: its first dt = 0.02 trajectories are all rejected), to
    where the chain's own step size is accepted."""
    for k in range(n):
        lat.flow_hmc_trajectory(beta, dt, ndrift, level, seed, first + k, reunit=True)


def plain_force_scale(lat, beta):
    """alpha = <F_Phi, F_W> / <F_W, F_W> at the current V: the multiple of the
    plain Wilson force closest (least squares over every link) to the flow
    force, for the two-time-scale trajectory, whose coarse step then carries
    only Phi - alpha beta S_W.  The momenta are left as they were."""
    cf = 2 * lat.nc / beta
    lat.flow_gradient(beta)
    fphi = lat.flow_link_gradient() * cf
    p0 = lat.momenta()
    lat.set_momenta(np.zeros_like(p0))
    lat.updateE(1.0)  # P = -TA(W): minus the plain force
    fw = -lat.momenta()
    lat.set_momenta(p0)
    return float(np.vdot(fw, fphi).real / np.vdot(fw, fw).real)


def flow_chain(lat, beta, dt, ndrift, level, seed, first, n, log=None, plaq=None, inner=None, alpha=1.0):
    """n FTHMC trajectories; |P| (and the plaquette, into `plaq`) measured on
    the physical configuration U = F(V).  inner: the two-time-scale
    trajectory, the flow's correction kicking every `inner` fine steps."""
    series, acc = [], 0
    lat.sync()
    t0 = tl = time.perf_counter()
    for k in range(n):
        r = lat.flow_hmc_trajectory(beta, dt, ndrift, level, seed, first + k, reunit=True, inner=inner, alpha=alpha)
        acc += r.accepted
        series.append(abs(lat.flow_polyakov()))
        if plaq is not None:
            plaq.append(1.0 - r.action / lat.n_plaquettes())
        now = time.perf_counter()
        if log and ((k + 1) % 50 == 0 or now - tl > 30):
            tl = now
            log(f"[learned] flow trajectory {k + 1}/{n} ({now - t0:.1f} s, acceptance {acc / (k + 1):.2f})")
    lat.sync()
    return np.array(series), acc / max(n, 1), time.perf_counter() - t0


def mean_err(x):
    """Mean and its error sqrt(var tau / n) with tau from tau_int."""
    x = np.asarray(x, np.float64)
    t = tau_int(x)
    t = 1.0 if not np.isfinite(t) or t < 1.0 else t
    return float(x.mean()), float(x.std() * math.sqrt(t / len(x)))


def tau_int(x, S=1.5):
    """tau = 1 + 2 sum rho(t) with Wolff's automatic window (the library's
    askja_tau_int); NaN below 8 samples."""
    if len(x) < 8:
        return float("nan")
    return A.tau_int(x, S)[0]


def train(lat, configs, beta, w0, epochs=2, lr=2e-3, log=None):
    """Adam on E_V[beta S_W(F(V)) - log J(V)] over the training configurations."""
    import torch

    fa = flow_action_function()
    w = torch.tensor(w0, device="cuda", dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([w], lr=lr)
    losses = []
    for ep in range(epochs):
        tot = 0.0
        for V in configs:
            lat.set_links(V)
            opt.zero_grad()
            loss = fa.apply(w, lat, beta)
            loss.backward()
            opt.step()
            tot += float(loss.detach())
        losses.append(tot / len(configs))
        if log:
            log(f"[learned] epoch {ep + 1}/{epochs} loss {losses[-1]:.3f}")
    return w.detach().cpu().numpy(), losses


def c5_study(args, log):
    """The two-dimensional large-N workload: SU(64), 128^2, beta = N^2 = 4096
    (the Gross-Witten point), DT 0.005, LEVEL 8.  A Haar start is
    thermalised by BF16x3 HMC (exact FP64 Metropolis); the spectral map (one
    sweep: both directions, four row offsets) is trained on configurations of
    plain HMC at beta_train, one Adam step per configuration; then plain FP64
    HMC, BF16x3 HMC and FTHMC with the trained map run from the same start.
    Each chain's plaquette is compared with the exact single-plaquette value
    of two-dimensional SU(64) (tests/golden/exact_2d_plaquette.json)."""
    import torch

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    gold = json.load(open(os.path.join(root, "tests", "golden", "exact_2d_plaquette.json")))
    n, ext, beta = 64, [128, 128], 4096.0
    want = [p["plaquette"] for p in gold["points"] if p["nc"] == n and p["beta"] == beta][0]
    dt, nd, level, seed = 0.005, 200, 8, args.seed
    lb = A.Lattice(n, ext, A.BF16X3)
    lb.haar(seed, 7)
    t0 = time.perf_counter()
    for t in range(args.therm):
        r = lb.hmc_trajectory(beta, dt, nd, level, seed, t)
        if (t + 1) % 10 == 0:
            log(f"[c5] thermalising {t + 1}/{args.therm}: plaquette {1.0 - r.action / lb.n_plaquettes():.7f} "
                f"({time.perf_counter() - t0:.0f} s)")
    start = lb.links()
    lf = A.Lattice(n, ext, A.FP64)
    mu, off, blk = spectral_masks(args.layers)
    tmax = args.rho_max if args.rho_max is not None else 0.4
    bsize = lf.flow_define(mu, off, blk, args.layers, tmax, kind="spectral")
    w0 = init_weights(args.layers, bsize, 3, seed=seed, scale=0.1)
    # training, one configuration at a time (2.1 GB each on the host)
    fa = flow_action_function()
    w = torch.tensor(w0, device="cuda", dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([w], lr=args.lr)
    losses = []
    t0 = time.perf_counter()
    for k in range(args.ntrain):
        for j in range(2):
            lb.hmc_trajectory(args.beta_train, dt, nd, level, seed, 10**6 + 2 * k + j)
        lf.set_links(lb.links())
        opt.zero_grad()
        loss = fa.apply(w, lf, beta)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
        log(f"[c5] training step {k + 1}/{args.ntrain}: loss {losses[-1]:.3f}")
    t_train = time.perf_counter() - t0
    wt = w.detach().cpu().numpy()
    lb.close()
    res = {}

    def summarise(name, plaq, acc, secs, ntraj, series=None):
        m, e = mean_err(plaq)
        d = {"ntraj": ntraj, "traj_per_s": ntraj / secs, "acceptance": acc, "plaquette": [m, e],
             "z_vs_exact": (m - want) / e if e > 0 else None}
        if series is not None and len(series) >= 8:
            d["abs_polyakov"] = mean_err(series)
            d["tau_int_polyakov"] = tau_int(series)
        res[name] = d
        log(f"[c5] {name}: plaquette {m:.7f} +- {e:.7f} (exact {want:.7f}), acceptance {acc:.2f}, "
            f"{ntraj / secs:.3f} traj/s")

    for name, prec in (("plain_fp64", A.FP64), ("plain_bf16x3", A.BF16X3)):
        if name not in args.chains.split(","):
            continue
        lat = lf if prec == A.FP64 else A.Lattice(n, ext, prec)
        lat.set_links(start)
        pq = []
        ser, acc, secs = plain_chain(lat, beta, dt, nd, level, seed, 3 * 10**6, args.ntraj, plaq=pq)
        summarise(name, pq, acc, secs, args.ntraj, ser)
        if lat is not lf:
            lat.close()
    if "flow_trained" in args.chains.split(","):
        lf.set_links(start)
        lf.flow_set_weights(wt)
        fq = []
        ser, acc, secs = flow_chain(lf, beta, args.flow_dt, args.flow_ndrift, level, seed, 4 * 10**6,
                                    args.ntraj_flow, log=log, plaq=fq, inner=args.flow_inner)
        summarise("flow_trained", fq, acc, secs, args.ntraj_flow, ser)
        lf.sync()
        lf.timer_start()
        lf.flow_gradient(beta)
        res["flow_trained"]["ms_per_flow_gradient"] = lf.timer_stop()
        lf.timer_start()
        lf.kick_drift(dt, dt, level)
        res["flow_trained"]["ms_per_plain_kick_drift"] = lf.timer_stop()
    out = {"workload": f"SU(64) D=2 128^2 beta=4096 DT={dt} {nd} MD steps LEVEL=8; FTHMC: spectral map, "
                       f"{args.layers} sweep(s) ({len(mu)} sub-layers), t_max {tmax}, trained on {args.ntrain} "
                       f"plain-HMC configs at beta={args.beta_train}, flow dt {args.flow_dt} x {args.flow_ndrift}"
                       + (f", the flow's correction every {args.flow_inner} steps" if args.flow_inner else ""),
           "exact_plaquette": want, "therm": args.therm, "train_seconds": t_train, "loss": losses, "results": res}
    print(json.dumps(out))


def main():
    import argparse

    ap = argparse.ArgumentParser(description="learned-proposal HMC vs plain HMC (SU(3), D=4)")
    ap.add_argument("--extent", type=int, default=16)
    ap.add_argument("--beta", type=float, default=6.0)
    ap.add_argument("--beta-train", type=float, default=5.7)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--ntrain", type=int, default=500)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--ntraj", type=int, default=400)
    ap.add_argument("--therm", type=int, default=200)
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--ndrift", type=int, default=50)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--kind", choices=["rank-one", "stout"], default="stout",
                    help="sub-layer form: stout smearing step (exact 8x8 log det) or rank-one rotation")
    ap.add_argument("--rho-max", type=float, default=None, help="default 0.15 (stout), 0.3 (rank-one)")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--save-weights", help="write the trained weights (and the recipe) to this JSON file")
    ap.add_argument("--no-chains", action="store_true", help="train (and save) only")
    ap.add_argument("--weights", help="load trained weights (save_weights format) instead of training")
    ap.add_argument("--chains", default="plain,flow_untrained,flow_trained",
                    help="comma-separated chains to run: plain, flow_untrained, flow_trained")
    ap.add_argument("--series-out", help="write each chain's |P| and plaquette series to this .npz file")
    ap.add_argument("--workload", choices=["c4", "c5"], default="c4",
                    help="c4: SU(3) D=4 (stout / rank-one map); c5: SU(64) D=2 128^2 (spectral map)")
    ap.add_argument("--ntraj-flow", type=int, default=20, help="c5: FTHMC trajectories")
    ap.add_argument("--flow-dt", type=float, default=0.005, help="c5: FTHMC step")
    ap.add_argument("--flow-ndrift", type=int, default=200, help="c5: FTHMC MD steps per trajectory")
    ap.add_argument("--flow-inner", type=int, default=None,
                    help="two-time-scale FTHMC: the flow's correction kicks every this many plain MD steps")
    args = ap.parse_args()
    if args.workload == "c5":
        import torch

        torch.cuda.init()
        if args.beta_train == 5.7:
            args.beta_train = 0.9 * 4096.0
        if args.chains == "plain,flow_untrained,flow_trained":
            args.chains = "plain_fp64,plain_bf16x3,flow_trained"
        return c5_study(args, lambda m: print(m, file=sys.stderr, flush=True))
    # torch first: the library then binds to torch's HIP runtime, so both
    # share one runtime in this process
    import torch

    torch.cuda.init()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    ext = [args.extent] * 4
    lat = A.Lattice(3, ext, A.FP64)
    # training set: plain HMC at beta_train
    lat.haar(args.seed, 3)
    plain_chain(lat, args.beta_train, args.dt, args.ndrift, args.level, args.seed, 0, args.therm, measure=False)
    configs = []
    if args.weights:
        w, recipe = load_weights(args.weights)
        args.kind, args.layers, args.rho_max = recipe["kind"], recipe["layers"], recipe["rho_max"]
        losses, t_train = recipe["loss"], recipe["train_seconds"]
        plain_chain(lat, args.beta_train, args.dt, args.ndrift, args.level, args.seed, 10**6, 2, measure=False)
        configs.append(lat.links())
    else:
        for k in range(args.ntrain):
            plain_chain(lat, args.beta_train, args.dt, args.ndrift, args.level, args.seed, 10**6 + 2 * k, 2,
                        measure=False)
            configs.append(lat.links())
        log(f"[learned] {len(configs)} training configurations at beta {args.beta_train}")
    mu, par, blk = sweep_masks(4, args.layers)
    rho_max = args.rho_max if args.rho_max is not None else (0.15 if args.kind == "stout" else 0.3)
    bsize = lat.flow_define(mu, par, blk, args.layers, rho_max, 1e-3, kind=args.kind)
    nf = (bsize - 65) // 32  # features: plaquette and two rectangles per plane
    w0 = init_weights(args.layers, bsize, nf, seed=args.seed)
    if not args.weights:
        t0 = time.perf_counter()
        w, losses = train(lat, configs, args.beta, w0, epochs=args.epochs, lr=args.lr, log=log)
        t_train = time.perf_counter() - t0
    if args.save_weights and not args.weights:
        save_weights(args.save_weights, w, {
            "nc": 3, "extents": ext, "beta": args.beta, "beta_train": args.beta_train, "layers": args.layers,
            "sublayers": len(mu), "ntrain": args.ntrain, "epochs": args.epochs, "lr": args.lr, "dt": args.dt,
            "ndrift": args.ndrift, "level": args.level, "seed": args.seed, "kind": args.kind, "rho_max": rho_max,
            "delta": 1e-3, "bsize": bsize, "loss": losses, "train_seconds": t_train})
        log(f"[learned] weights written to {args.save_weights}")
    if args.no_chains:
        return
    # chains at beta from a thermalised start
    lat.set_links(configs[-1])
    plain_chain(lat, args.beta, args.dt, args.ndrift, args.level, args.seed, 2 * 10**6, args.therm, measure=False)
    start = lat.links()
    res, series = {}, {}
    chains = args.chains.split(",")
    if "plain" in chains:
        pq = []
        pl, acc_p, t_p = plain_chain(lat, args.beta, args.dt, args.ndrift, args.level, args.seed, 3 * 10**6,
                                     args.ntraj, plaq=pq)
        res["plain"] = {"traj_per_s": args.ntraj / t_p, "acceptance": acc_p, "tau": tau_int(pl),
                        "abs_polyakov": mean_err(pl), "plaquette": mean_err(pq), "ntraj": args.ntraj}
        series["plain_abs_polyakov"], series["plain_plaquette"] = pl, np.asarray(pq)
    for name, weights in (("flow_untrained", w0), ("flow_trained", w)):
        if name not in chains:
            continue
        lat.set_links(start)
        lat.flow_set_weights(weights)
        flow_relax(lat, args.beta, args.level, args.seed, 5 * 10**6)
        log(f"[learned] {name} chain")
        fq = []
        fl, acc_f, t_f = flow_chain(lat, args.beta, args.dt, args.ndrift, args.level, args.seed, 4 * 10**6, args.ntraj,
                                    log=log, plaq=fq, inner=args.flow_inner)
        res[name] = {"traj_per_s": args.ntraj / t_f, "acceptance": acc_f, "tau": tau_int(fl),
                     "abs_polyakov": mean_err(fl), "plaquette": mean_err(fq), "ntraj": args.ntraj}
        series[name + "_abs_polyakov"], series[name + "_plaquette"] = fl, np.asarray(fq)
    if args.series_out:
        np.savez(args.series_out, **series)
    for k, v in res.items():
        v["tau_err"] = v["tau"] * math.sqrt(2.0 * (2 * A.tau_int(series[k + "_abs_polyakov"], 1.5)[1] + 1) / v["ntraj"])
        v["indep_configs_per_s"] = v["traj_per_s"] / max(v["tau"], 1.0)
        if k != "plain" and "plain" in res:  # agreement with the plain chain, in combined standard errors
            for obs in ("abs_polyakov", "plaquette"):
                (m, e), (m0, e0) = v[obs], res["plain"][obs]
                v[obs + "_z_vs_plain"] = (m - m0) / math.hypot(e, e0)
    out = {"workload": f"SU(3) D=4 {args.extent}^4 beta={args.beta}, FTHMC {args.layers} {args.kind} sweep layer(s) "
                       f"({len(mu)} masked sub-layers), trained on {args.ntrain} plain-HMC configs at "
                       f"beta={args.beta_train}, dt={args.dt}, {args.ndrift} MD steps",
           "train_seconds": t_train, "loss": losses, "results": res}
    if "flow_trained" in res and "plain" in res:
        out["gain_learned_vs_plain"] = res["flow_trained"]["indep_configs_per_s"] / res["plain"]["indep_configs_per_s"]
    if "flow_trained" in res and "flow_untrained" in res:
        out["gain_trained_vs_untrained"] = (res["flow_trained"]["indep_configs_per_s"]
                                            / res["flow_untrained"]["indep_configs_per_s"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
