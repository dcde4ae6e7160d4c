"""CPU restatement (numpy) of the reference's measurement program measure4d
(measurements4d.cpp:930-1306) over stored U-*.conf files.

TEST INFRASTRUCTURE ONLY: the checker for askja_amd/driver/askja-measure.
Follows measurements4d.cpp: loadconf 161-182, getevs 185-309 (Eigen's
ComplexEigenSolver replaced by numpy.linalg.eigvals), WilsonLoop 311-355,
scalarLoopcorr 357-408, PolyakovLoop 410-466, PlaqExp 885-927, main's two
passes 1125-1236 and the file rows 1222-1261.  Returns the numbers of the
Nrav, W24, P5corr and Hist files.
"""
import math

import numpy as np


def load_conf(path, nc, ext):
    V = int(np.prod(ext))
    a = np.fromfile(path, dtype=np.float64)
    assert a.size == V * len(ext) * 2 * nc * nc, path
    return a.view(np.complex128).reshape(V, len(ext), nc, nc)


def _plus(ext, i, d, steps=1):
    c = list(np.unravel_index(i, ext))
    c[d] = (c[d] + steps) % ext[d]
    return int(np.ravel_multi_index(c, ext))


def line(u, ext, start, d):
    w = np.eye(u.shape[-1], dtype=complex)
    pos = start
    for _ in range(ext[d]):
        w = w @ u[pos, d]
        pos = _plus(ext, pos, d)
    return w


def line_mean(u, ext, d):
    """PolyakovLoop (d = 0: mean over x1 of the lines through (0, x1, 0..)) and
    WilsonLoop(evs, d) (mean over x0 of the lines through (x0, 0, ..))."""
    along = 1 if d == 0 else 0
    nc = u.shape[-1]
    vals = []
    for k in range(ext[along]):
        c = [0] * len(ext)
        c[along] = k
        vals.append(abs(np.trace(line(u, ext, int(np.ravel_multi_index(c, ext)), d))) / nc)
    return float(np.mean(vals))


def plaq_exp(u, ext):
    nc, D = u.shape[-1], len(ext)
    ut = us = 0.0
    for x in range(u.shape[0]):
        for j in range(D):
            for k in range(j + 1, D):
                p = u[x, j] @ u[_plus(ext, x, j), k] @ u[_plus(ext, x, k), j].conj().T @ u[x, k].conj().T
                v = 1.0 - np.trace(p).real / nc
                if j == 0:
                    ut += v
                else:
                    us += v
    return ut, us


def slice_traces(u, ext):
    nc, D = u.shape[-1], len(ext)
    sl = u.shape[0] // ext[0]
    tr = np.trace(u[:, D - 1], axis1=-2, axis2=-1).real / nc
    return tr.reshape(ext[0], sl).mean(axis=1)


def evs(w):
    nc = w.shape[0]
    th = np.angle(np.linalg.eigvals(w))
    wind = th.sum() / (2 * math.pi)
    if wind > 0.1 or wind < -0.1:
        th = th - wind * 2 * math.pi / nc
    return np.sort(th)


def getevs(u, ext):
    nc = u.shape[-1]
    wd = np.zeros(nc)
    pd = np.zeros(nc)
    for x0 in range(ext[0]):
        c = [0] * len(ext)
        c[0] = x0
        wd += evs(line(u, ext, int(np.ravel_multi_index(c, ext)), 1))
    for x1 in range(ext[1]):
        c = [0] * len(ext)
        c[1] = x1
        pd += evs(line(u, ext, int(np.ravel_multi_index(c, ext)), 0))
    return wd / ext[0], pd / ext[1]


def measure(paths, nc, ext, beta):
    """paths: the configurations SKIP .. TAKEN-1 in order."""
    D = len(ext)
    vol = ext[0]
    width = math.pi / (2 * nc)
    bins = int(math.floor(2 * math.pi / width))
    whist, phist = np.zeros(bins, int), np.zeros(bins, int)
    rows = []
    for path in paths:
        u = load_conf(path, nc, ext)
        P, Wx = line_mean(u, ext, 0), line_mean(u, ext, 1)
        W2 = P * P + Wx * Wx + sum(line_mean(u, ext, d) ** 2 for d in range(2, D) if ext[d] > 1)
        ut, us = plaq_exp(u, ext)
        wd, pd = getevs(u, ext)
        for i in range(nc):
            whist[int(math.floor((wd[i] + math.pi) / width))] += 1
            phist[int(math.floor((pd[i] + math.pi) / width))] += 1
        rows.append({"P": P, "Wx": Wx, "W2": W2, "Ut": ut * beta / vol, "Us": us * beta / vol,
                     "P5": slice_traces(u, ext)})
    n = len(rows)
    mean = {k: float(np.mean([r[k] for r in rows])) for k in ("P", "Wx", "W2", "Ut", "Us")}
    p5 = np.mean([r["P5"] for r in rows], axis=0)
    err = {k: math.sqrt(sum((r[k] - mean[k]) ** 2 for r in rows) / (n * n)) for k in ("P", "Wx", "W2", "Ut", "Us")}
    n0 = ext[0]
    corr = [sum(0.5 * ((r["P5"][x] - p5[x]) * (r["P5"][0] - p5[0]) + (r["P5"][n0 - x - 1] - p5[n0 - x - 1])
                       * (r["P5"][0] - p5[0])) for r in rows) / (n * n) for x in range(n0 // 2)]
    running = []
    acc = {k: 0.0 for k in ("P", "Wx", "Ut", "Us")}
    for i, r in enumerate(rows):
        for k in acc:
            acc[k] += r[k]
        if i > 0:
            running.append([acc[k] / (i + 1) for k in ("P", "Wx", "Ut", "Us")])
    centers = [-math.pi + (i + 0.5) * width for i in range(bins)]
    return {"mean": mean, "err": err, "corr": corr, "running": running,
            "hist": [(centers[i], phist[i] / n, whist[i] / n) for i in range(bins)]}
