"""The `askja <params>` program (askja_amd/driver) against the reference's
file contract (sun.cpp:803-1373) and the oracle's first trajectory.

CPU tests: argument and params-file errors end the program before any GPU
call.  GPU tests: a short run writes the reference's output files with the
reference's layouts, its first trajectory matches the oracle bit for bit
(FP64_EXACT), and a resumed run continues the numbering.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

ASKJA = os.path.join(ROOT, "askja_amd", "driver", "askja")

PARAMS = """#Parameters_file
NC              3
DIM             4
N0              4
N1              4
N2              4
N3              4
BETA            5.9
DT              0.05
STEPS           {steps}
UPDATE          10
UPDATE2         20
CONFTAKE        1
LEVEL           6
VERBOSE         0
RUNCONTINUE     0
RANDOM          0
"""
TAG = "N3-Nt4-Ns4-BETA5.90-DT0.0500-U1.0000-CT1"


def _run(args, cwd, timeout=300):
    if not os.path.exists(ASKJA):
        pytest.fail("askja is not built (run __graft_entry__.build())")
    return subprocess.run([ASKJA] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def read_master(path, dim):
    """Restatement of measurements4d.cpp:752-882 (load_parameters of the
    measure programs): positional tokens, names ignored; with RUNCONTINUE 1
    it reads TIME-STEP, TAKEN, NACC, NAC1, NREC."""
    tok = open(path).read().split()
    it = iter(tok)
    next(it)
    vals = {}
    next(it); vals["NC"] = int(next(it))
    next(it); vals["DIM"] = int(next(it))
    assert vals["DIM"] == dim
    vals["x"] = []
    for _ in range(dim):
        next(it); vals["x"].append(int(next(it)))
    for k, f in (("BETA", float), ("DT", float), ("STEPS", int), ("UPDATE", int), ("UPDATE2", int),
                 ("CONFTAKE", int), ("LEVEL", int), ("VERBOSE", int), ("RUNCONTINUE", int), ("RANDOM", int)):
        next(it); vals[k] = f(next(it))
    if vals["RUNCONTINUE"]:
        for k in ("TIME-STEP", "TAKEN", "NACC", "NAC1", "NREC"):
            name = next(it)
            assert name == k
            vals[k] = int(next(it))
    return vals


@pytest.mark.gpu
def test_bf16x3_cold_start_warns(tmp_path):
    """--precision bf16x3 on a fresh run (cold start) at N >= 9 prints a
    warning that such a chain rejects most trajectories until disordered."""
    params = PARAMS.format(steps=20).replace("NC              3", "NC              12")
    params = params.replace("N1              4\nN2              4\nN3              4", "N1              2\nN2              2\nN3              2")
    (tmp_path / "p.txt").write_text(params)
    r = _run(["p.txt", "--precision", "bf16x3"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "warning: bf16x3 from the cold start" in r.stderr


def test_missing_and_truncated_params(tmp_path):
    r = _run([str(tmp_path / "nope.txt")], tmp_path)
    assert r.returncode != 0
    assert "could not be opened" in r.stdout and "Finished with errors" in r.stdout
    (tmp_path / "short.txt").write_text("#Parameters_file\nNC 3\nDIM 4\nN0 4\n")
    r = _run([str(tmp_path / "short.txt")], tmp_path)
    assert r.returncode != 0 and "Finished with errors" in r.stdout
    assert not (tmp_path / "output-DIM4").exists()
    r = _run([], tmp_path)
    assert r.returncode != 0


@pytest.mark.gpu
def test_run_files_and_first_trajectory(tmp_path, oracle):
    import askja_amd as A

    (tmp_path / "p.txt").write_text(PARAMS.format(steps=60))
    r = _run(["p.txt", "--precision", "fp64-exact"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Finished successfully" in r.stdout
    assert "tau_int(|P|) needs at least 8 trajectories (3)" in r.stdout
    out = tmp_path / "output-DIM4"
    cons = (out / f"constraints-{TAG}.dat").read_text().splitlines()
    assert cons[0] == "#Time\t Gauss-law\t Total Energy\t E_L\t E_M\t Unitary-Violation"
    rows = [list(map(float, l.split("\t"))) for l in cons[1:]]
    assert [round(r_[0], 6) for r_ in rows] == [0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0]
    assert all(len(r_) == 6 for r_ in rows)
    wil = (out / f"Wilson-{TAG}.dat").read_text().splitlines()
    assert wil[0] == "#Time\t <W>\t Wilson-Line-Phases"
    for l in wil[1:]:
        f = l.rstrip("\t").split("\t")
        assert len(f) == 1 + 3 + 1
        ph = list(map(float, f[2:]))
        # phases of an SU(3) matrix sum to 0 mod 2 pi (printed with %g)
        assert ph == sorted(ph)
        assert min(abs(sum(ph)), abs(abs(sum(ph)) - 2 * np.pi)) < 1e-4
    confs = sorted(glob.glob(str(out / f"U-{TAG}-*.conf")))
    assert len(confs) == 3
    V, D, N = 256, 4, 3
    for c in confs:
        assert os.path.getsize(c) == V * D * N * N * 16
    m = read_master(out / f"master-{TAG}.params", 4)
    assert m["RUNCONTINUE"] == 1 and m["TAKEN"] == 3 and m["TIME-STEP"] == 60
    assert m["NACC"] + m["NREC"] == 3 and m["BETA"] == 5.9 and m["x"] == [4, 4, 4, 4]

    # trajectory 1 on the oracle: cold start, reference momenta, 19 drifts
    o = oracle.Oracle(N, [4, 4, 4, 4], beta=5.9, level=6)
    o.initializeU()
    o.initializeE(120351616)
    h0 = o.hamiltonian("pi")
    t0_em = o.energy_mag()
    o.md_trajectory(0.05, 19)
    h1 = o.hamiltonian("Pi")
    lat = A.Lattice(N, [4, 4, 4, 4], A.FP64)
    u_draw = ctypes_uniform(A, lat, 120351616, 0)
    accepted = h1 < h0 or u_draw < np.exp(h0 - h1)
    conf0 = np.fromfile(confs[0], dtype=np.complex128).reshape(V, D, N, N)
    if accepted:
        assert np.array_equal(conf0, o.field("u"))
    else:
        assert np.array_equal(conf0, np.broadcast_to(np.eye(N), conf0.shape))
    assert ("ACCEPT" if accepted else "REJECT") in r.stdout.split("Final energy")[1].splitlines()[1]
    # t = 0 row: E_M = beta * EnergyMag of the cold start = 0, U-violation 0
    assert rows[0][4] == 0.0 and t0_em == 0.0 and rows[0][5] == 0.0


def ctypes_uniform(A, lat, seed, counter):
    import ctypes as C
    v = C.c_double()
    assert A.lib().askja_uniform(lat._h, seed, counter, C.byref(v)) == 0
    return v.value


@pytest.mark.gpu
def test_resume_continues_numbering(tmp_path):
    (tmp_path / "p.txt").write_text(PARAMS.format(steps=40))
    r = _run(["p.txt"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    out = tmp_path / "output-DIM4"
    master = out / f"master-{TAG}.params"
    assert read_master(master, 4)["TAKEN"] == 2
    r = _run([str(master)], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Continuing run at t=2.000000" in r.stdout
    confs = sorted(glob.glob(str(out / f"U-{TAG}-*.conf")))
    assert [os.path.basename(c).rsplit("-", 1)[1] for c in confs] == ["0.conf", "1.conf", "2.conf", "3.conf"]
    m = read_master(master, 4)
    assert m["TAKEN"] == 4 and m["TIME-STEP"] == 80
    cons = (out / f"constraints-{TAG}.dat").read_text().splitlines()
    times = [round(float(l.split("\t")[0]), 6) for l in cons[1:]]
    assert times == [0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0]
    # the seed travels after NTOT, where neither reference reader looks
    assert open(master).read().split()[-2] == "SEED"


@pytest.mark.gpu
def test_resume_reproduces_uninterrupted_run(tmp_path):
    """fp64-exact, CONFTAKE 1: a 4-trajectory run and a 2 + 2 run resumed from
    the first run's master file (sun.cpp:890-919, 1158-1171) write the same
    bytes: configurations U-*-2.conf and U-*-3.conf, the constraints and
    Wilson rows, and the master file's counters.  The resumed run continues
    the Philox stream at counter NTOT (the seed travels after NTOT), so the
    trajectories after the break draw the momenta and uniforms the
    uninterrupted run draws.  (With CONFTAKE > 1 the reference rounds NACC
    down to a multiple of CONFTAKE on resume, sun.cpp:904, and the
    trajectories after the last stored configuration are run again, so only
    the chain from that configuration on is reproduced.)"""
    whole, split = tmp_path / "whole", tmp_path / "split"
    whole.mkdir()
    split.mkdir()
    (whole / "p.txt").write_text(PARAMS.format(steps=80))
    (split / "p.txt").write_text(PARAMS.format(steps=40))
    r = _run(["p.txt", "--precision", "fp64-exact"], whole)
    assert r.returncode == 0, r.stdout + r.stderr
    r = _run(["p.txt", "--precision", "fp64-exact"], split)
    assert r.returncode == 0, r.stdout + r.stderr
    master = split / "output-DIM4" / f"master-{TAG}.params"
    r = _run([str(master), "--precision", "fp64-exact"], split)
    assert r.returncode == 0, r.stdout + r.stderr
    ow, os_ = whole / "output-DIM4", split / "output-DIM4"
    for k in range(4):
        a = (ow / f"U-{TAG}-{k}.conf").read_bytes()
        b = (os_ / f"U-{TAG}-{k}.conf").read_bytes()
        assert a == b, k
    for name in (f"constraints-{TAG}.dat", f"Wilson-{TAG}.dat"):
        assert (ow / name).read_text() == (os_ / name).read_text(), name
    mw, ms = read_master(ow / f"master-{TAG}.params", 4), read_master(master, 4)
    for k in ("TIME-STEP", "TAKEN", "NACC", "NAC1", "NREC"):
        assert mw[k] == ms[k], (k, mw[k], ms[k])
    # the whole master file agrees but for STEPS, which is the split run's
    # own (40 per invocation against 80)
    tw, ts = open(ow / f"master-{TAG}.params").read().split(), open(master).read().split()
    k = tw.index("STEPS") + 1
    assert (tw[k], ts[k]) == ("80", "40")
    assert tw[:k] + tw[k + 1:] == ts[:k] + ts[k + 1:]


@pytest.mark.gpu
def test_truncated_configuration_ends_resume(tmp_path):
    """A resumed run whose last configuration file is short (an interrupted
    write; the reference's loadconf reads it with gsl_matrix_complex_fread,
    sun.cpp:751-763) stops with a non-zero status before any output: the
    constraints and Wilson files and the master file are unchanged and no new
    configuration appears."""
    (tmp_path / "p.txt").write_text(PARAMS.format(steps=40))
    r = _run(["p.txt"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    out = tmp_path / "output-DIM4"
    conf = out / f"U-{TAG}-1.conf"
    data = conf.read_bytes()
    conf.write_bytes(data[: len(data) // 2])
    before = {p.name: p.read_bytes() for p in out.iterdir()}
    master = out / f"master-{TAG}.params"
    r = _run([str(master)], tmp_path)
    assert r.returncode != 0
    assert "Error loading fields file" in r.stdout and "Finished with errors" in r.stdout
    after = {p.name: p.read_bytes() for p in out.iterdir()}
    assert after == before


@pytest.mark.gpu
def test_nan_energies_reject_and_restore(tmp_path):
    """DT 1000 overflows the MD within a trajectory (links and momenta become
    inf / NaN): every trajectory's final energy is not finite, so it is
    rejected as in the reference (a NaN or +inf H fails both accept tests,
    sun.cpp:1258, 1288, 1304-1310), NREC counts every trajectory, and every
    stored configuration is the restored cold start, bit for bit."""
    params = PARAMS.format(steps=60).replace("DT              0.05", "DT              1000")
    (tmp_path / "p.txt").write_text(params)
    r = _run(["p.txt"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("REJECT") == 3 and "ACCEPT" not in r.stdout
    tag = TAG.replace("DT0.0500-U1.0000", "DT1000.0000-U20000.0000")
    out = tmp_path / "output-DIM4"
    m = read_master(out / f"master-{tag}.params", 4)
    assert m["NREC"] == 3 and m["NACC"] == 0 and m["TAKEN"] == 3
    V, D, N = 256, 4, 3
    cold = np.broadcast_to(np.eye(N, dtype=np.complex128), (V, D, N, N))
    for k in range(3):
        conf = np.fromfile(out / f"U-{tag}-{k}.conf", dtype=np.complex128).reshape(V, D, N, N)
        assert np.array_equal(conf, cold), k
    finals = [l for l in r.stdout.splitlines() if "Final energy" in l]
    assert len(finals) == 3 and all(("nan" in l or "inf" in l) for l in finals), finals


MPARAMS = """#Parameters_file
NC              3
DIM             4
N0              4
N1              6
N2              2
N3              2
BETA            5.9
DT              0.05
STEPS           100
UPDATE          20
UPDATE2         20
CONFTAKE        1
LEVEL           6
VERBOSE         0
RUNCONTINUE     0
RANDOM          0
"""
MTAG = "N3-Nt4-Ns6-BETA5.90-DT0.0500-U1.0000-CT1"


def _rows(path):
    return [list(map(float, l.rstrip("\t").split("\t"))) for l in open(path).read().splitlines() if not l.startswith("#")]


@pytest.mark.gpu
def test_measure_program_vs_oracle(tmp_path):
    """askja-measure (measure4d's command line and files) on configurations
    written by askja, against the oracle's restatement of measure4d on the
    same files: Nrav running and final rows, W24, P5corr, Hist (numbers as
    printed, 6 significant digits)."""
    from oracle import measure_oracle as mo

    (tmp_path / "p.txt").write_text(MPARAMS)
    r = _run(["p.txt", "--reunit"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    master = tmp_path / "output-DIM4" / f"master-{MTAG}.params"
    taken = read_master(master, 4)["TAKEN"]
    assert taken == 5
    meas = os.path.join(ROOT, "askja_amd", "driver", "askja-measure")
    r = subprocess.run([meas, str(master), "1", "0", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Finished successfully" in r.stdout, r.stdout + r.stderr
    ext = [4, 6, 2, 2]
    paths = [str(tmp_path / "output-DIM4" / f"U-{MTAG}-{k}.conf") for k in range(1, taken)]
    want = mo.measure(paths, 3, ext, 5.9)
    rav = tmp_path / "rav-DIM4"

    def close(a, b):
        return abs(a - b) <= 1e-5 * abs(b) + 1e-9

    nrav = _rows(rav / f"Nrav-{MTAG}.dat")
    assert len(nrav) == len(want["running"]) + 1
    for row, run in zip(nrav[:-1], want["running"]):
        assert all(close(a, b) for a, b in zip([row[1], row[3], row[5], row[7]], run)), (row, run)
        assert row[2] == row[4] == row[6] == row[8] == 0.0
    fin = nrav[-1]
    m, e = want["mean"], want["err"]
    exp = [taken, m["P"], e["P"], m["Wx"], e["Wx"], m["Ut"], e["Ut"], m["Us"], e["Us"], 4, 6, 5.9]
    assert all(close(a, b) for a, b in zip(fin, exp)), (fin, exp)
    w24 = _rows(rav / f"W24-{MTAG}.dat")
    assert all(close(a, b) for a, b in zip(w24[0], [3, 5.9, 6, m["W2"], e["W2"], taken - 1, 1])), w24
    p5 = _rows(rav / f"P5corr-{MTAG}.dat")
    assert [row[0] for row in p5] == [0, 1]
    assert all(close(row[1], c) for row, c in zip(p5, want["corr"])), (p5, want["corr"])
    hist = _rows(rav / f"Hist-{MTAG}.dat")
    assert len(hist) == len(want["hist"]) == 12
    for row, h in zip(hist, want["hist"]):
        assert all(close(a, b) for a, b in zip(row, h)), (row, h)
    assert abs(sum(row[1] for row in hist) - 3) < 1e-9 and abs(sum(row[2] for row in hist) - 3) < 1e-9


@pytest.mark.gpu
def test_run_reports_tau_int(tmp_path):
    """Twelve trajectories: the run ends with tau_int of |P| (Wolff's window,
    askja_tau_int) and the independent configurations per second."""
    import re

    (tmp_path / "p.txt").write_text(PARAMS.format(steps=240))
    r = _run(["p.txt"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"tau_int\(\|P\|\) = ([0-9.eE+-]+) \(window (\d+)\) over (\d+) trajectories, "
                  r"([0-9.eE+-]+) independent configurations/s", r.stdout)
    assert m, r.stdout[-2000:]
    tau, window, ntraj, rate = float(m.group(1)), int(m.group(2)), int(m.group(3)), float(m.group(4))
    assert ntraj == 12 and 0 < tau < 12 and 0 <= window < 6 and rate > 0


@pytest.mark.gpu
def test_reference_example_params_end_to_end(tmp_path):
    """The reference's own params file (params_N8S10T10B20k.txt, byte for
    byte: NC 8, DIM 10, 100000 MD steps, RANDOM 1) runs to the end: 100
    trajectories, 10 stored configurations of the U-*.conf size, the
    constraints/Wilson rows every UPDATE, the closing tau_int line."""
    import shutil

    shutil.copy(os.path.join(ROOT, "tests", "golden", "params_N8S10T10B20k.txt"), tmp_path / "params.txt")
    r = _run(["params.txt"], tmp_path, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Finished successfully" in r.stdout and "over 100 trajectories" in r.stdout
    out = tmp_path / "output-DIM10"
    tag = "N8-Nt10-Ns10-BETA20000.00-DT0.0100-U10.0000-CT10"
    confs = sorted(glob.glob(str(out / f"U-{tag}-*.conf")))
    assert len(confs) == 10
    assert all(os.path.getsize(c) == 100 * 10 * 8 * 8 * 16 for c in confs)
    rows = [l for l in open(out / f"constraints-{tag}.dat").read().splitlines() if not l.startswith("#")]
    assert len(rows) == 1 + 100000 // 500
