"""bench.py's command-line contract: one JSON line with the metric, the
roofline of the dominant kernel and the CPU baseline; more than one GPU is
refused with a message."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

BENCH = os.path.join(ROOT, "bench.py")


def _run(args, timeout):
    return subprocess.run([sys.executable, BENCH] + args, cwd=ROOT, capture_output=True, text=True,
                          timeout=timeout)


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2", "--steps", "1", "--warmup", "0"], 120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_learned_md_steps_must_hold_whole_correction_intervals():
    # 53 steps cannot be cut into intervals of --flow-inner 5: refused
    # before any GPU work
    r = _run(["--learned-md-steps", "53"], 120)
    assert r.returncode != 0
    assert "divisor of the 53 learned MD steps" in r.stderr
    r = _run(["--learned-md-steps", "0"], 120)
    assert r.returncode != 0 and "positive" in r.stderr


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_json_line_contract():
    """The line the driver records (--steps 20 --warmup 5): the learned and
    plain chains' tau_int come from series of at least 100 tau with a Wolff
    window of at least 1, tau >= 1, so the value is a measurement."""
    # the driver's --steps / --warmup on the first sweep layer of the trained
    # map and a shorter learned series than the default 1500 (still extended
    # to 100 tau), so that the test stays within a minute
    r = _run(["--steps", "20", "--warmup", "5", "--roofline-launches", "5", "--cpu-steps", "2",
              "--layers", "1", "--learned-min-series", "120"], 900)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 20 and d["warmup"] == 5
    assert d["ms_per_step"] > 0 and d["higher_is_better"] is True
    assert d["dtype"] == "f64" and d["vs_baseline"] is None
    assert "learned vs plain" in d["metric"] and "workload" in d["config"]
    for chain in ("learned", "plain"):
        c = d[chain]
        assert c["tau_int_polyakov"] >= 1.0 and c["tau_window"] >= 1, (chain, c)
        assert c["series"] >= 100 * c["tau_int_polyakov"], (chain, c)
        assert c["tau_span_ok"] and c["traj_per_s"] > 0 and 0.0 < c["acceptance"] <= 1.0
    assert d["learned"]["timed"] == 20 and "1 stout sweep layers" in d["config"]["workload"]
    assert d["value"] is not None and d["value"] > 0
    assert abs(d["value"] - d["learned"]["indep_configs_per_s"] / d["plain"]["indep_configs_per_s"]) < 1e-12
    for obs in ("abs_polyakov_z_vs_plain", "plaquette_z_vs_plain"):
        assert abs(d["learned"][obs]) < 5.0, (obs, d["learned"][obs])
    rl = d["roofline"]
    assert rl["bound"] == "hbm" and rl["unit"] == "GB/s" and rl["peak"] == 8000.0
    assert 0 < rl["frac"] < 1 and abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-12
    assert rl["traffic"] > 0
    assert d["secondary_c2"]["roofline"]["traffic"] > 0
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["cores"] >= 1 and cb["value"] > 0


@pytest.mark.gpu
def test_short_series_gives_null_value():
    """The plain 8x32^3 line over 12 trajectories: tau_int (about 20 on this
    chain) cannot be estimated from a series shorter than 100 tau, so the
    value is null with the reason; the JSON stays valid (no NaN)."""
    r = _run(["--workload", "c2", "--steps", "12", "--warmup", "2", "--roofline-launches", "5", "--cpu-steps", "1",
              "--mixed"], 600)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads([l for l in r.stdout.splitlines() if l.strip()][-1])
    assert d["value"] is None and "100 tau" in d["value_null_reason"]
    assert d["traj_per_s"] > 0 and d["roofline"]["traffic"] > 0
    assert d["mixed_precision_fp32"]["traj_per_s"] > 0
