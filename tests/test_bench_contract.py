"""bench.py's output contract, checked without a GPU: the JSON line's keys
and types, the roofline arithmetic (algorithmic bytes per launch over the
mean launch time, against 8 TB/s, with the counter traffic from
profiles/traffic.json), and the refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _args(**kw):
    import argparse
    d = dict(steps=20, warmup=5, single_pass=False, no_lookahead=False)
    d.update(kw)
    return argparse.Namespace(**d)


def test_line_has_the_contract_keys():
    out = bench.line(1.25, _args(), "2D 4096x4096 PPCG two-rectangle deck", 4096 * 4096, extra_field=1)
    for key, kind in (("metric", str), ("value", float), ("unit", str), ("n_gpus", int), ("steps", int),
                      ("warmup", int), ("ms_per_step", float), ("higher_is_better", bool), ("scaling", str),
                      ("dtype", str), ("data", str), ("config", dict)):
        assert isinstance(out[key], kind), key
    assert out["vs_baseline"] is None
    assert out["metric"] == json.load(open(os.path.join(ROOT, "BASELINE.json")))["metric"]
    assert out["unit"] == "s/timestep" and out["higher_is_better"] is False
    assert out["n_gpus"] == 1 and out["steps"] == 20 and out["warmup"] == 5
    assert out["ms_per_step"] == 1250.0 and out["dtype"] == "f64"
    assert out["config"]["workload"].startswith("2D 4096x4096") and out["config"]["cells"] == 4096 * 4096
    assert out["extra_field"] == 1
    json.dumps(out)


def test_roofline_arithmetic():
    rep = {"cells": 4096 * 4096,
           "kernels": {"ppcg_outer": {"ms": 2.0 * 265.0e-3, "count": 2, "bytes_per_cell": 64.0},
                       "cg_calc_r": {"ms": 0.1, "count": 2, "bytes_per_cell": 24.0}}}
    r = bench.roofline_of(rep, 4096)
    assert r["kernel"] == "ppcg_outer" and r["bound"] == "hbm" and r["unit"] == "GB/s"
    assert r["peak"] == 8000.0
    assert r["bytes_per_launch"] == 64.0 * 4096 * 4096
    assert abs(r["avg_launch_us"] - 265.0) < 1e-9
    want = 64.0 * 4096 * 4096 / 265.0e-6 / 1e9
    assert abs(r["achieved"] - round(want, 1)) < 1e-9
    assert abs(r["frac"] - round(want / 8000.0, 4)) < 1e-12
    traffic = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    assert r["traffic"] == traffic["ppcg_outer@4096"]
    assert bench.roofline_of({"cells": 1, "kernels": {}}, 1) is None


def test_kernel_table():
    rep = {"cells": 100, "kernels": {"k": {"ms": 4.0, "count": 2, "bytes_per_cell": 8.0}}}
    t = bench.kernel_table(rep)
    assert t == {"k": {"launches": 2, "avg_us": 2000.0, "gbs": round(800.0 / 2e-3 / 1e9, 1)}}


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n), "--no-cpu-baseline"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout.strip() == ""
