"""bench.py's output contract: one JSON line with the metric, the timing,
the workload, the roofline object of the dominant kernel; a request for more
than one GPU exits with a message (every box here has one)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.gpu
def test_bench_prints_one_json_line(gpu):
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "bench.py"), "--steps", "16", "--warmup", "2",
                        "--no-cpu", "--no-plain", "--no-serve", "--no-accuracy"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    out = json.loads(lines[0])
    for key, typ in (("metric", str), ("value", float), ("unit", str), ("n_gpus", int), ("steps", int),
                     ("warmup", int), ("ms_per_step", float), ("higher_is_better", bool), ("scaling", str),
                     ("dtype", str), ("data", str), ("config", dict), ("roofline", dict)):
        assert isinstance(out[key], typ), key
    assert out["n_gpus"] == 1 and out["steps"] == 16 and out["warmup"] == 2
    assert out["value"] > 0 and out["dtype"] == "f32" and out["vs_baseline"] is None
    assert "workload" in out["config"] and out["config"]["engine"] in ("fused", "hybrid", "stage", "resident")
    rl = out["roofline"]
    assert rl["bound"] in ("hbm", "valu") and rl["unit"] == "GB/s" and rl["peak"] == 8000.0
    assert abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-12
    if "valu" in rl:
        v = rl["valu"]
        assert v["unit"] == "lane-instructions/s" and 0 < v["frac"] < 1
        assert abs(v["achieved"] - v["lane_insts_per_point_step"] * rl["kernel_points"] / (rl["avg_launch_us"] * 1e-6)) \
            <= 1e-9 * v["achieved"]
    # 16 steps: below the 128 a call needs before it replays step-block graphs
    assert out["config"]["graphs_replayed"] is False
