"""bench.py's command-line contract: one GPU only, and one JSON line carrying
the fields the measurement needs (metric, value, roofline, dtype, config, the
copy-kernel ceiling ...)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr and not r.stdout.strip()


@pytest.mark.gpu
def test_small_run_prints_one_complete_json_line():
    print("[bench] 64^3, 4 steps", flush=True)
    r = subprocess.run([sys.executable, BENCH, "--size", "64", "--steps", "4", "--warmup", "1",
                        "--legs", "f64,f16s", "--no-cpu", "--no-adjoint"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    line = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "f16s_storage"):
        assert key in line, key
    assert line["n_gpus"] == 1 and line["steps"] == 4 and line["dtype"] == "f64"
    roof = line["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert 0 < roof["frac"] < 1 and abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-3
    assert line["config"]["workload"].startswith("D3Q19 AA SOA 64^3")
    f16 = line["f16s_storage"]
    assert f16["steps_compared"] >= 4 and 0 < f16["rel_l2_u_vs_f64"] < 1e-3
    copy = line["copy_kernel"]
    assert copy["unit"] == "GB/s" and 0 < copy["value"] < 8000.0
    assert abs(copy["frac"] - copy["value"] / 8000.0) < 1e-3
    assert abs(copy["frac_reached"] - roof["achieved"] / copy["value"]) < 1e-3
