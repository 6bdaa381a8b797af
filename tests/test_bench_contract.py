"""bench.py keeps its output contract: one JSON line with the headline metric, its unit and the run's
shape, a roofline object consistent with itself, and a clear refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout.strip() == ""


@pytest.mark.gpu
def test_json_line_contract():
    print("bench contract: bench.py --steps 2 --warmup 1 --only bf16 --no-cpu", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "2", "--warmup", "1", "--only",
                        "bf16", "--no-cpu"], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    baseline = json.load(open(os.path.join(ROOT, "BASELINE.json")))
    assert d["metric"] == baseline["metric"]
    assert d["unit"] == "GB/s" and d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["higher_is_better"] is True and d["scaling"] == "weak" and d["vs_baseline"] is None
    assert d["dtype"] == "bf16" and isinstance(d["data"], str) and d["data"].startswith("synthetic")
    assert isinstance(d["config"]["workload"], str) and d["config"]["dims"] == [32] * 6
    # value is the tensor's bytes over the measured time per decomposition
    assert d["value"] == pytest.approx(32 ** 6 * 4 / (d["ms_per_step"] * 1e-3) / 1e9, rel=2e-3)
    assert d["rel_error"] < 1e-4
    rf = d["roofline"]
    assert rf["bound"] in ("hbm", "mfma") and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], abs=1e-3)
    assert rf["achieved"] == pytest.approx(rf["bytes_per_launch"] / (rf["ms_per_launch"] * 1e-3) / 1e9, rel=2e-3)
    assert 0.2 < rf["frac"] < 1.0
