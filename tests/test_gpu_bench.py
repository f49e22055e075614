"""bench.py end to end at a small size: the JSON line it prints carries the
fields of the bench contract (value, roofline, the FP64 comparison), and the
whole-interval job finds every eigenvalue of the closed form."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("basis", ["fp64", "fp32"])
def test_bench_json_line_small_grid(basis):
    print("\n[gpu-test] bench.py on the 24^3 Laplacian, 5 slices, basis %s" % basis, flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--grid", "24",
                        "--nslices", "5", "--steps", "3", "--warmup", "1", "--basis", basis,
                        "--wave", "3", "--no-cpu-baseline", "--budget", "200"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    print(r.stderr[-3000:], flush=True)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.strip().splitlines() if ln.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["unit"] == "eigenpairs/s" and d["value"] > 0 and d["ms_per_step"] > 0
    assert "24^3" in d["config"]["workload"] and d["config"]["basis"] == basis
    job = d["job"]
    assert job["found"] == job["exact"] > 0 and not job["slices_not_exact"]
    assert job["max_rel_err"] < 1e-10
    roof = d["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert 0 < roof["frac"] < 1 and roof["kernel"].startswith("k_")
    assert d["step"]["slices"] == 3  # one wave of the 5 slices
    fp = d["fp64"]
    assert all(m["found"] == m["exact"] for m in fp["measured_slices"])
    assert fp["mixed_speedup_vs_projected_fp64"] > 0
