"""Host-side checks of bench.py that need no GPU: the multi-GPU refusal and
the CPU-baseline leg (the oracle timed on a bounded sample and projected to
a full slice)."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "--gpus 1" in r.stderr
    assert r.stdout.strip() == ""


def test_cpu_baseline_projection():
    """20^3 slice [0.6, 0.7]: the sample runs, reports the thread count and
    projects T = its * t_step + t_cgs(512)/512 * its^2/2."""
    its, found = 400, 20
    cb = bench.cpu_baseline(20, 0.6, 0.7, 22, its, found)
    assert cb["kind"] == "port" and cb["unit"] == "eigenpairs/s"
    assert cb["cores"] >= 1
    T = its * cb["t_step_s"] + cb["t_cgs512_s"] / 512.0 * its * its / 2.0
    assert abs(cb["projected_slice_s"] - T) <= 1e-9 * T
    assert abs(cb["value"] - found / T) <= 1e-9 * cb["value"]
    assert "%d OpenMP threads" % cb["cores"] in cb["sample"]
