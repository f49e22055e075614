"""Host-side checks of bench.py that need no GPU: the multi-GPU refusal, the
CPU-baseline sample on the oracle, and the FP64 projection model."""
import os
import subprocess
import sys

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "--gpus 1" in r.stderr
    assert r.stdout.strip() == ""


def test_cpu_sample_units():
    """20^3 slice [0.6, 0.7]: the oracle sample reports a positive time per
    degree step at the slice's FindPol degree, a Gram-Schmidt time per column
    and the thread count."""
    smp = bench.cpu_sample(20, 0.6, 0.7, 0, 22, steps=2, reorthk=128)
    assert smp["deg"] > 1 and smp["t_degree_step"] > 0 and smp["t_cgs_per_col"] > 0
    assert smp["threads"] >= 1 and smp["steps"] == 2 and smp["reorthk"] == 128


def test_cpu_model_counts_filter_and_gram_schmidt():
    """(steps + 1) filter applications at the degree, steps^2 / 2 Gram-Schmidt
    columns, summed over slices."""
    smp = {"t_degree_step": 2e-3, "t_cgs_per_col": 1e-4}
    t = bench.cpu_model(smp, [(100, 10), (200, 20)])
    assert abs(t - (11 * 100 * 2e-3 + 1e-4 * 50 + 21 * 200 * 2e-3 + 1e-4 * 200)) < 1e-12


def test_fp64_model_reproduces_its_inputs():
    """The per-slice model T = cf deg (its+1) + cr its^2/2 + co its, fitted on
    two slices, returns their own total when projected on them."""
    meas = [{"deg": 150, "its": 2000, "t_filter": 150 * 2001 * 2.1e-5, "t_reorth": 2000 ** 2 * 6e-7,
             "seconds": 150 * 2001 * 2.1e-5 + 2000 ** 2 * 6e-7 + 0.3},
            {"deg": 450, "its": 3000, "t_filter": 450 * 3001 * 2.1e-5, "t_reorth": 3000 ** 2 * 6e-7,
             "seconds": 450 * 3001 * 2.1e-5 + 3000 ** 2 * 6e-7 + 0.45}]
    tot, coef = bench.fp64_model(meas, [(150, 2000), (450, 3000)])
    assert abs(coef["cf_s_per_degree_step"] - 2.1e-5) < 1e-12
    assert abs(coef["cr_s_per_step2"] - 1.2e-6) < 1e-15
    assert abs(tot - sum(m["seconds"] for m in meas)) < 1e-9 * tot
    assert np.isfinite(tot)


def test_roofline_picks_the_kernel_with_most_time():
    kt = {"k_a<float,8>": {"launches": 10, "seconds": 1e-3, "bytes": 1.875e9},
          "k_b<float,4>": {"launches": 10, "seconds": 5e-4, "bytes": 1e9}}
    r = bench.roofline_of(kt, {"k_a<float,8>": 2e8})
    assert r["kernel"] == "k_a<float,8>" and r["traffic"] == 2e8
    assert abs(r["achieved"] - 1875.0) < 1e-6 and abs(r["frac"] - 1875.0 / 8000.0) < 1e-4


def test_first_full_wave_follows_the_library_order():
    """The bench times the first full wave: degree order, remainder first
    (as pEVSL_ChebLanNrBatch schedules waves)."""
    degs = [300, 226, 456, 240, 410, 253, 266, 279, 292, 304, 316, 328, 340, 352, 365, 376,
            388, 399, 422, 433]
    w = bench.first_full_wave(degs, 8)
    order = sorted(range(20), key=lambda s: degs[s])
    assert w == order[4:12]
    assert bench.first_full_wave(degs, 0) == list(range(20))
    assert bench.first_full_wave(degs, 20) == list(range(20))
    assert bench.first_full_wave(degs, 5) == order[0:5]   # no remainder
    assert bench.first_full_wave([3, 1, 2], 2) == [2, 0]  # remainder 1: slice 1 goes first
