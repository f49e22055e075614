"""pEVSL_SetBatchWave (batched slices solved in waves) and the out-of-memory
return of pEVSL_ChebLanNrBatch.

* waves of 2 and 3 slices give every slice's exact count (closed form,
  TESTS/COMMON/lapl.c:139-169) and the one-batch eigenvalues, with outputs
  in the caller's slice order;
* a batch whose Lanczos bases do not fit in device memory returns
  PEVSL_ERR_DEVICE_OOM with empty outputs instead of ending the process
  (run in a child process that first fills the device memory);
* filter groups replayed from hipGraphs (pEVSL_SetGraphFilter) give the
  eager solve bit for bit.
"""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from pevsl_amd import capi, driver

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def banner(msg):
    print("\n[gpu-test] %s" % msg, flush=True)


@pytest.mark.parametrize("prec", [capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP32])
def test_waves_equal_one_batch(lib, prec):
    banner("ChebLanNrBatch in waves vs one batch, 14^3, 5 slices, prec %d" % prec)
    g = 14
    ia, ja, a = driver.laplacian3d(g, g, g)
    M = driver.Matrix(lib, ia, ja, a)
    edges = np.linspace(0.5, 1.5, 6)
    # caller's order is not the degree order: waves sort by degree internally
    intvs = [[edges[i], edges[i + 1], 0.0, 12.0] for i in (3, 0, 4, 1, 2)]
    pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
    vinit = driver.rand_vinit(lib, M.n)
    runs = {}
    try:
        for wave in (0, 2, 3):
            inst = driver.Instance(lib, M, prec=prec, basis_prec=prec)
            try:
                assert lib.pEVSL_SetBatchWave(inst.p, wave) == 0
                rc, outs = driver.cheblan_batch(inst, intvs, [400] * 5, 1e-8, vinit, pols)
                assert rc == 0
                runs[wave] = outs
            finally:
                inst.close()
        inst = driver.Instance(lib, M)
        try:
            assert lib.pEVSL_SetBatchWave(inst.p, -1) == -1
        finally:
            inst.close()
    finally:
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
        M.free()
    for s, iv in enumerate(intvs):
        ex = orc.exact_eig_lap3(g, g, g, iv[0], iv[1])
        for wave in (0, 2, 3):
            o = runs[wave][s]
            assert o["info"] is not None and o["info"]["its"] > 0
            lam = np.sort(o["lam"])
            assert len(lam) == len(ex) > 0, (wave, s, len(lam), len(ex))
            np.testing.assert_allclose(lam, ex, rtol=1e-10)
            np.testing.assert_allclose(lam, np.sort(runs[0][s]["lam"]), rtol=1e-12)


CHILD = r"""
import ctypes as C, json, sys
sys.path.insert(0, %(root)r)
from pevsl_amd import capi, driver
lib = capi.load()
hip = C.CDLL("libamdhip64.so")
g = 40
ia, ja, a = driver.laplacian3d(g, g, g)
M = driver.Matrix(lib, ia, ja, a)
intvs = [[0.6, 0.8, 0.0, 12.0], [0.8, 1.0, 0.0, 12.0], [1.0, 1.2, 0.0, 12.0]]
pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
vinit = driver.rand_vinit(lib, M.n)
inst = driver.Instance(lib, M)
inst.stats_reset()
# one basis chunk: 512 columns of n = 64000 doubles = 262 MB.  Fill the
# device until about 400 MB are free, so that the first slice's chunk fits
# and the second's does not.
free, total = C.c_size_t(), C.c_size_t()
held = []
for piece in (1 << 30, 1 << 26):
    while True:
        assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
        if free.value < piece + (400 << 20):
            break
        p = C.c_void_p()
        if hip.hipMalloc(C.byref(p), C.c_size_t(piece)) != 0:
            break
        held.append(p)
hip.hipMemGetInfo(C.byref(free), C.byref(total))
rc, outs = driver.cheblan_batch(inst, intvs, [600] * 3, 1e-8, vinit, pols)
for p in held:
    hip.hipFree(p)
# the instance is usable again once memory is back (286 eigenvalues: the
# reference's step cap 5 * (nev + 2))
rc2, outs2 = driver.cheblan_batch(inst, intvs[:1], [1500], 1e-8, vinit, pols[:1])
print(json.dumps({"free_mb": free.value / 2**20, "rc": rc,
                  "nev": [len(o["lam"]) for o in outs], "rc2": rc2,
                  "nev2": len(outs2[0]["lam"])}), flush=True)
inst.close()
for p in pols:
    lib.pEVSL_FreePol(C.byref(p))
M.free()
"""


def test_batch_out_of_device_memory_returns_error(lib):
    banner("ChebLanNrBatch with the device memory nearly full: PEVSL_ERR_DEVICE_OOM")
    out = subprocess.run([sys.executable, "-c", CHILD % {"root": ROOT}], capture_output=True,
                         text=True, timeout=240)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0, out.stderr[-2000:]
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["free_mb"] < 520
    assert r["rc"] == capi.PEVSL_ERR_DEVICE_OOM
    assert r["nev"] == [0, 0, 0]
    assert "does not fit in device memory" in out.stderr
    assert r["rc2"] == 0
    assert r["nev2"] == len(orc.exact_eig_lap3(40, 40, 40, 0.6, 0.8)) > 0


@pytest.mark.parametrize("prec,kind", [(capi.PEVSL_PREC_FP32, 2), (capi.PEVSL_PREC_FP32, 5),
                                       (capi.PEVSL_PREC_FP64, 1)])
def test_graph_filter_equals_eager(lib, prec, kind):
    """pEVSL_SetGraphFilter: every step replays each filter group from a
    hipGraph captured on the group's first step, with the solver's own eager
    kernels and synchronizations between the replays; groups change as
    slices finish, which captures new graphs.  The solve must equal the
    eager one bit for bit (eigenvalues, their order, step counts)."""
    banner("batched solve, filter groups from hipGraphs vs eager, prec %d kernel %d"
           % (prec, kind))
    g = 16
    ia, ja, a = driver.laplacian3d(g, g, g)
    M = driver.Matrix(lib, ia, ja, a)
    edges = np.linspace(0.4, 1.6, 7)
    intvs = [[edges[i], edges[i + 1], 0.0, 12.0] for i in range(6)]
    pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
    vinit = driver.rand_vinit(lib, M.n)
    runs = {}
    try:
        for graphs in (0, 1):
            inst = driver.Instance(lib, M, prec=prec, basis_prec=prec, block_rowkernel=kind)
            try:
                assert lib.pEVSL_SetGraphFilter(inst.p, graphs) == 0
                rc, outs = driver.cheblan_batch(inst, intvs, [400] * 6, 1e-8, vinit, pols)
                assert rc == 0
                runs[graphs] = (outs, inst.kernel_stats(), inst.kernel_times())
            finally:
                inst.close()
        inst = driver.Instance(lib, M)
        try:
            assert lib.pEVSL_SetGraphFilter(inst.p, 2) == -1
        finally:
            inst.close()
    finally:
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
        M.free()
    (eager, ks0, kt0), (graph, ks1, kt1) = runs[0], runs[1]
    for iv, o0, o1 in zip(intvs, eager, graph):
        assert o0["info"]["its"] == o1["info"]["its"]
        np.testing.assert_array_equal(o1["lam"], o0["lam"])
        ex = orc.exact_eig_lap3(g, g, g, iv[0], iv[1])
        assert len(o1["lam"]) == len(ex) > 0
    # the replays are counted like the eager launches
    assert ks1["filter_steps"] == ks0["filter_steps"]
    assert ks1["filter_launches"] == ks0["filter_launches"]
    assert abs(ks1["filter_bytes"] - ks0["filter_bytes"]) <= 1e-9 * ks0["filter_bytes"]
    blk = [k_ for k_ in kt0 if "blk" in k_ or "clen" in k_]
    assert blk and all(k_ in kt1 for k_ in blk)
    for k_ in blk:
        assert kt1[k_]["launches"] == kt0[k_]["launches"]
