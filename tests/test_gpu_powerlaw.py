"""Irregular (power-law) matrices against the oracle's eigenvalues, at sizes
no dense solve reaches: the irregular-sparsity recipe at n = 200,000 (8
slices of the middle 5% of the spectrum) and at the workload's own size
n = 2,000,000 (middle slices of 16; the fixture holds the slices its
generator finished, about 1.6 hours each on 8 cores), from
tests/golden/make_powerlaw_golden.py (oracle ChebLanNr, FP64, same bounds,
slices and start vector).

* FP64, one slice per pEVSL_ChebLanNr call: the FindPol degree, the count
  and the eigenvalues of every slice equal the oracle's (to 1e-12 of the
  spectral radius);
* mixed (FP32 filter, FP64 basis), all the fixture's slices in one batch
  (the nnz-balanced block kernel for two or more slices): the same counts,
  eigenvalues to 1e-10 of the spectral radius, residuals below the stated
  1e-6 * rho.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from pevsl_amd import capi, driver

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = [n for n in (200000, 2000000)
         if os.path.exists(os.path.join(GOLD, "powerlaw_n%d.json" % n))]


def banner(msg):
    print("\n[gpu-test] %s" % msg, flush=True)


@pytest.fixture(scope="module", params=SIZES)
def case(lib, request):
    n = request.param
    with open(os.path.join(GOLD, "powerlaw_n%d.json" % n)) as f:
        gold = json.load(f)
    ia, ja, a = driver.powerlaw_sym(n, seed=4)
    assert int(ia[-1]) == gold["nnz"]
    M = driver.Matrix(lib, ia, ja, a)
    yield M, gold
    M.free()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_powerlaw_slices_match_oracle(lib, case, mode):
    M, gold = case
    banner("power-law n=%d, %d slices vs oracle, %s" % (gold["n"], len(gold["slices"]), mode))
    rho = max(abs(gold["lmin"]), abs(gold["lmax"]))
    intvs = [[s["intv"][0], s["intv"][1], gold["lmin"], gold["lmax"]] for s in gold["slices"]]
    pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
    try:
        assert [p_.deg for p_ in pols] == [s["deg"] for s in gold["slices"]]
        lib.pEVSL_RandSeed(0)
        vinit = driver.rand_vinit(lib, gold["n"])
        if mode == "fp64":
            inst = driver.Instance(lib, M)
            try:
                outs = []
                for iv, p_ in zip(intvs, pols):
                    rc, o = driver.cheblan_batch(inst, [iv], [300], 1e-8, vinit, [p_])
                    assert rc == 0
                    outs += o
            finally:
                inst.close()
            tol_lam, tol_res = 1e-12 * rho, 1e-8
        else:
            inst = driver.Instance(lib, M, prec=capi.PEVSL_PREC_FP32,
                                   basis_prec=capi.PEVSL_PREC_FP64)
            try:
                rc, outs = driver.cheblan_batch(inst, intvs, [300] * len(intvs), 1e-8, vinit,
                                                pols)
                assert rc == 0
                kt = inst.kernel_times()
            finally:
                inst.close()
            # a batch of two or more slices runs the block kernel (nnz-balanced
            # on these rows; three-term above degree 400, where Clenshaw sums
            # lose accuracy); one slice runs the one-column CSR-stream kernel
            if len(intvs) == 1:
                want = "k_cheb_stream<float>"
            elif max(p_.deg for p_ in pols) > 400:
                want = "k_blk_bal<float,%d,EpiCheb>" % len(intvs)
            else:
                want = "k_blk_bal<float,"
            assert any(k.startswith(want) for k in kt), kt.keys()
            tol_lam, tol_res = 1e-10 * rho, 1e-6 * rho
    finally:
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
    total = 0
    for s, o in zip(gold["slices"], outs):
        lam = np.sort(o["lam"])
        ref = np.asarray(s["lam"])
        print("slice %d: %d eigenvalues (oracle %d), %d steps (oracle %d)" % (
            s["slice"], len(lam), len(ref), o["info"]["its"], s["its"]), flush=True)
        assert len(lam) == len(ref), (s["slice"], len(lam), len(ref))
        if len(ref):
            assert np.max(np.abs(lam - ref)) <= tol_lam
            assert np.max(o["res"]) < tol_res
        total += len(ref)
    assert total > 0
