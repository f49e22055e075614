"""GPU tests at the full 100^3 sizes and of the boundary pieces around them.

* 100^3 FP64, KPM (Mdeg 300, nvec 64) slicing in 10 slices: the lowest- and
  highest-degree slices solved one at a time by pEVSL_ChebLanNr; exact
  counts (closed form, TESTS/COMMON/lapl.c:139-169), eigenvalues to 1e-10,
  residuals < 1e-8, and the pevsl_Stat counters of SRC/stats.c:13-103 filled.
* 100^3 mixed, learned-slicer slicing in 20 slices: two slices batched, with
  an FP64 and with an FP32 Lanczos basis, the second after pEVSL_Autotune.
* batched slices with different spectrum frames (per-slice cc/dd) equal the
  one-slice solves.
* pEVSL_Kpmdos over a host-callback operator equals the oracle's moments.
* pEVSL_SpslicerKpm (integral on the GPU) follows the closed-form cut rule.
* the C driver pevsl_amd/lap_plan (TESTS/Lap/LapPLanN.c flow) on 20^3.
"""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from pevsl_amd import capi, driver
from pevsl_amd import slicer as slc

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_slicer_pins import KPM_CASES, MU0, kpm_cuts_rule  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = 100
FRAME = (0.0, 12.0)


def banner(msg):
    print("\n[gpu-test] %s" % msg, flush=True)


@pytest.fixture(scope="module")
def lap100(lib):
    ia, ja, a = driver.laplacian3d(G, G, G)
    M = driver.Matrix(lib, ia, ja, a)
    yield M
    M.free()


def check_slice(o, a, b, tol_res):
    ex = driver.exact_eig_lap3(G, G, G, a, b)
    lam = np.sort(o["lam"])
    assert len(lam) == len(ex), (a, b, len(lam), len(ex))
    assert np.max(np.abs(lam - ex) / ex) <= 1e-10
    assert np.max(o["res"]) < tol_res
    return len(lam)


def test_c2_fp64_extreme_degree_slices(lib, lap100):
    banner("C2: 100^3 FP64, KPM 10 slices, lowest- and highest-degree slices")
    inst = driver.Instance(lib, lap100)
    try:
        intv = [0.6, 1.2, *FRAME]
        rc, mu, ecnt = inst.kpmdos(300, 1, 64, intv)
        assert rc == 0
        assert abs(ecnt - 15891) < 0.05 * 15891
        rc, sli = driver.spslicer_kpm(lib, mu, intv, 10, int(10 * ecnt))
        assert rc == 0
        pols = [driver.find_pol(lib, [sli[s], sli[s + 1], *FRAME])[0] for s in range(10)]
        degs = [p.deg for p in pols]
        ev_int = int(1 + ecnt / 10.0)
        mlan = min(max(5 * (ev_int + 2), 300), lap100.n)
        vinit = driver.rand_vinit(lib, lap100.n)
        for s in (int(np.argmin(degs)), int(np.argmax(degs))):
            banner("C2 slice %d [%.6f, %.6f] degree %d" % (s, sli[s], sli[s + 1], degs[s]))
            inst.stats_reset()
            rc, o = driver.cheblan_batch(inst, [[sli[s], sli[s + 1], *FRAME]], [mlan], 1e-8,
                                         vinit, [pols[s]])
            assert rc == 0
            nev = check_slice(o[0], sli[s], sli[s + 1], 1e-8)
            info, st, kt = o[0]["info"], inst.stats(), inst.kernel_times()
            print("slice %d: %d eigenvalues, its %d, %.1f s" % (s, nev, info["its"],
                                                                 info["t_total"]), flush=True)
            # pevsl_Stat as SRC/stats.c reports it
            assert st["n_mvA"] == info["nmatvec"] > degs[s] * info["its"]
            assert st["n_polAv"] == 1 + info["its"] + info["nbreak"]
            for f in ("t_polAv", "t_mvA", "t_reorth", "t_eig", "t_ritz", "t_iter"):
                assert st[f] > 0, f
            assert st["t_mvA"] <= st["t_polAv"] * 1.001
            assert info["tol_accept"] == 1e-8
            # the fused filter kernel ran every degree step
            assert kt["k_cheb_stream<double>"]["launches"] == \
                degs[s] * (1 + info["its"] + info["nbreak"])
        for p in pols:
            lib.pEVSL_FreePol(C.byref(p))
    finally:
        inst.close()


@pytest.fixture(scope="module")
def c3_slices(lib, lap100):
    """The learned slicer's 20 slices of [0.6, 1.2], judged by full KPM."""
    inst = driver.Instance(lib, lap100, prec=capi.PEVSL_PREC_FP32)
    net = slc.Mlp.load(slc.DEFAULT_WEIGHTS, lib)
    try:
        xintv = np.array([0.6, 1.2, *FRAME])
        sli, ecnt_mlp = slc.spslicer_mlp(inst, net, xintv, 20)
        rc, mu, ecnt = inst.kpmdos(300, 1, 64, xintv)
        assert rc == 0
        cnt, imb = slc.judge(lib, mu, xintv, lap100.n, sli)
    finally:
        net.free()
        inst.close()
    return sli, ecnt, imb


@pytest.mark.parametrize("basis", ["fp64", "fp32-autotuned"])
def test_c3_mixed_batched_two_slices(lib, lap100, c3_slices, basis):
    sli, ecnt, imb = c3_slices
    assert imb <= 0.15
    pols = [driver.find_pol(lib, [sli[s], sli[s + 1], *FRAME])[0] for s in range(20)]
    degs = [p.deg for p in pols]
    pick = sorted(range(20), key=lambda s: degs[s])[:2]
    banner("C3 mixed, %s basis: slices %s (degrees %s) batched" % (
        basis, pick, [degs[s] for s in pick]))
    bp = capi.PEVSL_PREC_FP64 if basis == "fp64" else capi.PEVSL_PREC_FP32
    inst = driver.Instance(lib, lap100, prec=capi.PEVSL_PREC_FP32, basis_prec=bp)
    try:
        if basis != "fp64":
            tr = capi.TuneResult()
            assert lib.pEVSL_Autotune(inst.p, 10, C.byref(tr)) == 0
            assert tr.block_width in (2, 4, 8, 16, 32, 64)
            assert tr.block_rowkernel in (0, 1, 2, 3, 4, 5)
            assert all(tr.t_blockseg[j] > 0 for j in range(1, 6))  # widths 4..64
            i = tr.block_width.bit_length() - 2
            t_sel = tr.t_blockseg[i] if tr.block_rowkernel == 5 else tr.t_blockclen[i]
            assert t_sel > 0 and tr.bytes_clen[min(i, 4)] > 0
            print("autotune: width %d kernel %d, %.2f us per step" % (
                tr.block_width, tr.block_rowkernel, 1e6 * t_sel), flush=True)
        ev_int = int(1 + ecnt / 20.0)
        mlan = min(max(5 * (ev_int + 2), 300), lap100.n)
        vinit = driver.rand_vinit(lib, lap100.n)
        rc, outs = driver.cheblan_batch(inst, [[sli[s], sli[s + 1], *FRAME] for s in pick],
                                        [mlan] * 2, 1e-8, vinit, [pols[s] for s in pick])
        assert rc == 0
        for s, o in zip(pick, outs):
            # stated tolerance of the FP32-filter path: 1e-6 * rho, rho = 12
            assert abs(o["info"]["tol_accept"] - 1.2e-5) < 1e-12
            nev = check_slice(o, sli[s], sli[s + 1], o["info"]["tol_accept"])
            print("slice %d: %d eigenvalues, its %d" % (s, nev, o["info"]["its"]), flush=True)
        kt = inst.kernel_times()
        assert any(k.startswith("k_clen_blkrow<float,") or k.startswith("k_clen_seg<float,")
                   for k in kt), kt
    finally:
        inst.close()
        for p in pols:
            lib.pEVSL_FreePol(C.byref(p))


@pytest.mark.parametrize("prec", [capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP32])
def test_batch_with_different_frames_equals_single_solves(lib, prec):
    """Slices whose filters map different spectrum frames (per-slice cc, dd:
    FindPol uses each slice's own [lmin, lmax], SRC/chebpol.c:386-389) are
    filtered with their own maps inside one batched call."""
    banner("batched slices with different frames, prec %d" % prec)
    g = 12
    ia, ja, a = driver.laplacian3d(g, g, g)
    M = driver.Matrix(lib, ia, ja, a)
    intvs = [[0.6, 1.2, 0.0, 12.0], [1.2, 1.9, 0.0, 12.0], [1.9, 2.6, -1.0, 12.5]]
    pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
    assert (pols[0].cc, pols[0].dd) == (pols[1].cc, pols[1].dd) != (pols[2].cc, pols[2].dd)
    vinit = driver.rand_vinit(lib, M.n)
    try:
        inst = driver.Instance(lib, M, prec=prec, block_width=4)
        try:
            rc, batched = driver.cheblan_batch(inst, intvs, [400] * 3, 1e-8, vinit, pols)
            assert rc == 0
            singles = []
            for iv, p in zip(intvs, pols):
                rc, o = driver.cheblan_batch(inst, [iv], [400], 1e-8, vinit, [p])
                assert rc == 0
                singles.append(o[0])
        finally:
            inst.close()
        for iv, ob, os_ in zip(intvs, batched, singles):
            ex = orc.exact_eig_lap3(g, g, g, iv[0], iv[1])
            assert len(ob["lam"]) == len(os_["lam"]) == len(ex)
            if prec == capi.PEVSL_PREC_FP64:
                np.testing.assert_allclose(np.sort(ob["lam"]), np.sort(os_["lam"]), rtol=1e-12)
            np.testing.assert_allclose(np.sort(ob["lam"]), ex, rtol=1e-10)
            np.testing.assert_allclose(np.sort(os_["lam"]), ex, rtol=1e-10)
    finally:
        for p in pols:
            lib.pEVSL_FreePol(C.byref(p))
        M.free()


def test_kpmdos_host_callback_matches_oracle(lib):
    banner("Kpmdos through pEVSL_SetAMatvec, 10^3")
    ia, ja, a = driver.laplacian3d(10, 10, 10)
    A = orc.Csr(ia, ja, a)
    n = A.n
    CB = C.CFUNCTYPE(None, capi.c_double_p, capi.c_double_p, C.c_void_p)
    calls = [0]

    def mv(x, y, data):
        calls[0] += 1
        xv = np.ctypeslib.as_array(x, shape=(n,))
        np.ctypeslib.as_array(y, shape=(n,))[:] = A.matvec(xv.copy())
    cb = CB(mv)
    intv = [0.6, 1.2, 0.0, 12.0]
    inst = driver.Instance(lib)
    try:
        lib.pEVSL_SetProbSizes(inst.p, n, n, 0)
        lib.pEVSL_SetAMatvec(inst.p, C.cast(cb, C.c_void_p), None)
        rc, mu, ecnt = inst.kpmdos(50, 1, 8, intv)
        assert rc == 0
        assert inst.stats()["n_mvA"] == calls[0] == 50 * 8
    finally:
        inst.close()
    orc.srand(0)
    orc_rc, omu, oecnt = orc.kpmdos(A, 50, 1, 8, intv)
    assert orc_rc == 0
    np.testing.assert_allclose(mu, omu, rtol=0, atol=1e-12 * np.max(np.abs(omu)))
    assert abs(ecnt - oecnt) <= 1e-10 * max(1.0, abs(oecnt))


@pytest.mark.parametrize("intv,n_int,npts", KPM_CASES)
def test_library_spslicer_kpm_closed_form(lib, intv, n_int, npts):
    rc, sli = driver.spslicer_kpm(lib, MU0, intv, n_int, npts)
    rc_ref, sli_ref = kpm_cuts_rule(intv, n_int, npts)
    assert rc == rc_ref == 0
    np.testing.assert_allclose(sli, sli_ref, rtol=0, atol=1e-13 * (intv[3] - intv[2]))


def test_library_spslicer_kpm_nonpositive_integral(lib):
    for intv in ([0.6, 1.2, 0.0, 12.0], [1.0, 5.0, 0.0, 12.0]):
        for scale in (0.0, -1.0):
            rc, sli = driver.spslicer_kpm(lib, scale * MU0[:11], intv, 4, 9)
            orc_rc, osli = orc.spslicer_kpm(scale * MU0[:11], intv, 4, 9)
            rc_ref, sli_ref = kpm_cuts_rule(intv, 4, 9, scale=scale)
            assert rc == orc_rc == rc_ref
            np.testing.assert_array_equal(sli, sli_ref)
            np.testing.assert_array_equal(osli, sli_ref)


@pytest.mark.parametrize("batch", [0, 1])
def test_lap_plan_c_driver(batch):
    """pevsl_amd/lap_plan: the LapPLanN.c call sequence through the C ABI on
    20^3, [0.6, 1.2] in 5 slices; every slice's count is the exact one."""
    banner("lap_plan C driver, batch %d" % batch)
    exe = os.path.join(ROOT, "pevsl_amd", "lap_plan")
    out = subprocess.run([exe, "-batch", str(batch)], capture_output=True, text=True, timeout=240)
    print(out.stdout[-2000:], flush=True)
    assert out.returncode == 0, out.stderr
    m = re.search(r"total (\d+) eigenvalues, (\d+) slice", out.stdout)
    assert int(m.group(1)) == len(orc.exact_eig_lap3(20, 20, 20, 0.6, 1.2)) == 118
    assert int(m.group(2)) == 0
