"""GPU parity tests: the HIP path through the C-ABI against the oracle on
the same inputs (same libc rand() stream), and against the closed-form
Laplacian spectrum.  Tolerances are written per test:
  * sparse products / filter with one lane per row: bitwise (same rounding
    sequence as the reference's scalar loop); with several lanes per row the
    row sum is a tree: relative 1e-14;
  * KPM moments: relative 1e-12 (reduction order of the dots differs);
  * eigenvalues (FP64 path): relative 1e-10 or better against the closed
    form, and the same count per slice as the oracle's ChebLanNr.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from pevsl_amd import capi, driver

pytestmark = pytest.mark.gpu


def banner(msg):
    print("\n[gpu case] " + msg, flush=True)


@pytest.fixture(scope="module")
def lap20(lib):
    ia, ja, a = driver.laplacian3d(20, 20, 20)
    M = driver.Matrix(lib, ia, ja, a)
    yield M, orc.Csr(ia, ja, a)
    M.free()


def parvec(lib, x):
    v = capi.Parvec()
    lib.pEVSL_ParvecCreateShell(len(x), len(x), 0, 0, C.byref(v),
                                x.ctypes.data_as(capi.c_double_p))
    return v


@pytest.mark.parametrize("lanes", [0, 1, 2, 4, 8])
def test_device_matvec_sin_recipe(lib, lap20, lanes):
    banner("ParcsrMatvec 20^3 x=sin(i) lanes=%d" % lanes)
    M, A = lap20
    x = np.sin(np.arange(M.n, dtype=np.float64))
    if lanes == 0:
        y = M.matvec(x)  # pEVSL_ParcsrMatvec, host vectors
    else:
        inst = driver.Instance(lib, M, lanes=lanes)
        hip = C.CDLL("libamdhip64.so")
        d = C.c_void_p()
        nb = x.nbytes
        assert hip.hipMalloc(C.byref(d), C.c_size_t(2 * nb)) == 0
        try:
            dx, dy = d.value, d.value + nb
            assert hip.hipMemcpy(C.c_void_p(dx), x.ctypes.data_as(C.c_void_p),
                                 C.c_size_t(nb), 1) == 0  # host to device
            assert lib.pEVSL_DeviceMatvec(inst.p, C.c_void_p(dx), C.c_void_p(dy)) == 0
            y = np.empty_like(x)
            assert hip.hipMemcpy(y.ctypes.data_as(C.c_void_p), C.c_void_p(dy),
                                 C.c_size_t(nb), 2) == 0  # device to host
        finally:
            hip.hipFree(d)
            inst.close()
    ref = A.matvec(x)
    if lanes == 1:
        assert np.array_equal(y, ref)
    else:
        assert np.max(np.abs(y - ref)) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("lanes,prec", [(1, 0), (2, 0), (8, 0), (2, 1),
                                        (capi.PEVSL_SPMV_STREAM, 0),
                                        (capi.PEVSL_SPMV_STREAM, 1)])
def test_cheb_av_against_oracle(lib, lap20, lanes, prec):
    banner("ChebAv 20^3 lanes=%d prec=%d" % (lanes, prec))
    M, A = lap20
    iv = [0.87, 0.99, 0.0, 12.0]
    pol, rc = driver.find_pol(lib, iv)
    opol = orc.Pol(iv)
    assert rc == 0 and pol.deg == opol.deg
    inst = driver.Instance(lib, M, prec=prec, lanes=lanes)
    try:
        v = np.random.default_rng(0).uniform(-1, 1, M.n)
        y = np.zeros(M.n)
        rc = lib.pEVSL_ChebAv(inst.p, C.byref(pol), C.byref(parvec(lib, v)),
                              C.byref(parvec(lib, y)), None)
        assert rc == 0
    finally:
        inst.close()
        lib.pEVSL_FreePol(C.byref(pol))
    ref = orc.cheb_av(A, opol, v)
    scale = np.max(np.abs(ref))
    if prec == 1:
        # FP32 filter: error relative to the filtered vector's size
        assert np.max(np.abs(y - ref)) <= 1e-4 * scale
    elif lanes in (1, capi.PEVSL_SPMV_STREAM):
        # one sequential sum per row: same rounding as the reference loop
        assert np.array_equal(y, ref)
    else:
        assert np.max(np.abs(y - ref)) <= 1e-12 * scale


@pytest.mark.parametrize("prec", [0, 1])
def test_kpmdos_and_slicer_against_oracle(lib, lap20, prec):
    banner("Kpmdos 20^3 Mdeg=300 nvec=60 prec=%d + SpslicerKpm" % prec)
    M, A = lap20
    intv = [0.6, 1.2, 0.0, 12.0]
    inst = driver.Instance(lib, M, prec=prec)  # pEVSL_Start seeds srand(0)
    try:
        rc, mu, ecnt = inst.kpmdos(300, 1, 60, intv)
    finally:
        inst.close()
    assert rc == 0
    orc.srand(0)
    rc, omu, oecnt = orc.kpmdos(A, 300, 1, 60, intv)
    if prec == 0:
        np.testing.assert_allclose(mu, omu, rtol=1e-12, atol=1e-15)
        assert abs(ecnt - oecnt) < 1e-9
        rc, sli = driver.spslicer_kpm(lib, mu, intv, 5, int(10 * ecnt))
        orc_rc, osli = orc.spslicer_kpm(omu, intv, 5, int(10 * oecnt))
        assert rc == orc_rc == 0
        np.testing.assert_allclose(sli, osli, rtol=0, atol=1e-12)
    else:
        np.testing.assert_allclose(mu, omu, rtol=0, atol=2e-5)
        assert abs(ecnt - oecnt) < 0.05


def test_spslicer_device_integral_matches_oracle(lib):
    banner("SpslicerKpm random moments, 5000 points")
    rng = np.random.default_rng(3)
    mu = rng.normal(size=301) * np.exp(-np.arange(301) / 50.0)
    mu[0] = abs(mu[0]) + 1.0
    intv = [0.6, 1.2, 0.0, 12.0]
    rc, sli = driver.spslicer_kpm(lib, mu, intv, 7, 5000)
    orc_rc, osli = orc.spslicer_kpm(mu, intv, 7, 5000)
    assert rc == orc_rc
    np.testing.assert_allclose(sli, osli, atol=1e-12)


def test_cheblan_single_slice_matches_oracle(lib):
    banner("ChebLanNr 12^3 [0.6,1.2] FP64 vs oracle")
    nx = 12
    ia, ja, a = driver.laplacian3d(nx, nx, nx)
    M = driver.Matrix(lib, ia, ja, a)
    inst = driver.Instance(lib, M)
    intv = [0.6, 1.2, 0.0, 12.0]
    try:
        pol, rc = driver.find_pol(lib, intv)
        vinit = driver.rand_vinit(lib, M.n)
        rc, outs = driver.cheblan_batch(inst, [intv], [300], 1e-8, vinit, [pol],
                                        keep_vectors=True)
        lib.pEVSL_FreePol(C.byref(pol))
    finally:
        inst.close()
    assert rc == 0
    o = outs[0]
    A = orc.Csr(ia, ja, a)
    ref = orc.cheblan_nr(A, intv, 300, 1e-8, vinit, orc.Pol(intv))
    ex = orc.exact_eig_lap3(nx, nx, nx, 0.6, 1.2)
    assert len(o["lam"]) == len(ref["lam"]) == len(ex)
    np.testing.assert_allclose(np.sort(o["lam"]), ex, rtol=1e-10)
    np.testing.assert_allclose(np.sort(o["lam"]), np.sort(ref["lam"]), rtol=1e-12)
    assert o["info"]["its"] == ref["its"] and o["info"]["nbreak"] == ref["nbreak"]
    assert np.all(o["res"] < 1e-8)
    # returned Ritz vectors: unit norm, residual as reported
    for lam, y, r in zip(o["lam"], o["vecs"], o["res"]):
        assert abs(np.linalg.norm(y) - 1.0) < 1e-12
        assert abs(np.linalg.norm(A.matvec(y) - lam * y) - r) < 1e-12
    M.free()


@pytest.mark.parametrize("batch", [False, True])
def test_c1_end_to_end_counts(lib, batch):
    banner("C1 20^3 [0.6,1.2] 5 slices batch=%s" % batch)
    r = driver.lap_slicing_run(20, 20, 20, 0.6, 1.2, 5, batch=batch)
    d = orc.run_driver(["-nx", 20, "-ny", 20, "-nz", 20, "-a", 0.6, "-b", 1.2,
                        "-nslices", 5], timeout=120)
    assert abs(r["ecount"] - d["ecount"]) < 1e-9
    np.testing.assert_allclose(r["sli"], d["sli"], atol=1e-12)
    total = 0
    for s, ds in zip(r["slices"], d["slices"]):
        assert s["deg"] == ds["deg"]
        assert s["nev"] == s["nev_exact"] == ds["nev"]
        assert s["max_rel_err"] < 1e-10
        assert np.all(s["res"] < 1e-8)
        total += s["nev"]
    assert total == 118


def test_c1_readme_slices(lib):
    banner("README slices 20^3: counts 18/27/27/19/27, degrees 61..110")
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "readme_c1.json")))
    r = driver.lap_slicing_run(20, 20, 20, 0.6, 1.2, 5, batch=True,
                               sli_override=g["sli"])
    for s, gs in zip(r["slices"], g["slices"]):
        assert s["deg"] == gs["deg"]
        assert s["nev"] == gs["found"] == gs["exact"]


@pytest.mark.parametrize("rowk", [0, 1, 3])
def test_batch_equals_single_slices(lib, rowk):
    banner("batched vs one-by-one solve, 16^3, 4 slices, block kernel rowk=%d" % rowk)
    r1 = driver.lap_slicing_run(16, 16, 16, 0.4, 1.0, 4, batch=False)
    r2 = driver.lap_slicing_run(16, 16, 16, 0.4, 1.0, 4, batch=True, block_rowkernel=rowk)
    for a, b in zip(r1["slices"], r2["slices"]):
        assert a["nev"] == b["nev"] == a["nev_exact"]
        np.testing.assert_allclose(a["lam"], b["lam"], rtol=1e-12)
        # block kernels sum each row sequentially: same Lanczos run
        assert a["info"]["its"] == b["info"]["its"]


def test_lanczos_dos_against_oracle(lib):
    banner("LanDosG 10^3 nvec=12 msteps=40 vs oracle")
    ia, ja, a = driver.laplacian3d(10, 10, 10)
    M = driver.Matrix(lib, ia, ja, a)
    intv = [0.6, 1.2, 0.0, 12.0]
    inst = driver.Instance(lib, M)
    try:
        rc, xd, yd, neig = inst.landos(12, 40, 200, intv)
    finally:
        inst.close()
        M.free()
    assert rc == 0
    orc.srand(0)
    orc_rc, oxd, oyd, oneig = orc.landos(orc.Csr(ia, ja, a), 12, 40, 200, intv)
    assert np.array_equal(xd, oxd)
    np.testing.assert_allclose(yd, oyd, rtol=1e-8, atol=1e-10)
    assert abs(neig - oneig) < 1e-6 * max(1.0, abs(oneig))
    sli = driver.spslicer_lan(lib, xd, yd, 4)
    assert np.all(np.diff(sli) > 0)


def test_host_matvec_callback_path(lib):
    banner("SetAMatvec host callback, 8^3 slice")
    ia, ja, a = driver.laplacian3d(8, 8, 8)
    A = orc.Csr(ia, ja, a)
    n = A.n
    CB = C.CFUNCTYPE(None, capi.c_double_p, capi.c_double_p, C.c_void_p)

    def mv(x, y, data):
        xv = np.ctypeslib.as_array(x, shape=(n,))
        yv = np.ctypeslib.as_array(y, shape=(n,))
        yv[:] = A.matvec(xv.copy())
    cb = CB(mv)
    inst = driver.Instance(lib)
    try:
        lib.pEVSL_SetProbSizes(inst.p, n, n, 0)
        lib.pEVSL_SetAMatvec(inst.p, C.cast(cb, C.c_void_p), None)
        intv = [0.6, 1.8, 0.0, 12.0]
        pol, rc = driver.find_pol(lib, intv)
        vinit = driver.rand_vinit(lib, n)
        rc, outs = driver.cheblan_batch(inst, [intv], [200], 1e-8, vinit, [pol])
        lib.pEVSL_FreePol(C.byref(pol))
    finally:
        inst.close()
    assert rc == 0
    ex = orc.exact_eig_lap3(8, 8, 8, 0.6, 1.8)
    assert len(outs[0]["lam"]) == len(ex)
    np.testing.assert_allclose(np.sort(outs[0]["lam"]), ex, rtol=1e-10)


def test_invalid_interval_returns_nothing(lib):
    banner("ChebLanNr invalid interval")
    ia, ja, a = driver.laplacian3d(6, 6, 6)
    M = driver.Matrix(lib, ia, ja, a)
    inst = driver.Instance(lib, M)
    try:
        pol, _ = driver.find_pol(lib, [0.6, 1.2, 0, 12])
        vinit = driver.rand_vinit(lib, M.n)
        rc, outs = driver.cheblan_batch(inst, [[1.2, 0.6, 0, 12]], [100], 1e-8, vinit, [pol])
        assert rc == 0 and len(outs[0]["lam"]) == 0
        lib.pEVSL_FreePol(C.byref(pol))
    finally:
        inst.close()
        M.free()


def test_reorth_kernel_orthogonality(lib):
    banner("40^3 slice with ~90 eigenvalues: counts, orthogonality")
    r = driver.lap_slicing_run(40, 40, 40, 0.6, 1.2, 8, slices_only=[3], batch=True,
                               keep_vectors=True)
    s = r["slices"][0]
    assert s["nev"] == s["nev_exact"]
    assert s["max_rel_err"] < 1e-10
    Y = s["vecs"]
    G = Y @ Y.T
    assert np.max(np.abs(G - np.eye(len(Y)))) < 1e-8


@pytest.mark.parametrize("n,k,i_max,near", [(5000, 40, 2, 0.0), (5000, 100, 2, 1e-9),
                                             (3001, 33, 1, 1e-9), (4096, 0, 2, 0.0),
                                             (20000, 300, 3, 1e-12)])
def test_cgs_dgks_against_oracle(lib, n, k, i_max, near):
    """pEVSL_CGS_DGKS (the device reorthogonalization kernels of the Lanczos
    solve, CGS_DGKS of SRC/miscla.c:211-240) against the oracle's restatement
    on the same Q and v.  `near` > 0 puts v within that distance of span(Q),
    so that the first pass removes almost all of it and the DGKS test asks
    for another pass; widths k = 33, 100 and 300 leave partial 32-column
    groups in the GEMV-T kernel; k = 0 leaves v as it is."""
    banner("CGS_DGKS n=%d k=%d i_max=%d near=%g" % (n, k, i_max, near))
    rng = np.random.default_rng(n + k)
    Q = np.linalg.qr(rng.standard_normal((n, max(k, 1))))[0][:, :k] if k else np.zeros((n, 0))
    Q = np.asfortranarray(Q)
    r = rng.standard_normal(n)
    v = Q @ rng.standard_normal(k) + near * r if near > 0 else r
    inst = driver.Instance(lib)
    try:
        lib.pEVSL_SetProbSizes(inst.p, n, n, 0)
        Qp = capi.Parvecs()
        lib.pEVSL_ParvecsCreateShell(n, max(k, 1), n, n, 0, 0, C.byref(Qp),
                                     Q.ctypes.data_as(capi.c_double_p) if k else None)
        Qp.n_vecs = k
        vg = np.ascontiguousarray(v.copy())
        nrm = C.c_double()
        assert lib.pEVSL_CGS_DGKS(inst.p, k, i_max, C.byref(Qp) if k else None,
                                  C.byref(parvec(lib, vg)), C.byref(nrm), None) == 0
    finally:
        inst.close()
    vo, nrm_o = orc.cgs_dgks(Q, v, i_max)
    scale = np.linalg.norm(v)
    np.testing.assert_allclose(vg, vo, rtol=0, atol=1e-13 * scale)
    assert abs(nrm.value - nrm_o) <= 1e-12 * scale
    if k:
        # what the passes leave of span(Q): rounding of the last pass
        assert np.max(np.abs(Q.T @ vg)) <= 1e-13 * scale


@pytest.mark.parametrize("upper", [0, 1])
@pytest.mark.parametrize("n,m", [(1000, 7), (40000, 130), (100003, 65), (250000, 100),
                                 (1000000, 800)])
def test_gram_mfma_against_rocblas(lib, n, m, upper):
    """The library's FP64 MFMA Gram kernel (v_mfma_f64_16x16x4 tiles, row
    parts summed in a fixed order) against the rocBLAS strided-batched DGEMM
    path on C = X^T Y for random X, Y: partial 64 x 64 tiles (m = 7, 65, 100,
    130), row counts that leave short parts and partial 64-row steps.  The
    two summation orders differ in rounding only.  upper: the kernel forms
    only the tiles on and above the diagonal (1, 3, 3, 6, 91 of 1, 9, 4, 4,
    169 here) and returns the upper triangle mirrored; the lower triangle
    must mirror it bit for bit."""
    banner("Gram X^T Y, MFMA kernel vs rocBLAS, n=%d m=%d upper=%d" % (n, m, upper))
    inst = driver.Instance(lib)
    try:
        lib.pEVSL_SetProbSizes(inst.p, n, n, 0)
        tb, tm, diff = C.c_double(), C.c_double(), C.c_double()
        assert lib.pEVSL_BenchGram(inst.p, n, m, 3, upper, C.byref(tb), C.byref(tm),
                                   C.byref(diff)) == 0
    finally:
        inst.close()
    flops = 2.0 * n * m * m
    print("rocBLAS %.3f ms (%.1f TFLOP/s), MFMA kernel %.3f ms (%.1f TFLOP/s of the whole "
          "product), max diff %.2e" % (tb.value * 1e3, flops / tb.value / 1e12, tm.value * 1e3,
                                       flops / tm.value / 1e12, diff.value), flush=True)
    assert diff.value < 1e-12


@pytest.mark.parametrize("gram_kernel", [0, 1])
def test_mixed_c1_gram_kernels(lib, gram_kernel):
    """The mixed path's FP64 Rayleigh-Ritz stage with either Gram kernel
    (rocBLAS, the MFMA kernel): C1 batched, every slice exact, eigenvalues to
    1e-10, residuals below 1e-6 * rho."""
    banner("C1 mixed precision, Rayleigh-Ritz Gram products on kernel %d" % gram_kernel)
    r = driver.lap_slicing_run(20, 20, 20, 0.6, 1.2, 5, batch=True, prec=capi.PEVSL_PREC_FP32,
                               gram_kernel=gram_kernel)
    for s_ in r["slices"]:
        assert s_["nev"] == s_["nev_exact"]
        assert s_["max_rel_err"] < 1e-10
        assert np.all(s_["res"] < 1.2e-5)


@pytest.mark.parametrize("n", [1, 2, 7, 300, 2000])
def test_tridiag_multisection_against_lapack(lib, n):
    banner("GPU Sturm multisection n=%d" % n)
    import scipy.linalg as sla
    rng = np.random.default_rng(n)
    d = rng.uniform(-1, 1, n)
    e = rng.uniform(0.0, 1, max(n - 1, 0))
    if n > 10:
        e[5] = 0.0          # split matrix
        d[20:40] = 0.3      # a cluster
        e[20:39] = 1e-9
    ref = sla.eigh_tridiagonal(d, e, eigvals_only=True)
    lam = np.zeros(n)
    cnt = C.c_int()
    tr = C.c_double()
    rc = lib.pEVSL_SymmTridEigvals(n, d.ctypes.data_as(capi.c_double_p),
                                   e.ctypes.data_as(capi.c_double_p), -np.inf,
                                   lam.ctypes.data_as(capi.c_double_p), C.byref(cnt),
                                   C.byref(tr))
    assert rc == 0 and cnt.value == n
    # bisection and LAPACK's MRRR each have O(sqrt(n) eps |T|) absolute error
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(lam - ref)) <= 64 * np.finfo(float).eps * scale * np.sqrt(n)
    # restricted trace above a threshold (the ChebLanNr stopping test)
    thr = 0.25
    rc = lib.pEVSL_SymmTridEigvals(n, d.ctypes.data_as(capi.c_double_p),
                                   e.ctypes.data_as(capi.c_double_p), thr,
                                   lam.ctypes.data_as(capi.c_double_p), C.byref(cnt),
                                   C.byref(tr))
    sel = ref[ref >= thr]
    assert cnt.value == len(sel)
    assert abs(tr.value - sel.sum()) <= 64 * np.finfo(float).eps * scale * np.sqrt(n) * max(1, len(sel))


@pytest.mark.parametrize("batch,rowk", [(False, -1), (True, 1), (True, 2), (True, 3), (True, 4),
                                        (True, 5)])
def test_mixed_precision_c1(lib, batch, rowk):
    banner("C1 mixed precision (FP32 filter + FP64 Rayleigh-Ritz) batch=%s block kernel %d"
           % (batch, rowk))
    # stated mixed-path tolerance: residual < 1e-6 * rho (rho = 12 here),
    # eigenvalue error < 1e-10;
    # block kernel 1 = three-term recurrence, 2 = Clenshaw form, 3 / 4 their
    # nnz-balanced forms, 5 = segmented Clenshaw form (all 5 slices in one
    # block of two 4-column pieces)
    r = driver.lap_slicing_run(20, 20, 20, 0.6, 1.2, 5, batch=batch, prec=capi.PEVSL_PREC_FP32,
                               block_rowkernel=rowk)
    for s in r["slices"]:
        assert s["nev"] == s["nev_exact"]
        assert s["max_rel_err"] < 1e-10
        assert np.all(s["res"] < 1.2e-5)


@pytest.mark.parametrize("basis", [capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP32])
def test_segmented_clenshaw_equals_row_kernel(lib, basis):
    """Block kernel 5 (one thread per (row, 4-column piece), every piece
    active from its own top degree) against kernel 2 (one thread per row,
    all columns of the block every degree) on 6 slices of uneven degree in
    one block of width 8 (pieces of 4 + 2 columns, 2 padding): each column
    goes through the same operations in the same order, so the two solves
    agree bit for bit (eigenvalues, their order, step counts)."""
    banner("segmented vs row Clenshaw kernel, 16^3, 6 slices, basis %d" % basis)
    g = 16
    ia, ja, a = driver.laplacian3d(g, g, g)
    M = driver.Matrix(lib, ia, ja, a)
    edges = np.linspace(0.4, 1.6, 7)
    intvs = [[edges[i], edges[i + 1], 0.0, 12.0] for i in range(6)]
    pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
    degs = [p_.deg for p_ in pols]
    assert len(set(degs)) == 6
    vinit = driver.rand_vinit(lib, M.n)
    runs = {}
    try:
        for kind in (2, 5):
            inst = driver.Instance(lib, M, prec=capi.PEVSL_PREC_FP32, basis_prec=basis,
                                   block_rowkernel=kind)
            try:
                rc, outs = driver.cheblan_batch(inst, intvs, [400] * 6, 1e-8, vinit, pols)
                assert rc == 0
                runs[kind] = (outs, inst.kernel_times(), inst.kernel_stats())
            finally:
                inst.close()
    finally:
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
        M.free()
    assert "k_clen_seg<float,8>" in runs[5][1], runs[5][1].keys()
    assert "k_clen_blkrow<float,8>" in runs[2][1], runs[2][1].keys()
    # both count every column's products once per degree
    assert runs[5][2]["filter_steps"] == runs[2][2]["filter_steps"]
    for iv, o2, o5 in zip(intvs, runs[2][0], runs[5][0]):
        assert o2["info"]["its"] == o5["info"]["its"]
        np.testing.assert_array_equal(o5["lam"], o2["lam"])
        ex = orc.exact_eig_lap3(g, g, g, iv[0], iv[1])
        assert len(o5["lam"]) == len(ex) > 0
        np.testing.assert_allclose(np.sort(o5["lam"]), ex, rtol=1e-10)


@pytest.mark.parametrize("batch", [False, True])
def test_mixed_precision_fp32_basis_c1(lib, batch):
    banner("C1 mixed precision with FP32 basis storage batch=%s" % batch)
    # same stated tolerance as the mixed path: residual < 1e-6 * 12, eigenvalue
    # error < 1e-10 against the closed form
    r = driver.lap_slicing_run(20, 20, 20, 0.6, 1.2, 5, batch=batch, prec=capi.PEVSL_PREC_FP32,
                               basis_prec=capi.PEVSL_PREC_FP32)
    for s in r["slices"]:
        assert s["nev"] == s["nev_exact"]
        assert s["max_rel_err"] < 1e-10
        assert np.all(s["res"] < 1.2e-5)
    ks = r["kstats"]
    # FP32 storage: the reorthogonalization streams 4-byte elements
    assert ks["cgs_bytes"] > 0


def test_fp32_basis_reorth_bytes_halved(lib):
    banner("BenchReorth FP64 vs FP32 basis, 40^3, k = 256")
    ia, ja, a = driver.laplacian3d(40, 40, 40)
    M = driver.Matrix(lib, ia, ja, a)
    out = {}
    try:
        for bp in (capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP32):
            inst = driver.Instance(lib, M, prec=capi.PEVSL_PREC_FP32, basis_prec=bp)
            try:
                sec, byt = C.c_double(), C.c_double()
                assert lib.pEVSL_BenchReorth(inst.p, 256, 5, C.byref(sec), C.byref(byt)) == 0
                out[bp] = (sec.value, byt.value)
            finally:
                inst.close()
    finally:
        M.free()
    assert abs(out[capi.PEVSL_PREC_FP32][1] / out[capi.PEVSL_PREC_FP64][1] - 0.5) < 1e-12
    assert out[capi.PEVSL_PREC_FP32][0] > 0


def test_fp64_rayleigh_ritz_refinement(lib):
    banner("C1 FP64 with the Rayleigh-Ritz Ritz stage")
    r = driver.lap_slicing_run(20, 20, 20, 0.6, 1.2, 5, batch=True, ritz_refine=1)
    for s in r["slices"]:
        assert s["nev"] == s["nev_exact"]
        assert s["max_rel_err"] < 1e-12
        assert np.all(s["res"] < 1e-8)


def test_stream_kernel_end_to_end(lib):
    banner("C1 FP64 with the CSR-stream filter kernel")
    r = driver.lap_slicing_run(20, 20, 20, 0.6, 1.2, 5, lanes=capi.PEVSL_SPMV_STREAM)
    d = orc.run_driver(["-nx", 20, "-ny", 20, "-nz", 20, "-a", 0.6, "-b", 1.2,
                        "-nslices", 5], timeout=120)
    for s, ds in zip(r["slices"], d["slices"]):
        assert s["nev"] == s["nev_exact"] == ds["nev"]
        assert s["info"]["its"] == ds["its"]


@pytest.mark.parametrize("dims,bndtype", [((20, 20, 20), 1), ((30, 30, 1), 2)])
def test_lan_trbounds_against_oracle(lib, dims, bndtype):
    nx, ny, nz = dims
    banner("LanTrbounds %dx%dx%d bndtype %d vs oracle" % (nx, ny, nz, bndtype))
    ia, ja, a = driver.laplacian3d(nx, ny, nz)
    M = driver.Matrix(lib, ia, ja, a)
    inst = driver.Instance(lib, M)  # pEVSL_Start seeds srand(0)
    try:
        v = driver.rand_vinit(lib, M.n)
        rc, lo, hi = inst.lan_trbounds(50, 200, 1e-8, v, bndtype)
    finally:
        inst.close()
        M.free()
    assert rc == 0
    orc.srand(0)
    ov = orc.rand_vec(nx * ny * nz)
    assert np.array_equal(ov, v)
    orc_rc, olo, ohi, its = orc.lan_trbounds(orc.Csr(ia, ja, a), 50, 200, 1e-8, ov, bndtype)
    assert orc_rc == 0
    # same algorithm and start vector; reduction orders differ, so the
    # converged bounds agree to the stopping tolerance
    scale = abs(olo) + abs(ohi)
    assert abs(lo - olo) < 1e-8 * scale and abs(hi - ohi) < 1e-8 * scale
    ev = driver.exact_eig_lap3(nx, ny, nz, -1.0, 13.0)
    assert abs(lo - ev[0]) < 1e-6 * scale and abs(hi - ev[-1]) < 1e-6 * scale


def test_lan_trbounds_unconverged_state_matches_oracle(lib):
    banner("LanTrbounds 60^3, 200 products (not converged) vs oracle")
    ia, ja, a = driver.laplacian3d(60, 60, 60)
    M = driver.Matrix(lib, ia, ja, a)
    inst = driver.Instance(lib, M)
    try:
        v = driver.rand_vinit(lib, M.n)
        rc, lo, hi = inst.lan_trbounds(50, 200, 1e-8, v, 1)
    finally:
        inst.close()
        M.free()
    orc.srand(0)
    ov = orc.rand_vec(M.n)
    orc_rc, olo, ohi, its = orc.lan_trbounds(orc.Csr(ia, ja, a), 50, 200, 1e-8, ov, 1)
    assert rc == orc_rc == 0 and its == 200
    assert abs(lo - olo) < 1e-9 * 12 and abs(hi - ohi) < 1e-9 * 12


def test_lan_trbounds_100cube_closed_form(lib):
    banner("LanTrbounds 100^3 vs closed-form extremes")
    ia, ja, a = driver.laplacian3d(100, 100, 100)
    M = driver.Matrix(lib, ia, ja, a)
    inst = driver.Instance(lib, M)
    try:
        # the ends of the 100^3 spectrum are 3e-3 apart: about 10^3 products
        # to reach the 1e-8 stopping rule with a 50-vector restart
        rc, lo, hi = inst.lan_trbounds(50, 5000, 1e-8)
    finally:
        inst.close()
        M.free()
    assert rc == 0
    t = np.sin(np.pi / (2 * 101.0)) ** 2 * 12
    tmax = np.sin(100 * np.pi / (2 * 101.0)) ** 2 * 12
    assert abs(lo - t) < 1e-6 * 12 and abs(hi - tmax) < 1e-6 * 12
    assert lo <= t + 1e-12 and hi >= tmax - 1e-12


@pytest.mark.parametrize("prec", [0, 1])
def test_kpmdos_uneven_rows_against_oracle(lib, prec):
    """KPM on the power-law n=3000 matrix (rows up to 1,041 entries): the
    nnz-balanced block step (blocks of 32 vectors; rows over 128 (FP64) /
    256 (FP32) entries take the whole-workgroup path) against the oracle on
    the same random stream.  FP64: relative 1e-12 (dot order differs);
    FP32: absolute 2e-5 on mu."""
    banner("Kpmdos power-law n=3000 Mdeg=100 nvec=64 prec=%d (balanced step)" % prec)
    ia, ja, a = driver.powerlaw_sym(3000, seed=11)
    M = driver.Matrix(lib, ia, ja, a)
    lo, hi = -8.0, 70.0  # spectrum [-6.71, 62.5]
    intv = [10.0, 14.0, lo, hi]
    try:
        inst = driver.Instance(lib, M, prec=prec)
        try:
            rc, mu, ecnt = inst.kpmdos(100, 1, 64, intv)
        finally:
            inst.close()
    finally:
        M.free()
    assert rc == 0
    orc.srand(0)
    rc, omu, oecnt = orc.kpmdos(orc.Csr(ia, ja, a), 100, 1, 64, intv)
    assert rc == 0
    if prec == 0:
        np.testing.assert_allclose(mu, omu, rtol=1e-12, atol=1e-15)
        assert abs(ecnt - oecnt) < 1e-8
    else:
        np.testing.assert_allclose(mu, omu, rtol=0, atol=2e-5)
        assert abs(ecnt - oecnt) < 0.05


@pytest.mark.parametrize("prec,basis,rowk", [
    (capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP64, -1),
    (capi.PEVSL_PREC_FP32, capi.PEVSL_PREC_FP32, -1),
    (capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP64, 1),
    (capi.PEVSL_PREC_FP32, capi.PEVSL_PREC_FP32, 2),
    (capi.PEVSL_PREC_FP32, capi.PEVSL_PREC_FP32, 5)])
def test_irregular_small_counts_match_dense(lib, prec, basis, rowk):
    """The irregular-sparsity recipe (power-law rows, Lanczos bounds, a 5%
    share of the spectrum, 16 KPM slices) at n = 3000, where the dense
    eigenvalues give every slice's exact count.  At this size the middle of
    the spectrum is empty, so the interval is centred at 20% of [lmin, lmax]
    (184 eigenvalues).  Rows reach ~2000 entries, so the default (-1) runs
    the nnz-balanced block kernels with their whole-workgroup long-row path
    (16 columns: rows over 256 FP64 / 512 FP32 entries); 1 and 2 are the
    row-per-thread kernels; 5 the segmented Clenshaw kernel, whose row
    blocks of 2048/Q nonzeros leave the longest rows unstaged."""
    banner("irregular n=3000, 16 slices, prec %d basis %d, block kernel %d, batched"
           % (prec, basis, rowk))
    r = driver.irregular_slicing_run(3000, nslices=16, seed=11, nvec=64, center=0.2, prec=prec,
                                     basis_prec=basis, tune=False, block_rowkernel=rowk)
    ia, ja, a = driver.powerlaw_sym(3000, seed=11)
    import scipy.sparse as sp
    ev = np.linalg.eigvalsh(sp.csr_matrix((a, ja, ia), shape=(3000, 3000)).toarray())
    lmin, lmax = r["bounds"]
    assert lmin <= ev[0] + 1e-6 * (abs(lmin) + abs(lmax))
    assert lmax >= ev[-1] - 1e-6 * (abs(lmin) + abs(lmax))
    scale = np.max(np.abs(ev))
    # mixed path: relative threshold 1e-6 of the spectral radius
    tol_res = 1e-8 if prec == capi.PEVSL_PREC_FP64 else 1e-6 * max(abs(lmin), abs(lmax))
    total = 0
    for s in r["slices"]:
        lo, hi = s["intv"]
        ex = ev[(ev >= lo) & (ev <= hi)]
        assert s["nev"] == len(ex), (s["slice"], s["nev"], len(ex))
        if len(ex):
            assert np.max(np.abs(s["lam"] - ex)) < 1e-10 * scale
            assert np.all(np.asarray(s["res"]) < tol_res)
        total += s["nev"]
    assert total == np.sum((ev >= r["intv"][0]) & (ev <= r["intv"][1])) > 100


@pytest.mark.parametrize("a,b,ns", [(0.10, 0.12, 3), (0.12, 0.14, 2)])
def test_matrix_market_flow_stiff1(lib, a, b, ns):
    """The TESTS/MM/MMPLanN.c flow on the reference's stiff1 matrix with the
    intervals of its TESTS/MM/matfile: every slice finds the dense-eigensolve
    count, eigenvalues to 1e-10 relative."""
    import json
    import os
    banner("MM flow stiff1 [%g, %g] %d slices" % (a, b, ns))
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = json.load(open(os.path.join(gdir, "mm_stiff1_eigs.json")))
    r = driver.mm_slicing_run(os.path.join(gdir, "mm_stiff1.mtx"), a, b, ns, idxin=1)
    lmin, lmax = r["bounds"]
    assert lmin <= g["lambda_min"] + 1e-8 and lmax >= g["lambda_max"] - 1e-8
    ev = np.array(g["eigs_0.10_0.14"])
    total = 0
    for s in r["slices"]:
        lo, hi = s["intv"]
        ex = ev[(ev >= lo) & (ev <= hi)]
        assert s["nev"] == len(ex), (s["slice"], s["nev"], len(ex))
        if len(ex):
            assert np.max(np.abs(s["lam"] - ex) / np.abs(ex)) < 1e-10
            assert np.all(np.asarray(s["res"]) < 1e-8)
        total += s["nev"]
    assert total == np.sum((ev >= a) & (ev <= b))



def test_rectangular_parcsr_matvec(lib):
    """TESTS0/Matvec/test4.c on one rank: a rectangular random sparse matrix
    (300 x 500), x_i = sin(i) (TESTS0/Matvec/test1.c:70-94), y = A x on the
    GPU against scipy: relative 1e-14.  A square-only instance rejects it."""
    banner("rectangular ParcsrMatvec 300 x 500")
    import scipy.sparse as sp
    R = sp.random(300, 500, density=0.03, random_state=np.random.default_rng(17), format="csr")
    R.sort_indices()
    ia, ja, a = R.indptr.astype(np.int32), R.indices.astype(np.int32), R.data.astype(np.float64)
    csr = capi.Csr(300, 500, ia.ctypes.data_as(capi.c_int_p), ja.ctypes.data_as(capi.c_int_p),
                   a.ctypes.data_as(capi.c_double_p))
    A = capi.Parcsr()
    assert lib.pEVSL_ParcsrCreate(300, 500, None, None, C.byref(A), capi.MPI_COMM_WORLD) == 0
    assert lib.pEVSL_ParcsrSetup(C.byref(csr), C.byref(A)) == 0
    try:
        x = np.sin(np.arange(500, dtype=np.float64))
        y = np.zeros(300)
        X, Y = capi.Parvec(), capi.Parvec()
        lib.pEVSL_ParvecCreateShell(500, 500, 0, 0, C.byref(X), x.ctypes.data_as(capi.c_double_p))
        lib.pEVSL_ParvecCreateShell(300, 300, 0, 0, C.byref(Y), y.ctypes.data_as(capi.c_double_p))
        lib.pEVSL_ParcsrMatvec(C.byref(A), C.byref(X), C.byref(Y))
        ref = R @ x
        assert np.linalg.norm(y - ref) <= 1e-14 * np.linalg.norm(ref)
        inst = driver.Instance(lib)
        try:
            assert lib.pEVSL_SetAParcsr(inst.p, C.byref(A)) == -1
        finally:
            inst.close()
    finally:
        lib.pEVSL_ParcsrFree(C.byref(A))


@pytest.mark.parametrize("batch", [False, True])
def test_empty_slice_in_a_spectral_gap(lib, batch):
    """A slice that lies in a gap of the spectrum (6^3 Laplacian: no
    eigenvalue in (0.5942, 1.1491)) returns 0 eigenvalues, alone and batched
    with non-empty neighbours, whose counts stay exact."""
    banner("empty slice in a spectral gap, 6^3, batch=%s" % batch)
    r = driver.lap_slicing_run(6, 6, 6, 0.3, 1.6, 3, batch=batch,
                               sli_override=[0.3, 0.75, 0.95, 1.6])
    counts = [s["nev"] for s in r["slices"]]
    assert counts[1] == 0
    for s in r["slices"]:
        assert s["nev"] == s["nev_exact"]
    assert counts[0] > 0 and counts[2] > 0


@pytest.mark.parametrize("a,b", [(2.5, 5.5), (4.5, 9.5), (2.5, 9.5)])
def test_matrix_smaller_than_krylov_dimension(lib, a, b):
    """2^3 Laplacian (n = 8, eigenvalues 3, 5 x3, 7 x3, 9): the Krylov
    dimension is capped at n and the breakdown restarts reach the repeated
    eigenvalues.  Counts equal the oracle's ChebLanNr (the reference
    algorithm).  On [2.5, 9.5], centred on the frame [0, 12], FindPol's
    degree-2 filter is even about 6, so p(3) = p(9) and p(5) = p(7): the
    Ritz vectors mix two eigenvalues and fail the residual test, and the
    reference algorithm, like this library, accepts none.  Off-centre
    intervals find every eigenvalue to 1e-12."""
    banner("n=8 Laplacian [%g, %g]" % (a, b))
    r = driver.lap_slicing_run(2, 2, 2, a, b, 1)
    s = r["slices"][0]
    d = orc.run_driver(["-nx", 2, "-ny", 2, "-nz", 2, "-a", a, "-b", b, "-nslices", 1],
                       timeout=60)
    assert s["nev"] == d["slices"][0]["nev"]
    ex = driver.exact_eig_lap3(2, 2, 2, a, b)
    if (a, b) == (2.5, 9.5):
        assert s["nev"] == 0 and len(ex) == 8
    else:
        assert s["nev"] == len(ex)
        np.testing.assert_allclose(np.sort(s["lam"]), ex, rtol=1e-12)


def _lap2d_with_empty_rows(m=40):
    """m x m 5-point Laplacian with the rows and columns of every 11th index
    and of the contiguous range [600, 1200) removed: about 690 empty rows,
    rows 768-1023 forming a whole CSR-stream row block (256 rows) with no
    entry.  Its spectrum is that of the remaining principal submatrix plus
    zeros."""
    import scipy.sparse as sp
    ia, ja, a = driver.laplacian3d(m, m, 1)
    n = len(ia) - 1
    L = sp.csr_matrix((a, ja, ia), shape=(n, n))
    keep = np.ones(n)
    keep[::11] = 0.0
    keep[600:1200] = 0.0
    D = sp.diags(keep)
    L = (D @ L @ D).tocsr()
    L.eliminate_zeros()
    L.sort_indices()
    return (L.indptr.astype(np.int32), L.indices.astype(np.int32), L.data.astype(np.float64),
            np.linalg.eigvalsh(L.toarray()))


@pytest.mark.parametrize("prec", [capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP32])
def test_cheb_av_empty_rows_against_oracle(lib, prec):
    """Stream-kernel filter on a matrix with empty rows and empty row blocks:
    bitwise with the oracle in FP64 (one sequential sum per row), 1e-4 of the
    filtered vector's size in FP32."""
    banner("ChebAv with empty rows, prec %d" % prec)
    ia, ja, a, _ = _lap2d_with_empty_rows()
    assert ia[1200] == ia[600]
    M = driver.Matrix(lib, ia, ja, a)
    iv = [0.5, 1.5, 0.0, 8.0]
    pol, rc = driver.find_pol(lib, iv)
    opol = orc.Pol(iv)
    assert rc == 0 and pol.deg == opol.deg
    inst = driver.Instance(lib, M, prec=prec, lanes=capi.PEVSL_SPMV_STREAM)
    try:
        v = np.random.default_rng(3).uniform(-1, 1, M.n)
        y = np.zeros(M.n)
        rc = lib.pEVSL_ChebAv(inst.p, C.byref(pol), C.byref(parvec(lib, v)),
                              C.byref(parvec(lib, y)), None)
        assert rc == 0
    finally:
        inst.close()
        lib.pEVSL_FreePol(C.byref(pol))
        M.free()
    ref = orc.cheb_av(orc.Csr(ia, ja, a), opol, v)
    if prec == capi.PEVSL_PREC_FP64:
        assert np.array_equal(y, ref)
    else:
        assert np.max(np.abs(y - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.parametrize("prec,basis,rowk,batch", [
    (capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP64, -1, False),
    (capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP64, 1, True),
    (capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP64, 3, True),
    (capi.PEVSL_PREC_FP32, capi.PEVSL_PREC_FP32, -1, False),
    (capi.PEVSL_PREC_FP32, capi.PEVSL_PREC_FP32, 2, True),
    (capi.PEVSL_PREC_FP32, capi.PEVSL_PREC_FP32, 4, True),
    (capi.PEVSL_PREC_FP32, capi.PEVSL_PREC_FP32, 5, True)])
def test_cheblan_empty_rows_counts_match_dense(lib, prec, basis, rowk, batch):
    """ChebLanNr on the matrix with empty rows, one slice at a time (stream
    kernel) or as a batch of two (row kernels 1 / 2, nnz-balanced kernels
    3 / 4): every slice's count equals the dense count, eigenvalues to 1e-10
    relative; residuals below 1e-8 (FP64) or 1e-6 * 8 (FP32 basis)."""
    banner("ChebLanNr with empty rows, prec %d basis %d rowk %d batch %s"
           % (prec, basis, rowk, batch))
    ia, ja, a, ev = _lap2d_with_empty_rows()
    M = driver.Matrix(lib, ia, ja, a)
    inst = driver.Instance(lib, M, prec=prec, basis_prec=basis, block_rowkernel=rowk,
                           lanes=capi.PEVSL_SPMV_STREAM)
    edges = [0.5, 1.0, 1.5]
    try:
        intvs = [[edges[i], edges[i + 1], 0.0, 8.0] for i in range(2)]
        pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
        counts = [int(np.sum((ev >= iv[0]) & (ev <= iv[1]))) for iv in intvs]
        maxits = [min(max(5 * (c + 2), 300), M.n) for c in counts]
        vinit = driver.rand_vinit(lib, M.n)
        groups = [[0, 1]] if batch else [[0], [1]]
        outs = []
        for g in groups:
            rc, o = driver.cheblan_batch(inst, [intvs[i] for i in g], [maxits[i] for i in g],
                                         1e-8, vinit, [pols[i] for i in g])
            assert rc == 0
            outs += o
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
    finally:
        inst.close()
        M.free()
    tol_res = 1e-8 if basis == capi.PEVSL_PREC_FP64 else 1e-6 * 8.0
    for iv, c, o in zip(intvs, counts, outs):
        lam = np.sort(o["lam"])
        ex = ev[(ev >= iv[0]) & (ev <= iv[1])]
        assert len(lam) == c > 0, (iv, len(lam), c)
        assert np.max(np.abs(lam - ex) / np.abs(ex)) < 1e-10
        assert np.all(np.asarray(o["res"]) < tol_res)
