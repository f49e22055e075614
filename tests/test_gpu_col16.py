"""16-bit column offsets (pEVSL_SetColumnOffsets16): the CSR-stream and
row-per-thread block kernels read each nonzero's column as a 16-bit offset
from its row block's first row when the matrix's offsets all fit.

* a banded matrix with offsets of both signs up to 30,000 and a row of 3,000
  entries (a row block of its own, reduced by the whole workgroup): the FP64
  and FP32 CSR-stream filters equal the 32-bit-index runs bit for bit, and
  the FP64 filter the oracle's ChebAv to 1e-12 (the long row is summed as a
  tree; short rows alone are bitwise, as the 20^3 parity tests show with
  the offsets on by default);
* a matrix whose offsets do not fit (a corner entry 69,999 columns away) has
  no 16-bit copy: the same checks with the 32-bit indices on both sides;
* batched solves (FP64 three-term row kernel, FP32 Clenshaw row kernel, two
  block widths) with and without the offsets: the same eigenvalues and step
  counts, and the counted filter bytes fall by 2 bytes per nonzero per
  launch.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from pevsl_amd import capi, driver

pytestmark = pytest.mark.gpu


def banner(msg):
    print("\n[gpu-test] %s" % msg, flush=True)


def sym_csr(n, pairs, diag):
    """CSR of the symmetric matrix with off-diagonal entries (i, j, v) and
    their mirrors, plus the diagonal; columns sorted within rows."""
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [diag]
    for i, j, v in pairs:
        rows += [i, j]
        cols += [j, i]
        vals += [v, v]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    ia = np.zeros(n + 1, dtype=np.int32)
    np.add.at(ia, r + 1, 1)
    return np.cumsum(ia).astype(np.int32), c.astype(np.int32), v.astype(np.float64)


def banded_matrix():
    n = 40000
    rng = np.random.default_rng(3)
    i1 = np.arange(n - 1)
    i2 = np.arange(n - 30000)
    lr = 20000 + np.arange(2, 1501)  # row 20000: 2 * 1499 entries beyond its neighbours
    pairs = [(i1, i1 + 1, rng.uniform(-1, 1, n - 1)),
             (i2, i2 + 30000, rng.uniform(-1, 1, n - 30000)),
             (np.full(lr.size, 20000), lr, rng.uniform(-0.1, 0.1, lr.size)),
             (np.full(lr.size, 20000), 40000 - lr, rng.uniform(-0.1, 0.1, lr.size))]
    return sym_csr(n, pairs, rng.uniform(2.0, 4.0, n))


def corner_matrix():
    n = 70000
    rng = np.random.default_rng(5)
    i1 = np.arange(n - 1)
    pairs = [(i1, i1 + 1, rng.uniform(-1, 1, n - 1)),
             (np.array([0]), np.array([n - 1]), np.array([0.5]))]
    return sym_csr(n, pairs, rng.uniform(2.0, 4.0, n))


def cheb_av(lib, M, pol, v, prec, col16):
    inst = driver.Instance(lib, M, prec=prec, lanes=capi.PEVSL_SPMV_STREAM, col16=col16)
    try:
        used = lib.pEVSL_ColumnOffsets16(inst.p)
        y = np.zeros(M.n)
        pv, py = capi.Parvec(), capi.Parvec()
        lib.pEVSL_ParvecCreateShell(M.n, M.n, 0, 0, C.byref(pv), v.ctypes.data_as(capi.c_double_p))
        lib.pEVSL_ParvecCreateShell(M.n, M.n, 0, 0, C.byref(py), y.ctypes.data_as(capi.c_double_p))
        assert lib.pEVSL_ChebAv(inst.p, C.byref(pol), C.byref(pv), C.byref(py), None) == 0
    finally:
        inst.close()
    return y, used


@pytest.mark.parametrize("which", ["banded", "corner"])
def test_stream_filter_with_offsets(lib, which):
    ia, ja, a = banded_matrix() if which == "banded" else corner_matrix()
    banner("CSR-stream filter, %s matrix n=%d nnz=%d, max row %d" % (
        which, len(ia) - 1, ia[-1], np.max(np.diff(ia))))
    if which == "banded":
        assert np.max(np.diff(ia)) > 2048  # the long-row branch
    M = driver.Matrix(lib, ia, ja, a)
    A = orc.Csr(ia, ja, a)
    try:
        # Gershgorin bound of the spectrum
        lam = float(np.max(np.add.reduceat(np.abs(a), ia[:-1])))
        iv = [2.0, 4.0, -lam, lam]
        pol, rc = driver.find_pol(lib, iv)
        assert rc == 0
        try:
            v = np.random.default_rng(1).uniform(-1, 1, M.n)
            for prec in (capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP32):
                y16, used = cheb_av(lib, M, pol, v, prec, True)
                y32, used_b = cheb_av(lib, M, pol, v, prec, False)
                assert used == (1 if which == "banded" else 0) and used_b == 0
                np.testing.assert_array_equal(y16, y32)
                if prec == capi.PEVSL_PREC_FP64:
                    ref = orc.cheb_av(A, orc.Pol(iv), v)
                    assert np.max(np.abs(y16 - ref)) <= 1e-12 * np.max(np.abs(ref))
        finally:
            lib.pEVSL_FreePol(C.byref(pol))
    finally:
        M.free()


@pytest.mark.parametrize("prec,kind,width", [(capi.PEVSL_PREC_FP64, 1, 4),
                                             (capi.PEVSL_PREC_FP32, 2, 4),
                                             (capi.PEVSL_PREC_FP32, 2, 8),
                                             (capi.PEVSL_PREC_FP32, 1, 8)])
def test_batched_solve_with_and_without_offsets(lib, prec, kind, width):
    banner("batched 16^3, 5 slices, prec %d kernel %d width %d, 16-bit offsets on/off" % (
        prec, kind, width))
    g = 16
    ia, ja, a = driver.laplacian3d(g, g, g)
    M = driver.Matrix(lib, ia, ja, a)
    edges = np.linspace(0.5, 1.5, 6)
    intvs = [[edges[i], edges[i + 1], 0.0, 12.0] for i in range(5)]
    pols = [driver.find_pol(lib, iv)[0] for iv in intvs]
    vinit = driver.rand_vinit(lib, M.n)
    runs = []
    try:
        for col16 in (True, False):
            inst = driver.Instance(lib, M, prec=prec, block_rowkernel=kind, block_width=width,
                                   col16=col16)
            try:
                assert lib.pEVSL_ColumnOffsets16(inst.p) == (1 if col16 else 0)
                rc, outs = driver.cheblan_batch(inst, intvs, [400] * 5, 1e-8, vinit, pols)
                assert rc == 0
                runs.append((outs, inst.kernel_stats()))
            finally:
                inst.close()
    finally:
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
        M.free()
    (o16, k16), (o32, k32) = runs
    for iv, x, y in zip(intvs, o16, o32):
        assert x["info"]["its"] == y["info"]["its"]
        np.testing.assert_array_equal(x["lam"], y["lam"])
        ex = orc.exact_eig_lap3(g, g, g, iv[0], iv[1])
        assert len(x["lam"]) == len(ex) > 0
        np.testing.assert_allclose(np.sort(x["lam"]), ex, rtol=1e-10)
    # every counted launch reads the matrix once: 2 index bytes per nonzero
    # fewer with the offsets
    assert k16["filter_launches"] == k32["filter_launches"] > 0
    saved = k32["filter_bytes"] - k16["filter_bytes"]
    assert saved == 2.0 * int(ia[-1]) * k16["filter_launches"], (saved, k16["filter_launches"])
