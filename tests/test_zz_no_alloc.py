"""pEVSL_ChebAvNoAlloc (the filter on HBM vectors with the library's own
work array) against pEVSL_ChebAv (host vectors, instance work buffers):
bitwise equal on two calls in a row, the second with a smaller matrix, in
FP64 and on the FP32 filter path; sizes beyond the work array and operators
it does not cover are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from pevsl_amd import capi, driver

pytestmark = pytest.mark.gpu


def banner(msg):
    print("\n[gpu-test] %s" % msg, flush=True)


def parvec(lib, x):
    v = capi.Parvec()
    lib.pEVSL_ParvecCreateShell(len(x), len(x), 0, 0, C.byref(v),
                                x.ctypes.data_as(capi.c_double_p))
    return v


def filter_both(lib, hip, g, prec, intv):
    """(pEVSL_ChebAv, pEVSL_ChebAvNoAlloc) outputs on the g^3 Laplacian."""
    ia, ja, a = driver.laplacian3d(g, g, g)
    M = driver.Matrix(lib, ia, ja, a)
    n = M.n
    pol, rc = driver.find_pol(lib, intv)
    assert rc == 0
    v = np.random.default_rng(g).uniform(-1.0, 1.0, n)
    y_ref = np.empty(n)
    y_na = np.full(n, np.nan)
    d = C.c_void_p()
    nb = v.nbytes
    inst = driver.Instance(lib, M, prec=prec)
    try:
        assert lib.pEVSL_ChebAv(inst.p, C.byref(pol), C.byref(parvec(lib, v)),
                                C.byref(parvec(lib, y_ref)), None) == 0
        assert hip.hipMalloc(C.byref(d), C.c_size_t(2 * nb)) == 0
        dv, dy = d.value, d.value + nb
        assert hip.hipMemcpy(C.c_void_p(dv), v.ctypes.data_as(C.c_void_p), C.c_size_t(nb), 1) == 0
        assert hip.hipMemcpy(C.c_void_p(dy), y_na.ctypes.data_as(C.c_void_p),
                             C.c_size_t(nb), 1) == 0
        print("n %d, degree %d, prec %d" % (n, pol.deg, prec), flush=True)
        assert lib.pEVSL_ChebAvNoAlloc(inst.p, C.byref(pol), C.c_void_p(dv), C.c_void_p(dy)) == 0
        assert hip.hipMemcpy(y_na.ctypes.data_as(C.c_void_p), C.c_void_p(dy), C.c_size_t(nb), 2) == 0
        st = inst.stats()
        assert st["n_polAv"] == 2 and st["n_mvA"] == 2 * pol.deg
    finally:
        if d.value:
            hip.hipFree(d)
        inst.close()
        lib.pEVSL_FreePol(C.byref(pol))
        M.free()
    return y_ref, y_na


@pytest.mark.parametrize("prec", [capi.PEVSL_PREC_FP64, capi.PEVSL_PREC_FP32])
def test_no_alloc_equals_cheb_av_two_calls(lib, prec):
    hip = C.CDLL("libamdhip64.so")
    # second call smaller: the work array still holds the first call's data
    for g, intv in ((40, [0.6, 1.2, 0.0, 12.0]), (20, [1.0, 1.5, 0.0, 12.0])):
        banner("ChebAvNoAlloc vs ChebAv, %d^3, prec %d" % (g, prec))
        y_ref, y_na = filter_both(lib, hip, g, prec, intv)
        assert np.all(np.isfinite(y_na))
        np.testing.assert_array_equal(y_na, y_ref)


def test_no_alloc_refuses_what_it_cannot_cover(lib):
    banner("ChebAvNoAlloc: size above the work array, host-callback operator")
    nmax = lib.pEVSL_ChebAvNoAllocMaxRows()
    assert nmax == 1 << 21
    n = nmax + 1
    ia = np.arange(n + 1, dtype=np.int32)
    ja = np.arange(n, dtype=np.int32)
    M = driver.Matrix(lib, ia, ja, np.full(n, 2.0))
    pol, rc = driver.find_pol(lib, [0.5, 1.0, 0.0, 4.0])
    assert rc == 0
    inst = driver.Instance(lib, M)
    try:
        # refused on the size alone: the vectors are never touched
        assert lib.pEVSL_ChebAvNoAlloc(inst.p, C.byref(pol), C.c_void_p(16), C.c_void_p(16)) == -5
    finally:
        inst.close()
        M.free()
    inst = driver.Instance(lib)
    CB = C.CFUNCTYPE(None, capi.c_double_p, capi.c_double_p, C.c_void_p)
    cb = CB(lambda x, y, data: None)
    try:
        lib.pEVSL_SetProbSizes(inst.p, 8, 8, 0)
        lib.pEVSL_SetAMatvec(inst.p, C.cast(cb, C.c_void_p), None)
        assert lib.pEVSL_ChebAvNoAlloc(inst.p, C.byref(pol), C.c_void_p(16), C.c_void_p(16)) == \
            capi.PEVSL_NOT_IMPLEMENT
    finally:
        inst.close()
        lib.pEVSL_FreePol(C.byref(pol))
