"""GPU test of the Fortran 90 interface (include/pevsl_f90.h): the routines
are called as a Fortran program calls them -- every argument by reference,
objects as INTEGER*8 handles -- through a TESTS/Lap-style flow, and the
results match the closed-form Laplacian spectrum and the C interface."""
import ctypes as C

import numpy as np
import pytest

from pevsl_amd import capi, driver

pytestmark = pytest.mark.gpu


def ref(x):
    return C.byref(x)


def test_fortran_interface_lap_flow(lib):
    print("\n[gpu case] Fortran interface: 10^3 Laplacian, [0.4, 0.8]", flush=True)
    nx = 10
    n = nx ** 3
    ia, ja, a = driver.laplacian3d(nx, nx, nx)
    comm = C.c_int(0)
    h, Ah, polh = C.c_uint64(), C.c_uint64(), C.c_uint64()
    lib.pevsl_start_f90_(ref(comm), ref(h))
    nn = C.c_int(n)
    lib.pevsl_parcsrcreate_f90_(ref(nn), ref(nn), None, None,
                                ia.ctypes.data_as(C.c_void_p), ja.ctypes.data_as(C.c_void_p),
                                a.ctypes.data_as(C.c_void_p), ref(comm), ref(Ah))
    try:
        lib.pevsl_seta_parcsr_f90_(ref(h), ref(Ah))
        # matvec through the handle equals the dense product
        x = np.random.default_rng(1).uniform(-1, 1, n)
        y = np.zeros(n)
        lib.pevsl_parcsrmatvec_f90_(x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p),
                                    ref(Ah))
        import scipy.sparse as sp
        A = sp.csr_matrix((a, ja, ia), shape=(n, n))
        assert np.allclose(y, A @ x, rtol=1e-14, atol=1e-14)
        mlan, nsteps, tol = C.c_int(50), C.c_int(300), C.c_double(1e-8)
        lmin, lmax = C.c_double(), C.c_double()
        lib.pevsl_lanbounds_f90_(ref(h), ref(mlan), ref(nsteps), ref(tol), ref(lmin), ref(lmax))
        lo = 12 * np.sin(np.pi / (2 * (nx + 1))) ** 2
        hi = 12 * np.sin(nx * np.pi / (2 * (nx + 1))) ** 2
        assert lmin.value <= lo + 1e-8 and lmax.value >= hi - 1e-8
        xintv = np.array([0.4, 0.8, lmin.value, lmax.value])
        Mdeg, damp, nvec = C.c_int(300), C.c_int(1), C.c_int(60)
        mu = np.zeros(301)
        ecnt = C.c_double()
        lib.pevsl_kpmdos_ecnt_f90_(ref(h), ref(Mdeg), ref(damp), ref(nvec),
                                   xintv.ctypes.data_as(C.c_void_p),
                                   mu.ctypes.data_as(C.c_void_p), ref(ecnt))
        exact = driver.exact_eig_lap3(nx, nx, nx, 0.4, 0.8)
        assert abs(ecnt.value - len(exact)) < 0.2 * len(exact)
        ti, te = C.c_double(0.8), C.c_double(0.2)
        lib.pevsl_findpol_f90_(xintv.ctypes.data_as(C.c_void_p), ref(ti), ref(te), ref(polh))
        maxit = C.c_int(min(max(5 * (len(exact) + 2), 300), n))
        lib.pevsl_cheblannr_f90_(ref(h), xintv.ctypes.data_as(C.c_void_p), ref(maxit), ref(tol),
                                 ref(polh))
        nev = C.c_int()
        lib.pevsl_get_nev_f90_(ref(h), ref(nev))
        assert nev.value == len(exact)
        val = np.zeros(nev.value)
        ld = C.c_int(n)
        vec = np.zeros((nev.value, n))
        lib.pevsl_copy_result_f90_(ref(h), val.ctypes.data_as(C.c_void_p),
                                   vec.ctypes.data_as(C.c_void_p), ref(ld))
        assert np.max(np.abs(np.sort(val) - exact) / exact) < 1e-10
        # eigenvectors: residuals of the returned pairs
        for lam, v in zip(val, vec):
            assert np.linalg.norm(A @ v - lam * v) < 1e-8
        lib.pevsl_get_nev_f90_(ref(h), ref(nev))
        assert nev.value == 0  # results handed over
        lib.pevsl_freepol_f90_(ref(polh))
    finally:
        lib.pevsl_finish_f90_(ref(h))
        lib.pevsl_parcsr_free_f90_(ref(Ah))
