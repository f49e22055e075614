"""CPU tests of libpevsl_amd.so: it loads, exports every symbol the headers
declare, and its host-side logic (filter design, slicers, sparse/vector
utilities) agrees with the oracle.  No GPU call is made here."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

from oracle import oracle as orc
from pevsl_amd import capi, driver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    names = set()
    for h in ("pevsl.h", "pevsl_amd.h", "pevsl_itsol.h", "pevsl_f90.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|double|size_t)\s+\*?\s*(\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) > 80
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding covers exactly the declared set
    assert names == set(capi.SIGNATURES), names ^ set(capi.SIGNATURES)


def test_struct_layouts_match_headers():
    # sizes of the public structs on x86-64 (field order of pevsl_struct.h)
    assert C.sizeof(capi.Csr) == 32
    assert C.sizeof(capi.Parvec) == 24
    assert C.sizeof(capi.Parvecs) == 32
    assert C.sizeof(capi.Polparams) == 96
    assert C.sizeof(capi.Stat) == 17 * 8 + 10 * 8
    assert capi.Polparams.deg.offset == 88


def pol_from_lib(lib, intv, **kw):
    pol, rc = driver.find_pol(lib, intv, **kw)
    return pol, rc


@pytest.mark.parametrize("intv", [[0.6, 0.7431715857928962, 0, 12],
                                  [1.098549274637319, 1.2, 0, 12],
                                  [0.0, 0.3, 0, 12], [11.6, 12.0, 0, 12],
                                  [0.6, 1.2, 0, 12], [5.0, 5.001, 0, 12],
                                  [-1.0, 0.2, 0, 12]])
def test_findpol_matches_oracle(lib, intv):
    pol, rc = pol_from_lib(lib, intv)
    opol = orc.Pol(intv)
    assert rc == opol.rc == 0
    assert pol.deg == opol.deg and pol.type == opol.type
    assert abs(pol.bar - opol.bar) <= 1e-14 * max(1, abs(opol.bar))
    assert abs(pol.gam - opol.gam) <= 1e-14
    np.testing.assert_allclose(driver.pol_mu(pol), opol.mu, rtol=1e-13, atol=1e-16)
    lib.pEVSL_FreePol(C.byref(pol))


def test_findpol_readme_degrees(lib):
    g = json.load(open(os.path.join(GOLD, "readme_c1.json")))
    sli = g["sli"]
    for s in g["slices"]:
        pol, rc = pol_from_lib(lib, [sli[s["slice"]], sli[s["slice"] + 1], 0, 12])
        assert rc == 0 and pol.deg == s["deg"]
        lib.pEVSL_FreePol(C.byref(pol))


def test_findpol_forced_degree_and_invalid(lib):
    pol, rc = pol_from_lib(lib, [0.6, 1.2, 0, 12], deg=20)
    opol = orc.Pol([0.6, 1.2, 0, 12], deg=20)
    assert rc == 0 and pol.deg == opol.deg
    lib.pEVSL_FreePol(C.byref(pol))
    _, rc = pol_from_lib(lib, [1.2, 0.6, 0, 12])
    assert rc == -1


def test_spslicer_lan_matches_oracle(lib):
    x = np.linspace(0.6, 1.2, 200)
    y = 1.0 + np.sin(7 * x) ** 2
    sli = driver.spslicer_lan(lib, x, y, 6)
    ref = orc.spslicer_lan(x, y, 6)
    assert np.array_equal(sli, ref)


def test_linspace_sort_part1d_bsearch(lib):
    a = np.zeros(7)
    lib.linspace(0.1, 0.7, 7, a.ctypes.data_as(capi.c_double_p))
    assert a[0] == 0.1 and a[-1] == 0.7 and np.allclose(np.diff(a), 0.1)
    v = np.array([3.0, -1.0, 2.0, 2.0, 0.5])
    ind = np.zeros(5, dtype=np.int32)
    vv = v.copy()
    lib.sort_double(5, vv.ctypes.data_as(capi.c_double_p), ind.ctypes.data_as(capi.c_int_p))
    assert list(vv) == sorted(v) and list(v[ind]) == sorted(v)
    # 1-D partitions (SRC/utils.c:121-124): 9 -> [3,2,2,2], 10 -> [3,3,2,2]
    for ln, sizes in ((9, [3, 2, 2, 2]), (10, [3, 3, 2, 2]), (11, [3, 3, 3, 2])):
        got = []
        for p in range(4):
            idx, j1, j2 = C.c_int(p), C.c_int(), C.c_int()
            lib.pEVSL_Part1d(ln, 4, C.byref(idx), C.byref(j1), C.byref(j2), 1)
            got.append(j2.value - j1.value)
        assert got == sizes
    x = np.array([1, 4, 9, 16], dtype=np.int32)
    assert lib.pEVSL_BinarySearch(x.ctypes.data_as(capi.c_int_p), 4, 9) == 2
    assert lib.pEVSL_BinarySearch(x.ctypes.data_as(capi.c_int_p), 4, 5) == -1


def test_host_csr_matvec_and_vectors(lib):
    ia, ja, a = driver.laplacian3d(8, 7, 6)
    n = len(ia) - 1
    csr = capi.Csr(n, n, ia.ctypes.data_as(capi.c_int_p), ja.ctypes.data_as(capi.c_int_p),
                   a.ctypes.data_as(capi.c_double_p))
    x = np.sin(np.arange(n, dtype=np.float64))
    y = np.zeros(n)
    lib.pEVSL_Matvec(C.byref(csr), x.ctypes.data_as(capi.c_double_p),
                     y.ctypes.data_as(capi.c_double_p))
    ref = orc.Csr(ia, ja, a).matvec(x)
    assert np.array_equal(y, ref)
    # Parvec utilities
    X, Y = capi.Parvec(), capi.Parvec()
    yy = np.cos(np.arange(n, dtype=np.float64))
    lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(X), x.ctypes.data_as(capi.c_double_p))
    lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(Y), yy.ctypes.data_as(capi.c_double_p))
    t = C.c_double()
    lib.pEVSL_ParvecDot(C.byref(X), C.byref(Y), C.byref(t))
    assert abs(t.value - x @ np.cos(np.arange(n))) < 1e-10
    lib.pEVSL_ParvecAxpy(2.0, C.byref(X), C.byref(Y))
    np.testing.assert_allclose(yy, np.cos(np.arange(n)) + 2 * x)
    assert lib.pEVSL_ParvecSameSize(C.byref(X), C.byref(Y)) == 1


def test_parvecs_gemv_gemtv(lib):
    rng = np.random.default_rng(0)
    n, k, ld = 50, 6, 53
    data = rng.normal(size=(k, ld))
    Vs = capi.Parvecs()
    lib.pEVSL_ParvecsCreateShell(n, k, ld, n, 0, 0, C.byref(Vs),
                                 data.ctypes.data_as(capi.c_double_p))
    x = rng.normal(size=k)
    y = rng.normal(size=n)
    y0 = y.copy()
    Y = capi.Parvec()
    lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(Y), y.ctypes.data_as(capi.c_double_p))
    lib.pEVSL_ParvecsGemv(-1.5, C.byref(Vs), k, x.ctypes.data_as(capi.c_double_p), 0.5,
                          C.byref(Y))
    V = data[:, :n].T
    np.testing.assert_allclose(y, 0.5 * y0 - 1.5 * V @ x, atol=1e-12)
    out = np.ones(k)
    lib.pEVSL_ParvecsGemtv(2.0, C.byref(Vs), k, C.byref(Y), 1.0,
                           out.ctypes.data_as(capi.c_double_p))
    np.testing.assert_allclose(out, 1.0 + 2.0 * V.T @ y, atol=1e-10)


def test_coo_csr_roundtrip_and_sortrow(lib):
    class Coo(C.Structure):
        _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("nnz", C.c_int),
                    ("ir", capi.c_int_p), ("jc", capi.c_int_p), ("vv", capi.c_double_p)]
    rng = np.random.default_rng(1)
    nnz, n = 40, 9
    ir = rng.integers(0, n, nnz).astype(np.int32)
    jc = rng.integers(0, n, nnz).astype(np.int32)
    vv = rng.normal(size=nnz)
    coo = Coo(n, n, nnz, ir.ctypes.data_as(capi.c_int_p), jc.ctypes.data_as(capi.c_int_p),
              vv.ctypes.data_as(capi.c_double_p))
    csr = capi.Csr()
    assert lib.pEVSL_CooToCsr(0, C.byref(coo), C.byref(csr)) == 0
    ia = np.ctypeslib.as_array(csr.ia, shape=(n + 1,)).copy()
    assert ia[-1] == nnz and np.all(np.diff(ia) >= 0)
    dense = np.zeros((n, n))
    np.add.at(dense, (ir, jc), vv)
    x = rng.normal(size=n)
    y = np.zeros(n)
    lib.pEVSL_Matvec(C.byref(csr), x.ctypes.data_as(capi.c_double_p),
                     y.ctypes.data_as(capi.c_double_p))
    np.testing.assert_allclose(y, dense @ x, atol=1e-12)
    lib.pEVSL_SortRow(C.byref(csr))
    ja = np.ctypeslib.as_array(csr.ja, shape=(nnz,))
    for r in range(n):
        assert np.all(np.diff(ja[ia[r]:ia[r + 1]]) >= 0)
    lib.pEVSL_FreeCsr(C.byref(csr))


def test_missing_instance_is_an_error(lib):
    # the generalized setters and LanTrbounds reject a missing instance
    assert lib.pEVSL_SetGenEig(None) == -1
    assert lib.pEVSL_SetBSol(None, None, None) == -1
    assert lib.pEVSL_SetLTSol(None, None, None) == -1
    assert lib.pEVSL_SetBMatvec(None, None, None) == -1
    assert lib.pEVSL_LanTrbounds(None, 10, 10, 1e-8, None, 0, None, None, None) == -1
    assert lib.pEVSL_LSPolInfo(None, None, None, None, None) == -1


@pytest.mark.parametrize("seed", [0, 1, 7, 123456789])
def test_rand_stream_equals_libc_rand(lib, seed):
    """pEVSL_ParvecRand draws from the library's own copy of the rand()
    stream; after pEVSL_RandSeed(s) it equals libc srand(s)/rand() (the
    reference's SRC/parvec.c:65-72 values), here checked against the C
    library itself over 2000 draws (past the 344-word warm-up state)."""
    libc = C.CDLL(None)
    libc.srand.argtypes = [C.c_uint]
    libc.rand.restype = C.c_int
    libc.srand(seed)
    t = 2147483647 / 2.0
    want = (np.array([libc.rand() for _ in range(2000)], dtype=np.float64) - t) / t
    n = 2000
    v = np.empty(n)
    pv = capi.Parvec()
    lib.pEVSL_RandSeed(seed)
    lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(pv), v.ctypes.data_as(capi.c_double_p))
    lib.pEVSL_ParvecRand(C.byref(pv))
    assert np.array_equal(v, want)
    # a draw from libc in between does not shift the library's stream
    lib.pEVSL_RandSeed(seed)
    libc.rand()
    lib.pEVSL_ParvecRand(C.byref(pv))
    assert np.array_equal(v, want)


def test_malformed_csr_is_rejected_before_the_device(lib):
    """pEVSL_ParcsrSetup checks ia/ja on the host before any allocation or
    device call, so malformed matrices return -2 (another row count -1)
    without a GPU."""
    import numpy as np

    def setup(n, ia, ja, a, ncol=None):
        ia = np.ascontiguousarray(ia, dtype=np.int32)
        ja = np.ascontiguousarray(ja, dtype=np.int32)
        a = np.ascontiguousarray(a, dtype=np.float64)
        A = capi.Parcsr()
        assert lib.pEVSL_ParcsrCreate(n, ncol or n, None, None, C.byref(A),
                                      capi.MPI_COMM_WORLD) == 0
        csr = capi.Csr(len(ia) - 1, ncol or n, ia.ctypes.data_as(capi.c_int_p),
                       ja.ctypes.data_as(capi.c_int_p), a.ctypes.data_as(capi.c_double_p))
        return lib.pEVSL_ParcsrSetup(C.byref(csr), C.byref(A))

    one = [1.0, 1.0, 1.0]
    assert setup(3, [0, 1, 2, 3], [0, 1, 3], one) == -2     # column 3 of 3
    assert setup(3, [0, 1, 2, 3], [0, -1, 2], one) == -2    # negative column
    assert setup(3, [1, 1, 2, 3], [0, 1, 2], one) == -2     # ia[0] != 0
    assert setup(3, [0, 2, 1, 3], [0, 1, 2], one) == -2     # ia decreasing
    assert setup(3, [0, 1, 2], [0, 1], one[:2]) == -1       # 2 rows for a 3-row matrix


@pytest.mark.parametrize("damping,thresh", [(0, (0.8, 0.2)), (1, (0.8, 0.2)), (2, (0.8, 0.2)),
                                            (2, (0.6, 0.15))])
def test_findpol_random_intervals_match_oracle(lib, damping, thresh):
    """FindPol (host code of the library) against the oracle on 50 seeded
    random intervals per setting: frames [lmin, lmax] of random centre and
    width, intervals of every relative width from 1e-4 to the whole frame,
    including ones that touch an end (left/right filter types), for each
    damping type and both threshold pairs the reference drivers use
    (0.8/0.2: Lap, 0.6/0.15: Gen_Lap and MM).  Degree, type and bar agree
    exactly or to rounding; the coefficients to 1e-13."""
    rng = np.random.default_rng(100 + damping + int(10 * thresh[0]))
    for _ in range(50):
        lmin = rng.uniform(-50, 50)
        lmax = lmin + 10 ** rng.uniform(-1, 3)
        w = (lmax - lmin) * 10 ** rng.uniform(-4, 0)
        a = rng.uniform(lmin, lmax - w)
        if rng.uniform() < 0.15:
            a = lmin  # left end
        elif rng.uniform() < 0.15:
            a = lmax - w  # right end
        intv = [a, a + w, lmin, lmax]
        kw = dict(damping=damping, thresh_int=thresh[0], thresh_ext=thresh[1], max_deg=3000)
        pol, rc = pol_from_lib(lib, intv, **kw)
        opol = orc.Pol(intv, **kw)
        assert rc == opol.rc, intv
        if rc == 0:
            assert pol.deg == opol.deg and pol.type == opol.type, intv
            assert abs(pol.bar - opol.bar) <= 1e-13 * max(1, abs(opol.bar)), intv
            assert abs(pol.gam - opol.gam) <= 1e-13, intv
            np.testing.assert_allclose(driver.pol_mu(pol), opol.mu, rtol=1e-12, atol=1e-15)
        lib.pEVSL_FreePol(C.byref(pol))


def test_coo_to_csr_rejects_out_of_range_indices(lib):
    """An index outside the declared size returns -1 before anything is
    counted or allocated (the reference's loop would index out of bounds)."""
    class Coo(C.Structure):
        _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("nnz", C.c_int),
                    ("ir", capi.c_int_p), ("jc", capi.c_int_p), ("vv", capi.c_double_p)]
    for ir, jc, idx in (([0, 3], [0, 1], 0), ([0, 1], [0, 5], 0), ([-1, 1], [0, 1], 0),
                        ([1, 4], [1, 2], 1), ([0, 2], [1, 2], 1)):
        ir = np.array(ir, dtype=np.int32)
        jc = np.array(jc, dtype=np.int32)
        vv = np.ones(2)
        coo = Coo(3, 3, 2, ir.ctypes.data_as(capi.c_int_p), jc.ctypes.data_as(capi.c_int_p),
                  vv.ctypes.data_as(capi.c_double_p))
        csr = capi.Csr()
        assert lib.pEVSL_CooToCsr(idx, C.byref(coo), C.byref(csr)) == -1, (ir, jc, idx)
        assert not csr.ia


def test_missing_library_fails_loudly(tmp_path):
    """No fallback: loading from a path without the library raises with the
    build command in the message."""
    import pytest
    with pytest.raises(RuntimeError) as e:
        capi.load(str(tmp_path / "libpevsl_amd.so"))
    assert "not found" in str(e.value) and "build" in str(e.value)
