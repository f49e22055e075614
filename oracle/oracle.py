"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results.  The product (pevsl_amd) never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")
DRIVER = os.path.join(_HERE, "_build", "lap_oracle")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class OrcCsr(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("ia", ip), ("ja", ip),
                ("a", dp)]


class OrcPol(C.Structure):
    _fields_ = [("max_deg", C.c_int), ("min_deg", C.c_int), ("damping", C.c_int),
                ("thresh_ext", C.c_double), ("thresh_int", C.c_double),
                ("intvtol", C.c_double), ("type", C.c_int), ("mu", dp),
                ("cc", C.c_double), ("dd", C.c_double), ("gam", C.c_double),
                ("bar", C.c_double), ("deg", C.c_int)]


class OrcLanInfo(C.Structure):
    _fields_ = [("kdim", C.c_int), ("its", C.c_int), ("nbreak", C.c_int),
                ("nconv", C.c_int), ("nmatvec", C.c_int), ("tr_last", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_lap3d.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcCsr)]
        L.orc_csr_free.argtypes = [C.POINTER(OrcCsr)]
        L.orc_csrmv.argtypes = [C.POINTER(OrcCsr), dp, dp]
        L.orc_exact_eig_lap3.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                         C.c_double, dp, C.c_int]
        L.orc_srand.argtypes = [C.c_uint]
        L.orc_rand_vec.argtypes = [C.c_int, dp]
        L.orc_tridiag_eig.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_cgs_dgks.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, dp, dp]
        L.orc_dampcf.argtypes = [C.c_int, C.c_int, dp]
        L.orc_kpmdos.argtypes = [C.POINTER(OrcCsr), C.c_int, C.c_int, C.c_int, dp,
                                 dp, dp]
        L.orc_intchx.argtypes = [C.c_int, dp, C.c_int, dp, dp]
        L.orc_sym_eig.argtypes = [C.c_int, dp, dp, dp]
        L.orc_lan_trbounds.argtypes = [C.POINTER(OrcCsr), C.c_int, C.c_int, C.c_double, dp,
                                       C.c_int, dp, dp, ip]
        L.orc_spslicer_kpm.argtypes = [dp, dp, C.c_int, dp, C.c_int, C.c_int]
        L.orc_set_pol_def.argtypes = [C.POINTER(OrcPol)]
        L.orc_find_pol.argtypes = [dp, C.POINTER(OrcPol)]
        L.orc_free_pol.argtypes = [C.POINTER(OrcPol)]
        L.orc_chebxPltd.argtypes = [C.c_int, dp, C.c_int, dp, dp]
        L.orc_cheb_av.argtypes = [C.POINTER(OrcCsr), C.POINTER(OrcPol), dp, dp, dp]
        L.orc_cheblan_nr.argtypes = [C.POINTER(OrcCsr), dp, C.c_int, C.c_double, dp,
                                     C.POINTER(OrcPol), ip, dp, dp, dp,
                                     C.POINTER(OrcLanInfo)]
        L.orc_simpson.argtypes = [dp, dp, C.c_int]
        L.orc_landos.argtypes = [C.POINTER(OrcCsr), C.c_int, C.c_int, C.c_int, dp, dp,
                                 dp, dp, C.c_double]
        L.orc_spslicer_lan.argtypes = [dp, dp, C.c_int, C.c_int, dp]
        _lib = L
    return _lib


def P(x):
    return x.ctypes.data_as(dp)


class Csr:
    """Oracle-side CSR; either the built-in Laplacian or given arrays."""

    def __init__(self, ia=None, ja=None, a=None, lap=None):
        self.m = OrcCsr()
        if lap is not None:
            lib().orc_lap3d(*lap, C.byref(self.m))
            n = self.m.nrows
            nnz = self.m.ia[n]
            self.ia = np.ctypeslib.as_array(self.m.ia, shape=(n + 1,)).copy()
            self.ja = np.ctypeslib.as_array(self.m.ja, shape=(nnz,)).copy()
            self.a = np.ctypeslib.as_array(self.m.a, shape=(nnz,)).copy()
            lib().orc_csr_free(C.byref(self.m))
        else:
            self.ia = np.ascontiguousarray(ia, dtype=np.int32)
            self.ja = np.ascontiguousarray(ja, dtype=np.int32)
            self.a = np.ascontiguousarray(a, dtype=np.float64)
        n = len(self.ia) - 1
        self.m = OrcCsr(n, n, self.ia.ctypes.data_as(ip), self.ja.ctypes.data_as(ip),
                        self.a.ctypes.data_as(dp))
        self.n = n

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().orc_csrmv(C.byref(self.m), P(x), P(y))
        return y


def srand(seed):
    lib().orc_srand(seed)


def rand_vec(n):
    x = np.empty(n)
    lib().orc_rand_vec(n, P(x))
    return x


def tridiag_eig(d, e, vectors=False):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(np.append(e, 0.0)[: max(len(d), 1)], dtype=np.float64)
    n = len(d)
    w = np.empty(n)
    z = np.empty((n, n)) if vectors else None
    lib().orc_tridiag_eig(n, P(d), P(e), P(w), P(z) if vectors else None)
    return (w, z.T.copy()) if vectors else w  # z columns = eigenvectors


def dampcf(m, damping):
    jac = np.empty(m + 1)
    lib().orc_dampcf(m, damping, P(jac))
    return jac


def kpmdos(A, Mdeg, damping, nvec, intv):
    mu = np.zeros(Mdeg + 1)
    e = C.c_double()
    iv = np.ascontiguousarray(intv, dtype=np.float64)
    rc = lib().orc_kpmdos(C.byref(A.m), Mdeg, damping, nvec, P(iv), P(mu), C.byref(e))
    return rc, mu, e.value


def sym_eig(A):
    """Eigenvalues (ascending) and eigenvectors of a dense symmetric matrix."""
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.empty(n)
    Z = np.empty((n, n), order="F")
    rc = lib().orc_sym_eig(n, A.ctypes.data_as(dp), w.ctypes.data_as(dp), Z.ctypes.data_as(dp))
    return rc, w, Z


def lan_trbounds(A, lanm, maxit, tol, vinit, bndtype=1):
    """Thick-restart Lanczos spectrum bounds; returns (rc, lmin, lmax, its)."""
    v = np.ascontiguousarray(vinit, dtype=np.float64)
    lo, hi = C.c_double(), C.c_double()
    its = C.c_int()
    rc = lib().orc_lan_trbounds(C.byref(A.m), lanm, maxit, tol, P(v), bndtype, C.byref(lo),
                                C.byref(hi), C.byref(its))
    return rc, lo.value, hi.value, its.value


def intchx(mu, xi):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    yi = np.empty_like(xi)
    lib().orc_intchx(len(mu) - 1, P(mu), len(xi), P(xi), P(yi))
    return yi


def spslicer_kpm(mu, intv, n_int, npts):
    sli = np.zeros(n_int + 1)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    iv = np.ascontiguousarray(intv, dtype=np.float64)
    rc = lib().orc_spslicer_kpm(P(sli), P(mu), len(mu) - 1, P(iv), n_int, npts)
    return rc, sli


class Pol:
    def __init__(self, intv, damping=2, thresh_int=0.8, thresh_ext=0.2, max_deg=3000,
                 deg=0):
        self.p = OrcPol()
        lib().orc_set_pol_def(C.byref(self.p))
        self.p.damping = damping
        self.p.thresh_int = thresh_int
        self.p.thresh_ext = thresh_ext
        self.p.max_deg = max_deg
        self.p.deg = deg
        iv = np.ascontiguousarray(intv, dtype=np.float64)
        self.rc = lib().orc_find_pol(P(iv), C.byref(self.p))
        if self.rc == 0:
            self.mu = np.ctypeslib.as_array(self.p.mu, shape=(self.p.deg + 1,)).copy()
        self.deg, self.type = self.p.deg, self.p.type
        self.bar, self.gam, self.cc, self.dd = self.p.bar, self.p.gam, self.p.cc, self.p.dd

    def __del__(self):
        try:
            lib().orc_free_pol(C.byref(self.p))
        except Exception:
            pass


def chebxPltd(mu, x):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_chebxPltd(len(mu) - 1, P(mu), len(x), P(x), P(y))
    return y


def cheb_av(A, pol, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    y = np.empty_like(v)
    w = np.empty(3 * len(v))
    lib().orc_cheb_av(C.byref(A.m), C.byref(pol.p), P(v), P(y), P(w))
    return y


def cgs_dgks(V, v, i_max=2):
    V = np.asfortranarray(V, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    n, k = V.shape
    nrm = C.c_double()
    w = np.empty(max(k, 1))
    lib().orc_cgs_dgks(n, k, i_max, V.ctypes.data_as(dp), n, P(v), C.byref(nrm), P(w))
    return v, nrm.value


def cheblan_nr(A, intv, maxit, tol, vinit, pol, want_vectors=False):
    n = A.n
    lam = np.zeros(maxit)
    res = np.zeros(maxit)
    Y = np.zeros((maxit, n)) if want_vectors else None
    nev = C.c_int()
    info = OrcLanInfo()
    iv = np.ascontiguousarray(intv, dtype=np.float64)
    vi = np.ascontiguousarray(vinit, dtype=np.float64)
    rc = lib().orc_cheblan_nr(C.byref(A.m), P(iv), maxit, tol, P(vi), C.byref(pol.p),
                              C.byref(nev), P(lam), P(res), P(Y) if want_vectors else None,
                              C.byref(info))
    k = nev.value
    return {"rc": rc, "lam": lam[:k].copy(), "res": res[:k].copy(),
            "vecs": Y[:k].copy() if want_vectors else None,
            "its": info.its, "kdim": info.kdim, "nbreak": info.nbreak,
            "nconv": info.nconv, "tr": info.tr_last}


def simpson(xi, yi):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    yi = np.array(yi, dtype=np.float64)
    lib().orc_simpson(P(xi), P(yi), len(xi))
    return yi


def landos(A, nvec, msteps, npts, intv, sigma_mult=1.0):
    xd = np.zeros(npts)
    yd = np.zeros(npts)
    ne = C.c_double()
    iv = np.ascontiguousarray(intv, dtype=np.float64)
    rc = lib().orc_landos(C.byref(A.m), nvec, msteps, npts, P(xd), P(yd), C.byref(ne),
                          P(iv), sigma_mult)
    return rc, xd, yd, ne.value


def spslicer_lan(xd, yd, n_int):
    sli = np.zeros(n_int + 1)
    xd = np.ascontiguousarray(xd, dtype=np.float64)
    yd = np.array(yd, dtype=np.float64)
    lib().orc_spslicer_lan(P(xd), P(yd), n_int, len(xd), P(sli))
    return sli


def exact_eig_lap3(nx, ny, nz, a, b):
    cnt = lib().orc_exact_eig_lap3(nx, ny, nz, a, b, None, 0)
    out = np.zeros(max(cnt, 1))
    lib().orc_exact_eig_lap3(nx, ny, nz, a, b, P(out), cnt)
    return out[:cnt]


def run_driver(args, timeout=None, threads=None):
    """Run the oracle driver binary (LapPLanN.c flow); returns parsed JSON.
    `threads` sets OMP_NUM_THREADS for the run (default: inherited)."""
    import json
    if not os.path.exists(DRIVER):
        build()
    env = None
    if threads is not None:
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True,
                       text=True, check=True, timeout=timeout, env=env)
    return json.loads(r.stdout)
