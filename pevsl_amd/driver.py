"""Host driver of the spectrum-slicing run through the C API.

`lap_slicing_run` follows TESTS/Lap/LapPLanN.c:15-323 for one process and one
group: Laplacian matrix, pEVSL_Start (srand(0)), KPM DOS (Mdeg 300, Jackson,
nvec 60) or Lanczos DOS, equal-count slicing, one random vinit shared by all
slices, then per slice FindPol (damping 2, thresholds 0.8/0.2, max_deg 3000)
and ChebLanNr with mlan = min(max(5 nev, 300), n), tol 1e-8.  Slices can also
be solved together with pEVSL_ChebLanNrBatch.

Every product, filter and Lanczos step runs in libpevsl_amd.so on the GPU;
this module only prepares host arrays, as the reference's C driver does.
"""
import ctypes as C
import time

import numpy as np

from . import capi


def laplacian3d(nx, ny, nz):
    """7-point (5-point for nz == 1) Dirichlet Laplacian, natural order
    idx = k*nx*ny + i*nx + j with j fastest, entries of a row ordered
    -z, -y, -x, diag, +x, +y, +z (TESTS/COMMON/lapl.c:4-70).
    Returns (ia, ja, a) as int32/int32/float64 arrays."""
    n = nx * ny * nz
    r = np.arange(n, dtype=np.int64)
    k = r // (nx * ny)
    i = (r - k * nx * ny) // nx
    j = r - k * nx * ny - i * nx
    dg = 6.0 if nz > 1 else 4.0
    cols, vals, ok = [], [], []
    for cond, off, v in ((k > 0, -nx * ny, -1.0), (i > 0, -nx, -1.0),
                         (j > 0, -1, -1.0), (np.ones(n, bool), 0, dg),
                         (j < nx - 1, 1, -1.0), (i < ny - 1, nx, -1.0),
                         (k < nz - 1, nx * ny, -1.0)):
        cols.append(r + off)
        vals.append(np.full(n, v))
        ok.append(cond)
    cols = np.stack(cols, 1)
    vals = np.stack(vals, 1)
    ok = np.stack(ok, 1)
    counts = ok.sum(1)
    ia = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(counts, out=ia[1:])
    ja = cols[ok].astype(np.int32)
    a = vals[ok].astype(np.float64)
    return ia, ja, a


def powerlaw_sym(n, mean=16.0, max_len=2000, alpha=2.2, diag_frac=0.1, seed=0):
    """Seeded irregular test matrix: row i of A0 has L_i entries, L_i drawn
    from a power law P(L) ~ L^-alpha truncated at max_len and scaled to the
    given mean, at uniformly random columns with values U(-1, 1);
    A = A0 + A0^T (duplicates summed) plus diag(diag_frac * sum_j |A_ij|).
    Returns (ia, ja, a) as int32/int32/float64 CSR with sorted columns."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)

    def tmean(xmin):  # mean of the continuous power law on [xmin, max_len]
        r = xmin / max_len
        return xmin * (alpha - 1) / (alpha - 2) * (1 - r ** (alpha - 2)) / (1 - r ** (alpha - 1))
    lo, hi = 0.5, mean
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if tmean(mid) < mean else (lo, mid)
    xmin = 0.5 * (lo + hi)
    u = rng.random(n)
    r = xmin / max_len
    # inverse CDF of the truncated power law
    x = xmin * (1 - u * (1 - r ** (alpha - 1))) ** (-1.0 / (alpha - 1))
    L = np.clip(np.rint(x), 1, max_len).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), L)
    cols = rng.integers(0, n, size=rows.size, dtype=np.int64)
    vals = rng.uniform(-1.0, 1.0, size=rows.size)
    A0 = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    del rows, cols, vals
    A = (A0 + A0.T).tocsr()
    del A0
    A.sum_duplicates()
    d = diag_frac * np.asarray(abs(A).sum(axis=1)).ravel()
    A = (A + sp.diags(d)).tocsr()
    A.sort_indices()
    return (A.indptr.astype(np.int32), A.indices.astype(np.int32),
            A.data.astype(np.float64))


def read_matrix_market(path, idxin=1):
    """Sparse real/integer coordinate Matrix Market file -> (ia, ja, a) CSR
    (int32/int32/float64, columns sorted), with the reference reader's rules
    (TESTS/COMMON/parmm.c:11-162): indices start at idxin (1 for MM1, 0 for
    MM0); a 'symmetric' file stores one triangle and every off-diagonal
    entry is mirrored.  Duplicate entries are summed.  As in the reference,
    any positive idxin means 1-based.  Indices are read as exact integers
    (the reference parses them through a float, exact only below 2^24)."""
    idxin = 1 if idxin > 0 else 0
    with open(path, "r") as f:
        banner = f.readline().split()
        if len(banner) < 5 or banner[0] != "%%MatrixMarket" or banner[1].lower() != "matrix":
            raise ValueError("%s: not a Matrix Market matrix" % path)
        fmt, field, symm = banner[2].lower(), banner[3].lower(), banner[4].lower()
        if fmt != "coordinate" or field not in ("real", "integer"):
            raise ValueError("%s: only sparse real/integer coordinate matrices" % path)
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nrow, ncol, nnz = (int(v) for v in line.split()[:3])
        data = np.loadtxt(f, dtype=np.float64, ndmin=2, max_rows=nnz)
    if data.shape[0] != nnz:
        raise ValueError("%s: %d entries, header says %d" % (path, data.shape[0], nnz))
    r = data[:, 0].astype(np.int64) - idxin
    c = data[:, 1].astype(np.int64) - idxin
    v = data[:, 2]
    if symm == "symmetric":
        off = r != c
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, v[off]])
    import scipy.sparse as sp
    A = sp.coo_matrix((v, (r, c)), shape=(nrow, ncol)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def mm_slicing_run(path, a, b, nslices, idxin=1, prec=capi.PEVSL_PREC_FP64,
                   basis_prec=capi.PEVSL_PREC_FP64, batch=False, log=None, verbose=False):
    """End-to-end run of the TESTS/MM/MMPLanN.c flow (KPM branch): Matrix
    Market matrix, bounds by pEVSL_LanTrbounds(50, 200, 1e-8), KPM DOS
    (Mdeg 50, nvec 60, Jackson) and SpslicerKpm with npts = 10 ecount,
    FindPol (damping 2, thresholds 0.6/0.15, max_deg 3000), ChebLanNr with
    mlan = min(max(5 nev, 300), n), tol 1e-8, one vinit for all slices."""
    lib = capi.load()
    t0 = time.time()
    ia, ja, av = read_matrix_market(path, idxin)
    M = Matrix(lib, ia, ja, av)
    n = M.n
    out = {"n": n, "nnz": M.nnz}
    inst = Instance(lib, M, prec=prec, basis_prec=basis_prec)
    try:
        vinit = rand_vinit(lib, n)
        rc, lmin, lmax = inst.lan_trbounds(50, 200, 1e-8, vinit, 1)
        if rc:
            raise RuntimeError("LanTrbounds error %d" % rc)
        xintv = np.array([a, b, lmin, lmax])
        rc, mu, ecount = inst.kpmdos(50, 1, 60, xintv)
        if rc:
            raise RuntimeError("kpmdos error %d" % rc)
        rc, sli = spslicer_kpm(lib, mu, xintv, nslices, int(10 * ecount))
        if rc:
            raise RuntimeError("spslicer error %d" % rc)
        out.update(bounds=[lmin, lmax], ecount=ecount, sli=sli, t_dos=time.time() - t0)
        ev_int = int(1 + ecount / float(nslices))
        nev = ev_int + 2
        mlan = min(max(5 * nev, 300), n)
        pols = [find_pol(lib, [sli[s], sli[s + 1], lmin, lmax], damping=2, thresh_int=0.6,
                         thresh_ext=0.15, max_deg=3000)[0] for s in range(nslices)]
        intvs = [[sli[s], sli[s + 1], lmin, lmax] for s in range(nslices)]
        groups = [list(range(nslices))] if batch else [[s] for s in range(nslices)]
        t3 = time.time()
        res = []
        for g in groups:
            rc, outs = cheblan_batch(inst, [intvs[i] for i in g], [mlan] * len(g), 1e-8, vinit,
                                     [pols[i] for i in g], False, verbose)
            if rc:
                raise RuntimeError("ChebLanNr error %d" % rc)
            for i, o in zip(g, outs):
                res.append({"slice": i, "intv": intvs[i][:2], "deg": pols[i].deg,
                            "nev": len(o["lam"]), "lam": np.sort(o["lam"]), "res": o["res"],
                            "info": o["info"], "mlan": mlan})
                if log:
                    log("slice %d: deg %d nev %d" % (i, pols[i].deg, len(o["lam"])))
        out["t_lanczos"] = time.time() - t3
        out["slices"] = res
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
    finally:
        inst.close()
        M.free()
    out["t_total"] = time.time() - t0
    return out


def exact_eig_lap3(nx, ny, nz, a, b):
    """Closed-form eigenvalues in [a - eps, b + eps], sorted
    (TESTS/COMMON/lapl.c:139-169)."""
    th = [0.5 * np.pi / (m + 1.0) for m in (nx, ny, nz)]
    tx = np.sin(np.arange(1, nx + 1) * th[0])
    ty = np.sin(np.arange(1, ny + 1) * th[1])
    tz = np.zeros(1) if nz == 1 else np.sin(np.arange(1, nz + 1) * th[2])
    ev = 4 * (tx[:, None, None] ** 2 + ty[None, :, None] ** 2 + tz[None, None, :] ** 2)
    ev = ev.ravel()
    eps = np.finfo(float).eps
    return np.sort(ev[(ev >= a - eps) & (ev <= b + eps)])


class Matrix:
    """A pevsl_Parcsr built through pEVSL_ParcsrCreate/Setup (device copy)."""

    def __init__(self, lib, ia, ja, a):
        self.lib = lib
        n = len(ia) - 1
        self._ia = np.ascontiguousarray(ia, dtype=np.int32)
        self._ja = np.ascontiguousarray(ja, dtype=np.int32)
        self._a = np.ascontiguousarray(a, dtype=np.float64)
        csr = capi.Csr(n, n, self._ia.ctypes.data_as(capi.c_int_p),
                       self._ja.ctypes.data_as(capi.c_int_p),
                       self._a.ctypes.data_as(capi.c_double_p))
        self.A = capi.Parcsr()
        rc = lib.pEVSL_ParcsrCreate(n, n, None, None, C.byref(self.A),
                                    capi.MPI_COMM_WORLD)
        if rc:
            raise RuntimeError("ParcsrCreate failed: %d" % rc)
        rc = lib.pEVSL_ParcsrSetup(C.byref(csr), C.byref(self.A))
        if rc:
            raise RuntimeError("ParcsrSetup failed: %d" % rc)
        self.n = n
        self.nnz = int(self._ia[-1])

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        X, Y = capi.Parvec(), capi.Parvec()
        self.lib.pEVSL_ParvecCreateShell(self.n, self.n, 0, 0, C.byref(X),
                                         x.ctypes.data_as(capi.c_double_p))
        self.lib.pEVSL_ParvecCreateShell(self.n, self.n, 0, 0, C.byref(Y),
                                         y.ctypes.data_as(capi.c_double_p))
        self.lib.pEVSL_ParcsrMatvec(C.byref(self.A), C.byref(X), C.byref(Y))
        return y

    def free(self):
        if self.A.dev:
            self.lib.pEVSL_ParcsrFree(C.byref(self.A))


class Instance:
    """pEVSL_Start / pEVSL_Finish around one matrix."""

    def __init__(self, lib=None, matrix=None, prec=capi.PEVSL_PREC_FP64,
                 rand_mode=capi.PEVSL_RAND_LIBC, seed=0, keep_vectors=False,
                 lanes=0, block_width=0, ritz_refine=-1, mixed_tol=None,
                 block_rowkernel=-1, basis_prec=capi.PEVSL_PREC_FP64, col16=True):
        self.lib = lib or capi.load()
        self.p = capi.DataP()
        rc = self.lib.pEVSL_Start(capi.MPI_COMM_WORLD, C.byref(self.p))
        if rc:
            raise RuntimeError("pEVSL_Start failed: %d" % rc)
        if matrix is not None:
            self.lib.pEVSL_SetAParcsr(self.p, C.byref(matrix.A))
        self.lib.pEVSL_SetFilterPrecision(self.p, prec)
        self.lib.pEVSL_SetRandMode(self.p, rand_mode, seed)
        self.lib.pEVSL_SetKeepVectors(self.p, 1 if keep_vectors else 0)
        self.lib.pEVSL_SetSpmvTuning(self.p, lanes, block_width)
        self.lib.pEVSL_SetRitzRefine(self.p, ritz_refine)
        self.lib.pEVSL_SetBlockKernel(self.p, block_rowkernel)
        self.lib.pEVSL_SetBasisPrecision(self.p, basis_prec)
        if not col16:
            self.lib.pEVSL_SetColumnOffsets16(self.p, 0)
        if mixed_tol is not None:
            self.lib.pEVSL_SetMixedTol(self.p, mixed_tol)

    def close(self):
        if self.p:
            self.lib.pEVSL_Finish(self.p)
            self.p = capi.DataP()

    def kpmdos(self, Mdeg, damping, nvec, intv):
        mu = np.zeros(Mdeg + 1)
        ecnt = C.c_double()
        iv = np.ascontiguousarray(intv, dtype=np.float64)
        rc = self.lib.pEVSL_Kpmdos(self.p, Mdeg, damping, nvec,
                                   iv.ctypes.data_as(capi.c_double_p), 1, 0, -1,
                                   mu.ctypes.data_as(capi.c_double_p), C.byref(ecnt))
        return rc, mu, ecnt.value

    def lan_trbounds(self, lanm=50, maxit=200, tol=1e-8, vinit=None, bndtype=1):
        """pEVSL_LanTrbounds: (rc, lmin, lmax) by thick-restart Lanczos.
        vinit defaults to a libc-rand() vector (pEVSL_ParvecRand)."""
        v = vinit if vinit is not None else rand_vinit(self.lib, self.n_rows())
        vv = np.ascontiguousarray(v, dtype=np.float64)
        pv = capi.Parvec()
        self.lib.pEVSL_ParvecCreateShell(len(vv), len(vv), 0, 0, C.byref(pv),
                                         vv.ctypes.data_as(capi.c_double_p))
        lo, hi = C.c_double(), C.c_double()
        rc = self.lib.pEVSL_LanTrbounds(self.p, lanm, maxit, tol, C.byref(pv), bndtype,
                                        C.byref(lo), C.byref(hi), None)
        return rc, lo.value, hi.value

    def n_rows(self):
        return self.p.contents.n

    def landos(self, nvec, msteps, npts, intv):
        xd = np.zeros(npts)
        yd = np.zeros(npts)
        neig = C.c_double()
        iv = np.ascontiguousarray(intv, dtype=np.float64)
        rc = self.lib.pEVSL_LanDosG(self.p, nvec, msteps, npts,
                                    xd.ctypes.data_as(capi.c_double_p),
                                    yd.ctypes.data_as(capi.c_double_p),
                                    C.byref(neig), iv.ctypes.data_as(capi.c_double_p),
                                    1, 0, -1)
        return rc, xd, yd, neig.value

    def kernel_stats(self):
        ks = capi.KernelStats()
        self.lib.pEVSL_GetKernelStats(self.p, C.byref(ks))
        return {f: getattr(ks, f) for f, _ in capi.KernelStats._fields_}

    def kernel_times(self):
        """Per-kernel HIP-event timings since the last stats reset:
        {name: {"launches", "seconds", "bytes"}}."""
        cnt = C.c_int(0)
        self.lib.pEVSL_GetKernelTimes(self.p, 0, None, C.byref(cnt))
        arr = (capi.KernelTime * max(1, cnt.value))()
        self.lib.pEVSL_GetKernelTimes(self.p, cnt.value, arr, C.byref(cnt))
        return {arr[i].name.decode(): {"launches": arr[i].launches, "seconds": arr[i].seconds,
                                       "bytes": arr[i].bytes} for i in range(cnt.value)}

    def stats(self):
        """The instance's pevsl_Stat as a dict."""
        st = self.p.contents.stats.contents
        return {f: getattr(st, f) for f, _ in capi.Stat._fields_}

    def stats_reset(self):
        self.lib.pEVSL_StatsReset(self.p)

    def lan_info(self, s=0):
        li = capi.LanInfo()
        if self.lib.pEVSL_GetLanInfo(self.p, s, C.byref(li)):
            return None
        return {f: getattr(li, f) for f, _ in capi.LanInfo._fields_}


def spslicer_kpm(lib, mu, intv, nslices, npts):
    sli = np.zeros(nslices + 1)
    iv = np.ascontiguousarray(intv, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    rc = lib.pEVSL_SpslicerKpm(sli.ctypes.data_as(capi.c_double_p),
                               mu.ctypes.data_as(capi.c_double_p), len(mu) - 1,
                               iv.ctypes.data_as(capi.c_double_p), nslices, npts)
    return rc, sli


def spslicer_lan(lib, xd, yd, nslices):
    sli = np.zeros(nslices + 1)
    xd = np.ascontiguousarray(xd, dtype=np.float64)
    yd = np.array(yd, dtype=np.float64)
    lib.pEVSL_SpslicerLan(xd.ctypes.data_as(capi.c_double_p),
                          yd.ctypes.data_as(capi.c_double_p), nslices, len(xd),
                          sli.ctypes.data_as(capi.c_double_p))
    return sli


def find_pol(lib, intv, damping=2, thresh_int=0.8, thresh_ext=0.2, max_deg=3000,
             deg=0):
    """pEVSL_SetPolDef + the driver's overrides + pEVSL_FindPol.  Returns the
    Polparams (caller frees with lib.pEVSL_FreePol) and the return code."""
    pol = capi.Polparams()
    lib.pEVSL_SetPolDef(C.byref(pol))
    pol.damping = damping
    pol.thresh_int = thresh_int
    pol.thresh_ext = thresh_ext
    pol.max_deg = max_deg
    pol.deg = deg
    iv = np.ascontiguousarray(intv, dtype=np.float64)
    rc = lib.pEVSL_FindPol(iv.ctypes.data_as(capi.c_double_p), C.byref(pol))
    return pol, rc


def pol_mu(pol):
    return np.ctypeslib.as_array(pol.mu, shape=(pol.deg + 1,)).copy()


def rand_vinit(lib, n):
    """vinit by pEVSL_ParvecRand (libc rand(), SRC/parvec.c:65-72)."""
    v = np.empty(n)
    pv = capi.Parvec()
    lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(pv), v.ctypes.data_as(capi.c_double_p))
    lib.pEVSL_ParvecRand(C.byref(pv))
    return v


def c_stdout():
    """libc's stdout FILE* (for the fstats argument)."""
    libc = C.CDLL(None)
    return C.c_void_p.in_dll(libc, "stdout")


def c_stderr():
    """libc's stderr FILE* (unbuffered; for the fstats argument)."""
    libc = C.CDLL(None)
    return C.c_void_p.in_dll(libc, "stderr")


def cheblan_batch(inst, intvs, maxits, tol, vinit, pols, keep_vectors=False,
                  verbose=False, fstats=None):
    """pEVSL_ChebLanNrBatch over len(pols) slices.  Returns a list of dicts
    with lam (unsorted, reference order), res, vecs (or None) and info.
    The solver's per-test lines go to stdout with verbose, or to the FILE*
    `fstats`."""
    lib = inst.lib
    ns = len(pols)
    iv = np.ascontiguousarray(np.asarray(intvs, dtype=np.float64).reshape(ns, 4))
    mi = np.ascontiguousarray(maxits, dtype=np.int32)
    n = len(vinit)
    vin = np.ascontiguousarray(vinit, dtype=np.float64)
    pv = capi.Parvec()
    lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(pv), vin.ctypes.data_as(capi.c_double_p))
    parr = (capi.Polparams * ns)(*pols)
    nev = np.zeros(ns, dtype=np.int32)
    lamo = (capi.c_double_p * ns)()
    reso = (capi.c_double_p * ns)()
    Wo = (C.POINTER(capi.Parvecs) * ns)()
    lib.pEVSL_SetKeepVectors(inst.p, 1 if keep_vectors else 0)
    rc = lib.pEVSL_ChebLanNrBatch(inst.p, ns, iv.ctypes.data_as(capi.c_double_p),
                                  mi.ctypes.data_as(capi.c_int_p), tol, C.byref(pv),
                                  parr, nev.ctypes.data_as(capi.c_int_p), lamo, Wo,
                                  reso, fstats if fstats is not None
                                  else (c_stdout() if verbose else None))
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    out = []
    for s in range(ns):
        k = int(nev[s])
        lam = np.ctypeslib.as_array(lamo[s], shape=(max(k, 1),))[:k].copy() if lamo[s] else np.zeros(0)
        res = np.ctypeslib.as_array(reso[s], shape=(max(k, 1),))[:k].copy() if reso[s] else np.zeros(0)
        vecs = None
        if Wo[s]:
            W = Wo[s].contents
            if k:
                vecs = np.ctypeslib.as_array(W.data, shape=(k, W.ld))[:, :n].copy()
            lib.pEVSL_ParvecsFree(Wo[s])
            libc.free(C.cast(Wo[s], C.c_void_p))
        if lamo[s]:
            libc.free(C.cast(lamo[s], C.c_void_p))
        if reso[s]:
            libc.free(C.cast(reso[s], C.c_void_p))
        out.append({"lam": lam, "res": res, "vecs": vecs, "info": inst.lan_info(s)})
    return rc, out


def irregular_slicing_run(n, nslices=16, seed=4, frac=0.05, center=0.5, nvec=128, Mdeg=300,
                          prec=capi.PEVSL_PREC_FP64, basis_prec=capi.PEVSL_PREC_FP64,
                          batch=True, wave=None, slices_only=None, log=None, verbose=False,
                          tune=True, block_rowkernel=-1, mlp=None, judge_tol=0.15):
    """The irregular-sparsity workload: powerlaw_sym(n) (mean row length 16
    before symmetrization, max 2000, exponent 2.2), bounds by
    pEVSL_LanTrbounds (lanm 50, maxit 200, tol 1e-8, as TESTS/MM/MMPLanN.c:111),
    interval = a `frac` share of [lmin, lmax] centred at the `center` share
    (0.5: the middle of the spectrum), KPM (Mdeg, nvec) slicing
    into `nslices`, FindPol and (batched) ChebLanNr as in lap_slicing_run.
    block_rowkernel selects the block SpMM (pEVSL_SetBlockKernel; an
    autotune run overrides it).  With `mlp` (a weights path) the slices
    are the learned slicer's (cheap KPM + network), judged by the full KPM
    run: when the full-KPM counts of the proposed slices are more than
    judge_tol above their mean, the KPM slices are used instead.  Returns the DOS, slices and per-slice results
    (no closed form: callers check counts against their own reference)."""
    lib = capi.load()
    t0 = time.time()
    ia, ja, av = powerlaw_sym(n, seed=seed)
    M = Matrix(lib, ia, ja, av)
    out = {"n": n, "nnz": M.nnz, "t_gen": time.time() - t0}
    inst = Instance(lib, M, prec=prec, basis_prec=basis_prec, block_rowkernel=block_rowkernel)
    try:
        if tune:
            tr = capi.TuneResult()
            lib.pEVSL_Autotune(inst.p, 10, C.byref(tr))
            out["autotune"] = {"lanes": tr.lanes, "block_width": tr.block_width,
                               "block_kernel": tr.block_rowkernel}
        t1 = time.time()
        rc, lmin, lmax = inst.lan_trbounds(50, 200, 1e-8)
        if rc:
            raise RuntimeError("LanTrbounds error %d" % rc)
        c, h = lmin + center * (lmax - lmin), 0.5 * frac * (lmax - lmin)
        a, b = c - h, c + h
        xintv = np.array([a, b, lmin, lmax])
        rc, mu, ecount = inst.kpmdos(Mdeg, 1, nvec, xintv)
        if rc:
            raise RuntimeError("kpmdos error %d" % rc)
        rc, sli = spslicer_kpm(lib, mu, xintv, nslices, max(int(10 * ecount), 2 * nslices + 1))
        if rc:
            raise RuntimeError("spslicer error %d" % rc)
        if mlp is not None:
            from . import slicer as slc
            net = slc.Mlp.load(mlp, lib)
            try:
                msli, mcount = slc.spslicer_mlp(inst, net, xintv, nslices)
            finally:
                net.free()
            cnt, imb = slc.judge(lib, mu, xintv, n, msli)
            kcnt, kimb = slc.judge(lib, mu, xintv, n, sli)
            out["judge"] = {"counts": cnt.tolist(), "imbalance": imb, "kpm_imbalance": kimb,
                            "ecount_mlp": mcount, "fallback": bool(imb > judge_tol)}
            if imb <= judge_tol:
                sli = msli
        out.update(bounds=[lmin, lmax], intv=[a, b], ecount=ecount, sli=sli, mu=mu,
                   t_dos=time.time() - t1)
        ev_int = int(1 + ecount / float(nslices))
        vinit = rand_vinit(lib, n)
        nev = ev_int + 2
        mlan = min(max(5 * nev, 300), n)
        todo = list(range(nslices)) if slices_only is None else list(slices_only)
        pols = [find_pol(lib, [sli[s], sli[s + 1], lmin, lmax])[0] for s in todo]
        intvs = [[sli[s], sli[s + 1], lmin, lmax] for s in todo]
        if batch:
            w_ = wave or len(todo)
            groups = [list(range(i, min(i + w_, len(todo)))) for i in range(0, len(todo), w_)]
        else:
            groups = [[i] for i in range(len(todo))]
        t3 = time.time()
        res = []
        for g in groups:
            rc, outs = cheblan_batch(inst, [intvs[i] for i in g], [mlan] * len(g), 1e-8, vinit,
                                     [pols[i] for i in g], False, verbose)
            if rc:
                raise RuntimeError("ChebLanNr error %d" % rc)
            for i, o in zip(g, outs):
                res.append({"slice": todo[i], "intv": intvs[i][:2], "deg": pols[i].deg,
                            "nev": len(o["lam"]), "lam": np.sort(o["lam"]), "res": o["res"],
                            "info": o["info"], "mlan": mlan})
                if log:
                    log("slice %d: deg %d its %s nev %d" % (
                        todo[i], pols[i].deg, o["info"]["its"] if o["info"] else "?",
                        len(o["lam"])))
        out["t_lanczos"] = time.time() - t3
        out["slices"] = res
        out["kstats"] = inst.kernel_stats()
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
    finally:
        inst.close()
        M.free()
    out["t_total"] = time.time() - t0
    return out


def lap_slicing_run(nx, ny, nz, a, b, nslices, nvec=60, Mdeg=300, dostype=0,
                    prec=capi.PEVSL_PREC_FP64, batch=False, slices_only=None,
                    keep_vectors=False, rand_mode=capi.PEVSL_RAND_LIBC, seed=0,
                    sli_override=None, block_width=0, log=None, verbose=False,
                    maxit_cap=None, lanes=0, ritz_refine=-1, mixed_tol=None,
                    block_rowkernel=-1, mlp=None, judge=True, judge_tol=0.15, wave=None,
                    basis_prec=capi.PEVSL_PREC_FP64, lanczos_bounds=False, gram_kernel=-1):
    """End-to-end run of the TESTS/Lap/LapPLanN.c flow on the GPU.  Returns a
    dict with the DOS, slices, per-slice filters, eigenvalues and timings.
    gram_kernel selects the Rayleigh-Ritz Gram products (pEVSL_SetGramKernel).

    dostype 0: KPM (Mdeg, nvec) + SpslicerKpm; 1: Lanczos DOS + SpslicerLan;
    3: learned slicer (`mlp`: a slicer.Mlp or a weights path).  With
    `judge`, dostype 3 also runs the full KPM, and when the full-KPM counts of
    the proposed slices are more than `judge_tol` above their mean it falls
    back to SpslicerKpm on the full moments.

    With `batch`, slices are solved `wave` at a time (all at once when None)
    as the columns of one pEVSL_ChebLanNrBatch call, so that the Lanczos
    bases of one wave fit in HBM.  With `lanczos_bounds` the frame
    [lmin, lmax] comes from pEVSL_LanTrbounds (50-vector restart, 1e-8)
    instead of the hand-set Laplacian bounds."""
    lib = capi.load()
    t0 = time.time()
    ia, ja, av = laplacian3d(nx, ny, nz)
    M = Matrix(lib, ia, ja, av)
    n = M.n
    lmin, lmax = 0.0, (8.0 if nz == 1 else 12.0)
    tol = 1e-8
    inst = Instance(lib, M, prec=prec, rand_mode=rand_mode, seed=seed,
                    block_width=block_width, lanes=lanes, ritz_refine=ritz_refine,
                    mixed_tol=mixed_tol, block_rowkernel=block_rowkernel,
                    basis_prec=basis_prec)
    lib.pEVSL_SetGramKernel(inst.p, gram_kernel)
    out = {"n": n, "nnz": M.nnz, "t_setup": time.time() - t0}
    try:
        if lanczos_bounds:
            tb = time.time()
            rc, lmin, lmax = inst.lan_trbounds(50, 3000, 1e-8)
            if rc:
                raise RuntimeError("LanTrbounds error %d" % rc)
            out["bounds"] = {"lmin": lmin, "lmax": lmax, "seconds": time.time() - tb}
        xintv = np.array([a, b, lmin, lmax])
        t1 = time.time()
        if dostype == 0:
            rc, mu, ecount = inst.kpmdos(Mdeg, 1, nvec, xintv)
            if rc:
                raise RuntimeError("kpmdos error %d" % rc)
            npts = int(10 * ecount)
            rc, sli = spslicer_kpm(lib, mu, xintv, nslices, npts)
            if rc:
                raise RuntimeError("spslicer error %d" % rc)
            out["mu"] = mu
        elif dostype == 3:
            from . import slicer as slc
            net = slc.Mlp.load(mlp, lib) if isinstance(mlp, str) else mlp
            try:
                sli, ecount = slc.spslicer_mlp(inst, net, xintv, nslices)
            finally:
                if isinstance(mlp, str):
                    net.free()
            out["t_slicer"] = time.time() - t1
            if judge:
                rc, mu, ecount_full = inst.kpmdos(Mdeg, 1, nvec, xintv)
                if rc:
                    raise RuntimeError("kpmdos error %d" % rc)
                cnt, imb = slc.judge(lib, mu, xintv, n, sli)
                fallback = imb > judge_tol
                out["judge"] = {"counts": cnt.tolist(), "imbalance": imb,
                                "ecount_full": ecount_full, "ecount_mlp": ecount,
                                "fallback": bool(fallback)}
                if fallback:
                    rc, sli = spslicer_kpm(lib, mu, xintv, nslices, int(10 * ecount_full))
                    if rc:
                        raise RuntimeError("spslicer error %d" % rc)
                ecount = ecount_full
                out["mu"] = mu
        else:
            npts = 200
            rc, xd, yd, ecount = inst.landos(nvec, 40, npts, xintv)
            sli = spslicer_lan(lib, xd, yd, nslices)
        out["t_dos"] = time.time() - t1
        if sli_override is not None:
            sli = np.asarray(sli_override, dtype=np.float64)
        out["ecount"] = ecount
        out["sli"] = sli
        ev_int = int(1 + ecount / float(nslices))
        vinit = rand_vinit(lib, n)
        todo = list(range(nslices)) if slices_only is None else list(slices_only)
        nev = ev_int + 2
        mlan = min(max(5 * nev, 300), n)
        if maxit_cap:
            mlan = min(mlan, maxit_cap)
        pols, intvs = [], []
        t2 = time.time()
        for sl in todo:
            iv = [sli[sl], sli[sl + 1], lmin, lmax]
            pol, rc = find_pol(lib, iv)
            pols.append(pol)
            intvs.append(iv)
        out["t_findpol"] = time.time() - t2
        res_slices = []
        t3 = time.time()
        if batch:
            w_ = wave or len(todo)
            groups = [list(range(i, min(i + w_, len(todo)))) for i in range(0, len(todo), w_)]
        else:
            groups = [[i] for i in range(len(todo))]
        for g in groups:
            rc, outs = cheblan_batch(inst, [intvs[i] for i in g], [mlan] * len(g), tol,
                                     vinit, [pols[i] for i in g], keep_vectors,
                                     verbose)
            if rc:
                raise RuntimeError("ChebLanNr error %d" % rc)
            for i, o in zip(g, outs):
                sl = todo[i]
                lam_sorted = np.sort(o["lam"])
                ex = exact_eig_lap3(nx, ny, nz, sli[sl], sli[sl + 1])
                d = {"slice": sl, "intv": [float(sli[sl]), float(sli[sl + 1])],
                     "deg": pols[i].deg, "type": pols[i].type, "bar": pols[i].bar,
                     "gam": pols[i].gam, "mlan": mlan, "nev": len(o["lam"]),
                     "nev_exact": len(ex), "lam": lam_sorted, "lam_raw": o["lam"],
                     "res": o["res"], "exact": ex, "info": o["info"],
                     "vecs": o["vecs"]}
                if len(ex) == len(lam_sorted) and len(ex):
                    d["max_rel_err"] = float(np.max(np.abs(lam_sorted - ex) / np.abs(ex)))
                res_slices.append(d)
                if log:
                    log("slice %d: deg %d its %s nev %d exact %d" % (
                        sl, pols[i].deg, o["info"]["its"] if o["info"] else "?",
                        len(o["lam"]), len(ex)))
        out["t_lanczos"] = time.time() - t3
        out["slices"] = res_slices
        out["kstats"] = inst.kernel_stats()
        for p_ in pols:
            lib.pEVSL_FreePol(C.byref(p_))
    finally:
        inst.close()
        M.free()
    out["t_total"] = time.time() - t0
    return out


SVFUNC = C.CFUNCTYPE(None, capi.c_double_p, capi.c_double_p, C.c_void_p)


def chain_laplacian(n):
    """tridiag(-1, 4, -1) of order n: the B of TESTS/Gen_Lap/LapPLanN.c
    (ParcsrLaplace(n, 1, 1): the stencil of TESTS/COMMON/lapl.c:4-70 on an
    n x 1 x 1 grid, whose diagonal is 4 since nz = 1); spectrum in (2, 6)."""
    return laplacian3d(n, 1, 1)


class HostSolve:
    """A B-solve (SVFunc) on the host, kept alive with its callback: the
    pattern of a caller-supplied direct solver (the reference's tests use
    MUMPS).  solve(b) -> x with B x = b."""

    def __init__(self, solve):
        self._solve = solve

        def cb(pb, px, data):
            n = self.n
            b = np.ctypeslib.as_array(pb, shape=(n,))
            x = np.ctypeslib.as_array(px, shape=(n,))
            x[:] = self._solve(b.copy())
        self.n = 0
        self.func = SVFUNC(cb)

    def ptr(self):
        return C.cast(self.func, C.c_void_p)


def gen_lap_run(nx, ny, nz, a, b, nslices, bsol="lspol", nvec=60, Mdeg=50, tol=1e-8,
                lspol_deg=50, lspol_tol=1e-12, chebiter_deg=40, verbose=False, log=None):
    """End-to-end run of TESTS/Gen_Lap/LapPLanN.c (KPM branch): A = 3-D
    Laplacian, B = chain_laplacian(n) (spectrum in (2, 6)), L^T solve by
    pEVSL_SetupLSPolSqrt(50, 1e-6, 2, 6) as the reference, B-solve by the
    library's LSPol approximation of B^{-1} on [2, 6] ("lspol", on the
    device), Chebyshev iteration on [2, 6] ("chebiter", chebiter_deg steps,
    on the device) or an exact host Cholesky solve ("host", the place of the
    reference's MUMPS solve); bounds by LanTrbounds(50, 200, 1e-8), KPM
    (Mdeg 50, nvec 60), SpslicerKpm(npts = 10 ecount), FindPol (damping 2,
    thresholds 0.6/0.15, max_deg 3000), ChebLanNr per slice with
    mlan = min(max(5 nev, 300), n)."""
    import scipy.linalg as sla
    lib = capi.load()
    n = nx * ny * nz
    MA = Matrix(lib, *laplacian3d(nx, ny, nz))
    MB = Matrix(lib, *chain_laplacian(n))
    inst = Instance(lib, MA)
    keep = []
    lt = C.c_void_p()
    bdata = C.c_void_p()
    cdata = C.c_void_p()
    out = {"n": n}
    try:
        lib.pEVSL_SetBParcsr(inst.p, C.byref(MB.A))
        if bsol == "lspol":
            lib.pEVSL_SetupLSPolRec(lspol_deg, lspol_tol, 2.0, 6.0, C.byref(MB.A),
                                    C.byref(bdata))
            lib.pEVSL_SetBSol(inst.p, C.cast(lib.pEVSL_LSPolSol, C.c_void_p), bdata)
        elif bsol == "chebiter":
            lib.pEVSL_ChebIterSetup(2.0, 6.0, chebiter_deg, C.byref(MB.A), C.byref(cdata))
            lib.pEVSL_SetBSol(inst.p, C.cast(lib.pEVSL_ChebIterSolv2, C.c_void_p), cdata)
        else:
            import scipy.sparse as sp
            ia, ja, av = chain_laplacian(n)
            Bd = sp.csr_matrix((av, ja, ia), shape=(n, n)).toarray()
            cho = sla.cho_factor(Bd)
            hs = HostSolve(lambda r: sla.cho_solve(cho, r))
            hs.n = n
            keep.append(hs)
            lib.pEVSL_SetBSol(inst.p, hs.ptr(), None)
        lib.pEVSL_SetupLSPolSqrt(50, 1e-6, 2.0, 6.0, C.byref(MB.A), C.byref(lt))
        lib.pEVSL_SetLTSol(inst.p, C.cast(lib.pEVSL_LSPolSol, C.c_void_p), lt)
        lib.pEVSL_SetGenEig(inst.p)
        vinit = rand_vinit(lib, n)
        t0 = time.time()
        rc, lmin, lmax = inst.lan_trbounds(50, 200, 1e-8, vinit, 1)
        if rc:
            raise RuntimeError("LanTrbounds error %d" % rc)
        xintv = np.array([a, b, lmin, lmax])
        rc, mu, ecount = inst.kpmdos(Mdeg, 1, nvec, xintv)
        if rc:
            raise RuntimeError("kpmdos error %d" % rc)
        rc, sli = spslicer_kpm(lib, mu, xintv, nslices, int(10 * ecount))
        if rc:
            raise RuntimeError("spslicer error %d" % rc)
        out.update(bounds=[lmin, lmax], ecount=ecount, sli=sli, mu=mu, t_dos=time.time() - t0)
        nev = int(1 + ecount / float(nslices)) + 2
        mlan = min(max(5 * nev, 300), n)
        res = []
        for s in range(nslices):
            iv = [sli[s], sli[s + 1], lmin, lmax]
            pol, rcp = find_pol(lib, iv, damping=2, thresh_int=0.6, thresh_ext=0.15,
                                max_deg=3000)
            rc, o = cheblan_batch(inst, [iv], [mlan], tol, vinit, [pol], False, verbose)
            if rc:
                raise RuntimeError("ChebLanNr error %d" % rc)
            o = o[0]
            res.append({"slice": s, "intv": iv[:2], "deg": pol.deg, "nev": len(o["lam"]),
                        "lam": np.sort(o["lam"]), "res": o["res"], "info": o["info"]})
            if log:
                log("slice %d: deg %d its %d nev %d" % (s, pol.deg, o["info"]["its"],
                                                       len(o["lam"])))
            lib.pEVSL_FreePol(C.byref(pol))
        out["slices"] = res
        out["t_total"] = time.time() - t0
    finally:
        inst.close()
        if bdata:
            lib.pEVSL_LSPolFree(bdata)
        if cdata:
            lib.pEVSL_ChebIterFree(cdata)
        if lt:
            lib.pEVSL_LSPolFree(lt)
        MA.free()
        MB.free()
    return out
