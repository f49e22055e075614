"""Learned spectrum slicer: Python side of the C API in include/pevsl_amd.h
(pEVSL_Mlp*, pEVSL_SpslicerMlp, pEVSL_SlicerJudge ...).

The network and all its arithmetic (forward, backward, Adam) run in
libpevsl_amd.so on the GPU; this module wraps the handles and holds the
training recipe:

* matrices: Dirichlet Laplacians (3-D grids 12..64 per side, 2-D grids
  40..400 per side) plus a diagonal potential w * uniform(0, 1) per row, with
  w = 0 for about a third of them and uniform(0, 2) otherwise; frame
  [lmin, lmax] = Gershgorin bounds [0, 12 + w] (3-D) or [0, 8 + w] (2-D);
* per matrix: one full KPM run (pEVSL_Kpmdos, degree 300, Jackson, 64
  vectors) gives the targets, one cheap run (degree 20, 8 vectors) the
  inputs;
* per matrix, `nint` random intervals with widths log-uniform between 1% and
  30% of the frame; target = full-KPM bin masses of the interval
  (normalised) and log(fraction of the spectrum inside);
* network 138 -> 3 x 256 (GELU) -> 513, trained with Adam on the GPU.

An irregular family can be added (--family powerlaw): driver.powerlaw_sym
matrices with n log-uniform in [1e5, 5e5], framed by pEVSL_LanTrbounds, half
of their intervals a 2-10% share centred in the middle of the frame.

Command line (writes the weights and a JSON record):
    python -m pevsl_amd.slicer --out pevsl_amd/weights/slicer_lap.bin
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import capi, driver

NBINS = capi.PEVSL_SLICER_NBINS
NIN = capi.PEVSL_SLICER_NIN
NOUT = NBINS + 1
_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(_HERE, "weights", "slicer_lap.bin")

_fp = capi.c_float_p
_dp = capi.c_double_p


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t=_fp):
    return None if a is None else a.ctypes.data_as(t)


class Mlp:
    """Handle of a pEVSL_Mlp (device-resident network)."""

    def __init__(self, lib=None, nin=NIN, nhidden=256, nlayers=3, nout=NOUT, nsoft=NBINS,
                 seed=0, handle=None):
        self.lib = lib or capi.load()
        self.h = C.c_void_p()
        if handle is not None:
            self.h = handle
        else:
            rc = self.lib.pEVSL_MlpCreate(nin, nhidden, nlayers, nout, nsoft, seed,
                                          C.byref(self.h))
            if rc:
                raise RuntimeError("pEVSL_MlpCreate failed: %d" % rc)
        vals = [C.c_int() for _ in range(5)]
        nw = C.c_size_t()
        self.lib.pEVSL_MlpShape(self.h, *[C.byref(v) for v in vals], C.byref(nw))
        self.nin, self.nhidden, self.nlayers, self.nout, self.nsoft = [v.value for v in vals]
        self.nweights = nw.value

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, lib=None):
        lib = lib or capi.load()
        h = C.c_void_p()
        rc = lib.pEVSL_MlpLoad(path.encode(), C.byref(h))
        if rc:
            raise RuntimeError("pEVSL_MlpLoad(%s) failed: %d" % (path, rc))
        return cls(lib, handle=h)

    def save(self, path):
        rc = self.lib.pEVSL_MlpSave(self.h, path.encode())
        if rc:
            raise RuntimeError("pEVSL_MlpSave(%s) failed: %d" % (path, rc))

    def free(self):
        if self.h:
            self.lib.pEVSL_MlpFree(self.h)
            self.h = C.c_void_p()

    @property
    def shape(self):
        return (self.nin, self.nhidden, self.nlayers, self.nout)

    def get_weights(self):
        w = np.empty(self.nweights, np.float32)
        if self.lib.pEVSL_MlpGetWeights(self.h, _ptr(w), self.nweights):
            raise RuntimeError("pEVSL_MlpGetWeights failed")
        return w

    def set_weights(self, w):
        w = _f32(w)
        if self.lib.pEVSL_MlpSetWeights(self.h, _ptr(w), w.size):
            raise RuntimeError("pEVSL_MlpSetWeights: size %d != %d" % (w.size, self.nweights))

    def forward(self, X, Q=None):
        X = _f32(X).reshape(-1, self.nin)
        Qa = None if Q is None else _f32(Q).reshape(-1, self.nout)
        Y = np.empty((X.shape[0], self.nout), np.float32)
        rc = self.lib.pEVSL_MlpForward(self.h, _ptr(X), _ptr(Qa), X.shape[0], _ptr(Y))
        if rc:
            raise RuntimeError("pEVSL_MlpForward failed: %d" % rc)
        return Y

    def loss_grad(self, X, Q, T):
        X = _f32(X).reshape(-1, self.nin)
        Qa = None if Q is None else _f32(Q)
        T = _f32(T)
        g = np.empty(self.nweights, np.float32)
        loss = C.c_float()
        rc = self.lib.pEVSL_MlpLossGrad(self.h, _ptr(X), _ptr(Qa), _ptr(T), X.shape[0],
                                        C.byref(loss), _ptr(g))
        if rc:
            raise RuntimeError("pEVSL_MlpLossGrad failed: %d" % rc)
        return loss.value, g

    def train_step(self, X, Q, T, lr):
        X = _f32(X).reshape(-1, self.nin)
        Qa = None if Q is None else _f32(Q)
        T = _f32(T)
        loss = C.c_float()
        rc = self.lib.pEVSL_MlpTrainStep(self.h, _ptr(X), _ptr(Qa), _ptr(T), X.shape[0],
                                         lr, C.byref(loss))
        if rc:
            raise RuntimeError("pEVSL_MlpTrainStep failed: %d" % rc)
        return loss.value

    def train(self, X, Q, T, batch=256, epochs=100, lr=1e-3, seed=0, verbose=0):
        X = _f32(X)
        Qa = None if Q is None else _f32(Q)
        T = _f32(T)
        losses = np.zeros(epochs, np.float32)
        rc = self.lib.pEVSL_MlpTrain(self.h, _ptr(X), _ptr(Qa), _ptr(T), X.shape[0], batch,
                                     epochs, lr, seed, _ptr(losses), verbose)
        if rc:
            raise RuntimeError("pEVSL_MlpTrain failed: %d" % rc)
        return losses


# ------------------------------------------------------------ slicer calls
def bin_masses(lib, mu, intv, nbins=NBINS):
    mu = np.ascontiguousarray(mu, np.float64)
    iv = np.ascontiguousarray(intv, np.float64)
    m = np.empty(nbins)
    fr = C.c_double()
    rc = lib.pEVSL_KpmBinMasses(_ptr(mu, _dp), len(mu) - 1, _ptr(iv, _dp), nbins,
                                _ptr(m, _dp), C.byref(fr))
    if rc:
        raise RuntimeError("pEVSL_KpmBinMasses failed: %d" % rc)
    return m, fr.value


def features_from_mu(lib, mu, n, intv):
    mu = np.ascontiguousarray(mu, np.float64)
    iv = np.ascontiguousarray(intv, np.float64)
    f = np.empty(NIN, np.float32)
    q = np.empty(NOUT, np.float32)
    rc = lib.pEVSL_SlicerFeaturesFromMu(_ptr(mu, _dp), len(mu) - 1, n, _ptr(iv, _dp),
                                        _ptr(f), _ptr(q))
    if rc:
        raise RuntimeError("pEVSL_SlicerFeaturesFromMu failed: %d" % rc)
    return f, q


def features(inst, intv):
    iv = np.ascontiguousarray(intv, np.float64)
    f = np.empty(NIN, np.float32)
    q = np.empty(NOUT, np.float32)
    rc = inst.lib.pEVSL_SlicerFeatures(inst.p, _ptr(iv, _dp), _ptr(f), _ptr(q))
    if rc:
        raise RuntimeError("pEVSL_SlicerFeatures failed: %d" % rc)
    return f, q


def cuts_from_bins(lib, masses, a, b, nslices):
    m = np.ascontiguousarray(masses, np.float64)
    iv = np.array([a, b], np.float64)
    sli = np.empty(nslices + 1)
    rc = lib.pEVSL_SlicerFromBins(_ptr(m, _dp), len(m), _ptr(iv, _dp), nslices, _ptr(sli, _dp))
    if rc < 0:
        raise RuntimeError("pEVSL_SlicerFromBins failed: %d" % rc)
    return sli


def spslicer_mlp(inst, mlp, intv, nslices):
    """Proposed slices and eigenvalue count of intv[0:2] (one cheap KPM run
    on the GPU plus one network evaluation)."""
    iv = np.ascontiguousarray(intv, np.float64)
    sli = np.empty(nslices + 1)
    ecnt = C.c_double()
    rc = inst.lib.pEVSL_SpslicerMlp(inst.p, mlp.h, _ptr(iv, _dp), nslices, _ptr(sli, _dp),
                                    C.byref(ecnt))
    if rc:
        raise RuntimeError("pEVSL_SpslicerMlp failed: %d" % rc)
    return sli, ecnt.value


def judge(lib, mu, intv, n, sli):
    """Full-KPM counts per slice and imbalance (max / mean - 1)."""
    mu = np.ascontiguousarray(mu, np.float64)
    iv = np.ascontiguousarray(intv, np.float64)
    s = np.ascontiguousarray(sli, np.float64)
    cnt = np.empty(len(s) - 1)
    imb = C.c_double()
    rc = lib.pEVSL_SlicerJudge(_ptr(mu, _dp), len(mu) - 1, _ptr(iv, _dp), float(n),
                               len(s) - 1, _ptr(s, _dp), _ptr(cnt, _dp), C.byref(imb))
    if rc:
        raise RuntimeError("pEVSL_SlicerJudge failed: %d" % rc)
    return cnt, imb.value


def kpm_slices(lib, mu, intv, nslices, ecnt):
    """The method without the network: pEVSL_SpslicerKpm on given moments."""
    rc, sli = driver.spslicer_kpm(lib, mu, intv, nslices, max(int(10 * ecnt), 2 * nslices + 1))
    if rc:
        raise RuntimeError("pEVSL_SpslicerKpm failed: %d" % rc)
    return sli


# --------------------------------------------------------------- matrices
def lap_potential(nx, ny, nz, w, seed):
    """Laplacian plus diagonal potential w * uniform(0, 1); returns CSR and
    the Gershgorin frame."""
    ia, ja, a = driver.laplacian3d(nx, ny, nz)
    n = nx * ny * nz
    if w > 0:
        rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(ia))
        dpos = np.nonzero(ja == rows)[0]
        a = a.copy()
        a[dpos] += w * np.random.default_rng(seed).random(n)
    top = (12.0 if nz > 1 else 8.0) + w
    return ia, ja, a, (0.0, top)


def draw_matrix(rng, grid=None):
    """Training draw: 3-D sides 12-64 (80%) or 2-D 40-400; grid = (lo, hi):
    3-D sides drawn in [lo, hi] only (the evaluation range)."""
    if grid is not None:
        nx, ny, nz = (int(v) for v in rng.integers(grid[0], grid[1] + 1, 3))
        w = 0.0 if rng.random() < 0.35 else float(rng.uniform(0.0, 2.0))
        return nx, ny, nz, w
    if rng.random() < 0.8:
        nx, ny, nz = (int(v) for v in rng.integers(12, 65, 3))
    else:
        nx, ny = (int(v) for v in rng.integers(40, 401, 2))
        nz = 1
    w = 0.0 if rng.random() < 0.35 else float(rng.uniform(0.0, 2.0))
    return nx, ny, nz, w


def draw_interval(rng, lmin, lmax):
    L = lmax - lmin
    wd = L * math.exp(rng.uniform(math.log(0.01), math.log(0.3)))
    a = lmin + rng.uniform(0.0, L - wd)
    return a, a + wd


def draw_interval_irregular(rng, lmin, lmax):
    """Half the intervals of the irregular family as draw_interval, half a
    2-10% share of the frame centred in its middle 40% (the regime of the
    irregular workload: a 5% share at the centre)."""
    if rng.random() < 0.5:
        return draw_interval(rng, lmin, lmax)
    L = lmax - lmin
    c = lmin + L * rng.uniform(0.3, 0.7)
    h = 0.5 * L * rng.uniform(0.02, 0.1)
    return c - h, c + h


def gen_moments(lib, nmat, seed, log=print, full=(300, 64), cheap=None, family="lap",
                nmin=100000, nmax=500000, grid=None):
    """Per matrix: full and cheap KPM moments from this library's Kpmdos.
    family "lap": Laplacians with potentials framed by Gershgorin bounds;
    "powerlaw": driver.powerlaw_sym matrices with n log-uniform in
    [nmin, nmax], framed by pEVSL_LanTrbounds(50, 200, 1e-8) (marked w = -1)."""
    cheap = cheap or (capi.PEVSL_SLICER_MDEG, capi.PEVSL_SLICER_NVEC)
    rng = np.random.default_rng(seed)
    mats = []
    t0 = time.time()
    for j in range(nmat):
        if family == "powerlaw":
            n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
            ia, ja, a = driver.powerlaw_sym(n, seed=seed * 100003 + j)
            nx, ny, nz, w = n, 1, 1, -1.0
        else:
            nx, ny, nz, w = draw_matrix(rng, grid)
            ia, ja, a, (lmin, lmax) = lap_potential(nx, ny, nz, w, seed * 100003 + j)
        M = driver.Matrix(lib, ia, ja, a)
        inst = driver.Instance(lib, M)
        try:
            if family == "powerlaw":
                rc, lmin, lmax = inst.lan_trbounds(50, 200, 1e-8)
                if rc:
                    raise RuntimeError("LanTrbounds failed")
            frame = [lmin, lmax, lmin, lmax]
            rc, mu_full, _ = inst.kpmdos(full[0], 1, full[1], frame)
            rc2, mu_cheap, _ = inst.kpmdos(cheap[0], 1, cheap[1], frame)
            if rc or rc2:
                raise RuntimeError("Kpmdos failed")
        finally:
            inst.close()
            M.free()
        mats.append({"dims": (nx, ny, nz), "w": w, "frame": (lmin, lmax), "n": nx * ny * nz,
                     "mu_full": mu_full, "mu_cheap": mu_cheap})
        if (j + 1) % (10 if family == "powerlaw" or grid else 100) == 0 or j + 1 == nmat:
            log("moments: %d/%d matrices, %.1f s" % (j + 1, nmat, time.time() - t0))
            sys.stdout.flush()
    return mats


def save_moments(path, mats):
    np.savez(path, dims=np.array([m["dims"] for m in mats]), w=np.array([m["w"] for m in mats]),
             frame=np.array([m["frame"] for m in mats]), n=np.array([m["n"] for m in mats]),
             mu_full=np.array([m["mu_full"] for m in mats]),
             mu_cheap=np.array([m["mu_cheap"] for m in mats]))


def load_moments(path):
    z = np.load(path)
    return [{"dims": tuple(int(v) for v in z["dims"][i]), "w": float(z["w"][i]),
             "frame": tuple(z["frame"][i]), "n": int(z["n"][i]), "mu_full": z["mu_full"][i],
             "mu_cheap": z["mu_cheap"][i]} for i in range(len(z["n"]))]


def make_pairs(lib, mats, nint, seed, log=print, nint_irregular=None):
    """Training pairs (features, prior, target) for `nint` random intervals
    per matrix (`nint_irregular` for the irregular family); targets are
    full-KPM bin masses (see module doc)."""
    rng = np.random.default_rng(seed + 7)
    X, Q, T, meta = [], [], [], []
    for j, m in enumerate(mats):
        lmin, lmax = m["frame"]
        for _ in range(nint_irregular or nint if m["w"] < 0 else nint):
            lo, hi = (draw_interval_irregular if m["w"] < 0 else draw_interval)(rng, lmin, lmax)
            iv = [lo, hi, lmin, lmax]
            ms, fr = bin_masses(lib, m["mu_full"], iv)
            ms = np.maximum(ms, 0.0)
            if not (fr > 1e-7 and ms.sum() > 0):
                continue
            f, q = features_from_mu(lib, m["mu_cheap"], m["n"], iv)
            X.append(f)
            Q.append(q)
            T.append(np.concatenate([ms / ms.sum(), [math.log(fr)]]))
            meta.append((j, m["dims"][0], m["dims"][1], m["dims"][2], m["w"], lo, hi, lmin,
                         lmax))
        if log and ((j + 1) % 1000 == 0 or j + 1 == len(mats)):
            log("pairs: %d/%d matrices, %d samples" % (j + 1, len(mats), len(X)))
            sys.stdout.flush()
    return (np.array(X, np.float32), np.array(Q, np.float32), np.array(T, np.float32),
            np.array(meta))


def evaluate(lib, mlp, X, Q, T, meta, nslices=10):
    """Slice quality of the network vs the cheap KPM alone, both judged by
    the full-KPM targets: imbalance of full-KPM counts and error of the
    count estimate."""
    Y = mlp.forward(X, Q)
    res = {"net_imb": [], "cheap_imb": [], "net_cnt_err": [], "cheap_cnt_err": []}
    ncount = []
    for i in range(len(X)):
        ncount.append(math.exp(T[i, NBINS]) * meta[i][1] * meta[i][2] * meta[i][3])
        a, b = float(meta[i][5]), float(meta[i][6])
        tmass = T[i, :NBINS].astype(np.float64)
        ctm = np.concatenate([[0.0], np.cumsum(tmass)])
        edges = a + (b - a) * np.arange(NBINS + 1) / NBINS

        def imb(sli):
            c = np.diff(np.interp(sli, edges, ctm))
            return c.max() / c.mean() - 1.0

        res["net_imb"].append(imb(cuts_from_bins(lib, Y[i, :NBINS], a, b, nslices)))
        res["cheap_imb"].append(imb(cuts_from_bins(lib, np.exp(Q[i, :NBINS]), a, b, nslices)))
        res["net_cnt_err"].append(abs(math.exp(Y[i, NBINS] - T[i, NBINS]) - 1.0))
        res["cheap_cnt_err"].append(abs(math.exp(Q[i, NBINS] - T[i, NBINS]) - 1.0))
    out = {}
    ncount = np.array(ncount)
    for name, sel in (("all", ncount >= 0), ("count_ge_200", ncount >= 200)):
        o = {"samples": int(sel.sum())}
        for k, v in res.items():
            v = np.array(v)[sel]
            if len(v):
                o[k] = {"median": float(np.median(v)), "p90": float(np.percentile(v, 90)),
                        "max": float(v.max())}
        out[name] = o
    return out


def eval_only(lib, args):
    """Slice quality of saved weights on matrices drawn outside (or inside)
    the training range, judged by full KPM, against the cheap KPM alone."""
    grid = None
    if args.family == "lap" and (args.grid_min or args.grid_max):
        grid = (args.grid_min or 12, args.grid_max or args.grid_min)
    t0 = time.time()
    mats = gen_moments(lib, args.nmat, args.seed, family=args.family, nmin=args.nmin,
                       nmax=args.nmax, grid=grid)
    X, Q, T, meta = make_pairs(lib, mats, args.nint, args.seed)
    mlp = Mlp.load(args.eval, lib)
    try:
        slices = (10, 16) if args.family == "powerlaw" else (10, 20)
        rec = {"weights": os.path.basename(args.eval), "family": args.family,
               "matrices": len(mats), "samples": int(len(X)), "seed": args.seed,
               "sizes": {"n_min": int(min(m["n"] for m in mats)),
                         "n_max": int(max(m["n"] for m in mats))},
               "seconds": round(time.time() - t0, 1)}
        if grid:
            rec["grid"] = list(grid)
        for ns in slices:
            rec["eval_%d_slices" % ns] = evaluate(lib, mlp, X, Q, T, meta, ns)
    finally:
        mlp.free()
    if args.eval_out:
        with open(args.eval_out, "w") as f:
            json.dump(rec, f, indent=1)
    print(json.dumps(rec))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--nmat", type=int, default=400)
    ap.add_argument("--nint", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20241)
    ap.add_argument("--holdout", type=float, default=0.1)
    ap.add_argument("--moments", default=None,
                    help="comma-separated npz files of per-matrix moments: read when "
                         "they all exist, else the first is generated and written")
    ap.add_argument("--gen-only", action="store_true",
                    help="only generate and write the moments")
    ap.add_argument("--family", default="lap", choices=("lap", "powerlaw"),
                    help="matrix family of generated moments")
    ap.add_argument("--nint-irregular", type=int, default=None,
                    help="intervals per matrix of the irregular family (default --nint)")
    ap.add_argument("--split", default="tail", choices=("tail", "mod10"),
                    help="held-out matrices: the last --holdout share, or every 10th")
    ap.add_argument("--nmin", type=int, default=100000)
    ap.add_argument("--nmax", type=int, default=500000)
    ap.add_argument("--eval", default=None, metavar="WEIGHTS",
                    help="no training: evaluate these weights on freshly drawn matrices "
                         "(--family, --nmat, --nint, and --grid-min/--grid-max or --nmin/--nmax)")
    ap.add_argument("--grid-min", type=int, default=None,
                    help="with --eval, family lap: 3-D grid sides drawn in [grid-min, grid-max]")
    ap.add_argument("--grid-max", type=int, default=None)
    ap.add_argument("--eval-out", default=None, help="with --eval: write the record here")
    args = ap.parse_args(argv)
    for name in ("nmat", "nint", "epochs", "batch"):
        if getattr(args, name) < 1:
            ap.error("--%s must be at least 1" % name)
    if args.nint_irregular is not None and args.nint_irregular < 1:
        ap.error("--nint-irregular must be at least 1")
    if not 0.0 < args.holdout < 1.0:
        ap.error("--holdout must lie in (0, 1)")
    if not 1 <= args.nmin <= args.nmax:
        ap.error("--nmin and --nmax must satisfy 1 <= nmin <= nmax")
    if args.grid_min is not None or args.grid_max is not None:
        lo = args.grid_min if args.grid_min is not None else 12
        hi = args.grid_max if args.grid_max is not None else lo
        if not 2 <= lo <= hi:
            ap.error("--grid-min and --grid-max must satisfy 2 <= min <= max")
    lib = capi.load()
    t0 = time.time()
    if args.eval:
        return eval_only(lib, args)
    files = args.moments.split(",") if args.moments else []
    if files and all(os.path.exists(f) for f in files) and not args.gen_only:
        mats = [m for f in files for m in load_moments(f)]
    else:
        mats = gen_moments(lib, args.nmat, args.seed, family=args.family, nmin=args.nmin,
                           nmax=args.nmax)
        if files:
            save_moments(files[0], mats)
        if args.gen_only:
            return 0
    X, Q, T, meta = make_pairs(lib, mats, args.nint, args.seed,
                               nint_irregular=args.nint_irregular)
    t_data = time.time() - t0
    nmat = len(mats)
    idx = meta[:, 0].astype(int)
    if args.split == "mod10":
        tr = idx % 10 != 9
    else:
        nval = max(1, int(round(args.holdout * nmat)))
        tr = idx < nmat - nval
    va = ~tr
    mlp = Mlp(lib, seed=args.seed)
    t1 = time.time()
    losses = mlp.train(X[tr], Q[tr], T[tr], args.batch, args.epochs, args.lr, args.seed,
                       verbose=10)
    t_train = time.time() - t1
    print("training: %d samples, %d epochs, loss %.5f -> %.5f, %.1f s" % (
        int(tr.sum()), args.epochs, losses[0], losses[-1], t_train))
    ev = {}
    for ns in (10, 20):
        ev["heldout_%d_slices" % ns] = evaluate(lib, mlp, X[va], Q[va], T[va], meta[va], ns)
    ev["train_10_slices"] = evaluate(lib, mlp, X[tr][:3000], Q[tr][:3000], T[tr][:3000],
                                     meta[tr][:3000], 10)
    irr = meta[:, 4].astype(float) < 0
    if np.any(va & irr):
        for ns in (10, 16):
            ev["heldout_irregular_%d_slices" % ns] = evaluate(
                lib, mlp, X[va & irr], Q[va & irr], T[va & irr], meta[va & irr], ns)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    mlp.save(args.out)
    rec = {"weights": os.path.basename(args.out), "arch": [NIN, 256, 3, NOUT, NBINS],
           "recipe": {k: getattr(args, k) for k in ("nmat", "nint", "nint_irregular", "epochs",
                                                     "batch", "lr", "seed", "holdout", "split",
                                                     "moments")},
           "matrices": nmat, "samples": {"train": int(tr.sum()), "heldout": int(va.sum())},
           "seconds": {"dataset": round(t_data, 1), "train": round(t_train, 1)},
           "loss": {"first_epoch": float(losses[0]), "last_epoch": float(losses[-1]),
                    "every_10": [float(v) for v in losses[::10]]}}
    rec.update(ev)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(ev))
    mlp.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
