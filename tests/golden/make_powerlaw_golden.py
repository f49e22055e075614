"""Writes powerlaw_n<n>.json: eigenvalues of slices of the irregular matrix
recipe, computed by the oracle (the FP64 C restatement of SRC/cheblanNr.c,
SRC/chebpol.c and SRC/lantrbnd.c).

Matrix: driver.powerlaw_sym(n, seed=4), the generator of the
irregular-sparsity workload (power-law row lengths, mean 16, up to 2,000
entries, uniform random columns); n = 2,000,000 is the workload's own size.
Bounds: the oracle's thick-restart Lanczos (lanm 50, maxit 200, tol 1e-8,
start vector srand(1)).  Interval: the middle 5% of [lmin, lmax] in equal
slices; the chosen slices are solved by the oracle's ChebLanNr (FindPol
defaults of TESTS/Lap/LapPLanN.c:217-221, maxit 300, tol 1e-8, start vector
srand(0)).

The GPU tests (tests/test_gpu_powerlaw.py) solve the same slices and compare
counts and eigenvalues.  Run from the repository root:

    python tests/golden/make_powerlaw_golden.py 200000 8 0,1,2,3,4,5,6,7   # ~5 min, 8 threads
    python tests/golden/make_powerlaw_golden.py 2000000 16 7,8            # ~2 h per slice, 8 threads
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from pevsl_amd import driver  # noqa: E402

SEED = 4


def main():
    N, NSLICES = int(sys.argv[1]), int(sys.argv[2])
    SOLVE = [int(x) for x in sys.argv[3].split(",")]
    t0 = time.time()
    ia, ja, a = driver.powerlaw_sym(N, seed=SEED)
    A = orc.Csr(ia, ja, a)
    orc.srand(1)
    rc, lmin, lmax, its = orc.lan_trbounds(A, 50, 200, 1e-8, orc.rand_vec(N))
    assert rc == 0
    print("bounds [%.12g, %.12g] in %d steps, %.0f s" % (lmin, lmax, its, time.time() - t0),
          flush=True)
    c, h = lmin + 0.5 * (lmax - lmin), 0.025 * (lmax - lmin)
    edges = np.linspace(c - h, c + h, NSLICES + 1)
    out = {"recipe": "driver.powerlaw_sym(%d, seed=%d)" % (N, SEED), "n": N,
           "nnz": int(ia[-1]), "lmin": lmin, "lmax": lmax, "bounds_its": its,
           "interval": [edges[0], edges[-1]], "nslices": NSLICES, "slices": []}
    path = os.path.join(ROOT, "tests", "golden", "powerlaw_n%d.json" % N)
    for s in SOLVE:
        intv = [float(edges[s]), float(edges[s + 1]), lmin, lmax]
        pol = orc.Pol(intv)
        orc.srand(0)
        vinit = orc.rand_vec(N)
        t1 = time.time()
        r = orc.cheblan_nr(A, intv, 300, 1e-8, vinit, pol)
        assert r["rc"] == 0 and r["its"] < 300, (r["rc"], r["its"])
        lam = np.sort(r["lam"])
        print("slice %d [%.10g, %.10g]: degree %d, %d eigenvalues, %d steps, %.0f s" % (
            s, intv[0], intv[1], pol.deg, len(lam), r["its"], time.time() - t1), flush=True)
        out["slices"].append({"slice": s, "intv": intv[:2], "deg": pol.deg, "its": r["its"],
                              "nbreak": r["nbreak"], "lam": lam.tolist(),
                              "max_res": float(np.max(r["res"])) if len(lam) else None})
        # written after every slice: a long run keeps the slices it finished
        with open(path + ".tmp", "w") as f:
            json.dump(out, f, indent=1)
        os.replace(path + ".tmp", path)
        print("wrote %s (%d slices) in %.0f s" % (path, len(out["slices"]), time.time() - t0),
              flush=True)


if __name__ == "__main__":
    main()
