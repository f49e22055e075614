"""Benchmark of the spectrum-slicing hot path on one MI355X.

Workload (default): the 100x100x100 7-point Laplacian (n = 1e6, nnz =
6.94e6, int32 CSR), interval [0.6, 1.2] of [0, 12] cut into 10 slices by KPM
(Mdeg 300, nvec 64, Jackson) -- the plain FP64 GPU configuration.  Set-up
(matrix upload, autotune, KPM + slicing + FindPol for all slices) is timed
and reported separately.  One *step* solves one slice: filtered Lanczos with
full reorthogonalization and Ritz/residual checks (pEVSL_ChebLanNrBatch with
one slice, or --batch slices as columns of one block SpMM); step i solves
slice (first + i*batch) mod nslices.  `value` = accepted eigenpairs per
second over the timed steps.  --mode mixed runs the filter and KPM in FP32.

After the timed FP64 steps the same slices are solved once more on the mixed
path (FP32 filter in Clenshaw form, FP32 basis storage, FP64 Lanczos
coefficients / reorthogonalization sums / Rayleigh-Ritz / residuals); the
`mixed` object reports its rate and the ratio to FP64 (--no-mixed skips it).

`roofline` is the dominant kernel of the timed region (largest event time):
the fused filter step or one of the two CGS kernels.  Its `traffic` (HBM
bytes per launch) comes from the hardware-counter summary committed in
profiles/pmc_traffic.json, measured with pevsl_amd/profile_probe.py.

Prints one JSON line (rank 0 of a single process).
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(grid, a, b, nev_slice, its, nev_found, t_budget=25.0):
    """Times the FP64 CPU oracle on a bounded sample of the same slice: a few
    filtered Lanczos steps and CGS passes at k = 512, then projects the full
    slice from the GPU-measured step count `its`:
        T = its * t_step + t_cgs(512)/512 * its^2/2 .
    The oracle's long loops (CSR product, vector updates, CGS) run on
    OMP_NUM_THREADS host threads with results independent of the count;
    `cores` reports the count it ran on."""
    from oracle import oracle as orc
    orc.build() if not os.path.exists(orc.DRIVER) else None
    steps = 2
    t0 = time.time()
    d = orc.run_driver(["-nx", grid, "-ny", grid, "-nz", grid, "-a", a, "-b", b,
                        "-dostype", 2, "-nev", nev_slice, "-maxsteps", steps,
                        "-reorthk", 512], timeout=600)
    wall = time.time() - t0
    s = d["slices"][0]
    # t_lan covers steps+1 filter applications (FILTER_VINIT) and the Ritz
    # stage of a 2-step basis; per Lanczos step:
    t_step = s["t_lan"] / (steps + 1)
    t_cgs = d["t_cgs_pass"]
    T = its * t_step + (t_cgs / 512.0) * its * its / 2.0
    threads = int(d.get("threads", 1))
    return {"value": nev_found / T, "unit": "eigenpairs/s", "cores": threads, "kind": "port",
            "sample": ("oracle (FP64 C, %d OpenMP threads) on the same slice [%.6f, %.6f]: %d "
                       "filtered Lanczos steps (deg %d) + 3 CGS passes at k=512, %.1f s wall; "
                       "projected to the GPU run's %d steps: %.0f s per slice" % (
                           threads, a, b, steps, s["deg"], wall, its, T)),
            "t_step_s": t_step, "t_cgs512_s": t_cgs, "projected_slice_s": T}


def load_traffic():
    """Per-launch HBM bytes of the hot kernels from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return {}
    out = {k: v.get("hbm_bytes_per_launch") for k, v in d.get("kernels", {}).items()}
    out["_source"] = "profiles/pmc_traffic.json (%s)" % d.get("workload", "")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--mode", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--grid", type=int, default=100)
    ap.add_argument("--nslices", type=int, default=10)
    ap.add_argument("--first-slice", type=int, default=4)
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--nvec", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this library runs on one GPU only (multi-GPU sharding is "
              "designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from pevsl_amd import capi, driver
    lib = capi.load()
    g = args.grid
    prec = capi.PEVSL_PREC_FP32 if args.mode == "mixed" else capi.PEVSL_PREC_FP64
    lmin, lmax, a, b = 0.0, 12.0, 0.6, 1.2
    tol = 1e-8

    # ---- set-up (not in the timed region)
    t_setup0 = time.time()
    ia, ja, av = driver.laplacian3d(g, g, g)
    M = driver.Matrix(lib, ia, ja, av)
    inst = driver.Instance(lib, M, prec=prec)
    tune = capi.TuneResult()
    lib.pEVSL_Autotune(inst.p, 20, C.byref(tune))
    if args.batch == 1:
        lib.pEVSL_SetSpmvTuning(inst.p, tune.lanes, 1)
    t_kpm0 = time.time()
    rc, mu, ecount = inst.kpmdos(300, 1, args.nvec, [a, b, lmin, lmax])
    assert rc == 0, rc
    rc, sli = driver.spslicer_kpm(lib, mu, [a, b, lmin, lmax], args.nslices, int(10 * ecount))
    assert rc == 0, rc
    t_kpm = time.time() - t_kpm0
    ev_int = int(1 + ecount / float(args.nslices))
    mlan = min(max(5 * (ev_int + 2), 300), M.n)
    vinit = driver.rand_vinit(lib, M.n)
    pols = [driver.find_pol(lib, [sli[s], sli[s + 1], lmin, lmax])[0]
            for s in range(args.nslices)]
    t_setup = time.time() - t_setup0
    log("setup %.2fs (kpm+slicing %.2fs), ecount %.1f, lanes %d, degs %s" % (
        t_setup, t_kpm, ecount, tune.lanes, [p.deg for p in pols]))

    def step(i, inst=inst):
        sl = [(args.first_slice + (i * args.batch) + j) % args.nslices for j in range(args.batch)]
        intvs = [[sli[s], sli[s + 1], lmin, lmax] for s in sl]
        rc, outs = driver.cheblan_batch(inst, intvs, [mlan] * len(sl), tol, vinit,
                                        [pols[s] for s in sl])
        assert rc == 0, rc
        res = []
        for s, o in zip(sl, outs):
            ex = driver.exact_eig_lap3(g, g, g, sli[s], sli[s + 1])
            lam = np.sort(o["lam"])
            err = float(np.max(np.abs(lam - ex) / np.abs(ex))) if len(lam) == len(ex) and len(ex) else None
            res.append({"slice": s, "found": len(lam), "exact": len(ex), "deg": pols[s].deg,
                        "its": o["info"]["its"], "max_rel_err": err,
                        "max_res": float(np.max(o["res"])) if len(o["res"]) else None,
                        "max_rel_res": float(np.max(np.asarray(o["res"]) / np.abs(o["lam"])))
                        if len(o["res"]) else None})
            log("slice %d: found %d exact %d its %d deg %d err %s" % (
                s, len(lam), len(ex), o["info"]["its"], pols[s].deg, err))
        return res

    for w in range(args.warmup):
        step(args.steps + w)  # warm-up on slices outside the timed set
    lib.pEVSL_StatsReset(inst.p)
    t0 = time.time()
    results = []
    for i in range(args.steps):
        results += step(i)
    elapsed = time.time() - t0
    ks = inst.kernel_stats()
    found = sum(r["found"] for r in results)
    value = found / elapsed

    cgs_bw = ks["cgs_bytes"] / ks["t_cgs"] / 1e9 if ks["t_cgs"] > 0 else 0.0
    filt_bw = ks["filter_bytes"] / ks["t_filter"] / 1e9 if ks["t_filter"] > 0 else 0.0
    passes = max(1, ks["cgs_passes"])
    flaunch = max(1, ks["filter_launches"])
    traffic = load_traffic()
    cgs = {"bound": "hbm", "kernel": "CGS-DGKS pass (k_lan_gemvt + k_lan_gemvn)",
           "achieved": round(cgs_bw, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(cgs_bw / HBM_PEAK_GBS, 4),
           "traffic": traffic.get("cgs_pass_fp64_per_column"),
           "bytes_per_pass_avg": ks["cgs_bytes"] / passes,
           "sec_per_pass_avg": ks["t_cgs"] / passes}
    fkern = "k_cheb_stream<double>" if (prec == capi.PEVSL_PREC_FP64 and args.batch == 1) \
        else "block filter kernels"
    filt = {"bound": "hbm", "kernel": "fused Chebyshev step (%s)" % fkern,
            "achieved": round(filt_bw, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(filt_bw / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("k_cheb_stream<double>") if fkern.startswith("k_cheb") else None,
            "bytes_per_launch": ks["filter_bytes"] / flaunch,
            "sec_per_launch_avg": ks["t_filter"] / flaunch,
            "launches": ks["filter_launches"], "column_steps": ks["filter_steps"],
            "t_filter_s": ks["t_filter"]}
    # the dominant single kernel: the filter step, unless one CGS kernel (about
    # half of a pass) takes longer
    roof = dict(filt) if ks["t_filter"] >= 0.5 * ks["t_cgs"] else dict(cgs)
    roof["traffic_source"] = traffic.get("_source")
    # what a plain copy kernel reaches on arrays beyond the 256 MB Infinity
    # Cache (2 x 2 GiB, HIP events): the practical ceiling beside the paper peak
    csec, cbytes = C.c_double(0.0), C.c_double(0.0)
    if lib.pEVSL_BenchCopy(C.c_size_t(1 << 28), 10, C.byref(csec), C.byref(cbytes)) == 0 \
            and csec.value > 0:
        copy_bw = cbytes.value / csec.value / 1e9
        for r_ in (roof, filt, cgs):
            r_["copy_GBs"] = round(copy_bw, 1)
            r_["frac_of_copy"] = round(r_["achieved"] / copy_bw, 4)
    mixed = None
    if not args.no_mixed:
        instm = driver.Instance(lib, M, prec=capi.PEVSL_PREC_FP32,
                                basis_prec=capi.PEVSL_PREC_FP32)
        tm = capi.TuneResult()
        lib.pEVSL_Autotune(instm.p, 20, C.byref(tm))
        if args.batch == 1:
            lib.pEVSL_SetSpmvTuning(instm.p, tm.lanes, 1)
        t1 = time.time()
        mres = []
        for i in range(args.steps):
            mres += step(i, instm)
        tmix = time.time() - t1
        ksm = instm.kernel_stats()
        instm.close()
        mfound = sum(r["found"] for r in mres)
        mixed = {"value": round(mfound / tmix, 3), "unit": "eigenpairs/s",
                 "seconds": round(tmix, 2), "speedup_vs_fp64": round((mfound / tmix) / value, 3),
                 "found": mfound, "exact": sum(r["exact"] for r in mres),
                 "max_rel_err": max((r["max_rel_err"] or 0.0) for r in mres),
                 "max_residual": max((r["max_res"] or 0.0) for r in mres),
                 "max_rel_residual": max((r["max_rel_res"] or 0.0) for r in mres),
                 "tolerance": "residual < 1e-6 * max(|lmin|, |lmax|) = 1.2e-5 (FP32-filtered basis), "
                              "eigenvalue rel err < 1e-10",
                 "phase_s": {"filter": round(ksm["t_filter"], 3), "cgs": round(ksm["t_cgs"], 3),
                             "eig": round(ksm["t_eig"], 3), "ritz": round(ksm["t_ritz"], 3)},
                 "filter_GBs": round(ksm["filter_bytes"] / max(ksm["t_filter"], 1e-9) / 1e9, 1),
                 "cgs_GBs": round(ksm["cgs_bytes"] / max(ksm["t_cgs"], 1e-9) / 1e9, 1),
                 "autotune": {"block_width": tm.block_width, "block_kernel": tm.block_rowkernel}}
    cpu = None
    if not args.no_cpu_baseline:
        r0 = results[0]
        s0 = r0["slice"]
        try:
            cpu = cpu_baseline(g, float(sli[s0]), float(sli[s0 + 1]), ev_int + 2, r0["its"],
                               r0["found"])
        except Exception as e:  # the baseline must not hide the GPU result
            cpu = {"value": None, "error": str(e)}
    out = {
        "metric": "eigenpairs/sec + filtered-SpMM GB/s (% HBM roofline), mixed vs FP64, 1e-10 rel",
        "value": round(value, 3), "unit": "eigenpairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == capi.PEVSL_PREC_FP64 else "f32-filter/f64-lanczos",
        "data": "synthetic (3-D Laplacian, libc rand() start vectors)",
        "config": {"workload": "Laplacian %d^3 (n=%d, nnz=%d), [0.6,1.2] in %d slices; "
                               "one slice per step" % (g, M.n, M.nnz, args.nslices),
                   "slices_timed": [r["slice"] for r in results], "mode": args.mode,
                   "batch": args.batch, "kpm_nvec": args.nvec, "mlan": mlan,
                   "parallelism": "1 GPU"},
        "correctness": {"found": found, "exact": sum(r["exact"] for r in results),
                        "max_rel_err": max((r["max_rel_err"] or 0.0) for r in results),
                        "max_residual": max((r["max_res"] or 0.0) for r in results),
                        "max_rel_residual": max((r["max_rel_res"] or 0.0) for r in results),
                        "per_slice": results},
        "setup_s": round(t_setup, 3), "kpm_slicing_s": round(t_kpm, 3),
        # the whole-job rate with the one-time matrix upload, KPM, slicing and
        # FindPol of all slices charged to the timed slices
        "value_incl_setup": round(found / (elapsed + t_setup), 3),
        "ecount": ecount,
        "phase_s": {"filter": round(ks["t_filter"], 3), "reorth": round(ks["t_reorth"], 3),
                    "cgs": round(ks["t_cgs"], 3), "eig": round(ks["t_eig"], 3),
                    "ritz": round(ks["t_ritz"], 3)},
        "autotune": {"lanes": tune.lanes, "block_width": tune.block_width},
        "roofline": roof, "filter_roofline": filt, "cgs_roofline": cgs, "mixed": mixed,
        "cpu_baseline": cpu,
    }
    for p in pols:
        lib.pEVSL_FreePol(C.byref(p))
    inst.close()
    M.free()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
