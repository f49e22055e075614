"""Short run of the hot kernels for hardware-counter profiling.

A full bench slice launches ~10^5 kernels, too many to collect counters on;
this probe launches each hot kernel a few times on the 100^3 Laplacian
(n = 1e6, nnz = 6.94e6) so that `rocprofv3 --pmc` finishes in seconds:

  copy       pEVSL_BenchCopy: plain copy of 2^26 doubles (512 MiB read,
             512 MiB written per launch), the known-size reference that
             calibrates the byte counters;
  filter64   pEVSL_ChebAv with a degree-168 filter on one FP64 column
             (CSR-stream kernel, the plain FP64 path's filter step);
  cgs64      pEVSL_BenchReorth, k = 1024, FP64 basis;
  mixed      pEVSL_Autotune of the FP32 filter (every candidate kernel,
             including the width-8 Clenshaw block SpMM) and
             pEVSL_BenchReorth, k = 1024, FP32 basis.

  irregular  (--irregular) the power-law matrix of the irregular workload
             (n = 2e6, nnz = 6.6e7): the FP32 Clenshaw filter step at block
             width 16 with the row-per-thread kernel and with the
             nnz-balanced kernel (pEVSL_SetBlockKernel 2 and 4), 5 launches
             each after one warm-up (pEVSL_BenchBlock).

usage: python -m pevsl_amd.profile_probe [--irregular]
"""
import ctypes as C
import sys

import numpy as np

from . import capi, driver


def irregular():
    """Row kernel vs nnz-balanced kernel on the irregular workload's matrix:
    pEVSL_BenchBlock times back-to-back FP32 Clenshaw steps of width 16 with
    the instance's block kernel."""
    lib = capi.load()
    sec, byt = C.c_double(), C.c_double()
    lib.pEVSL_BenchCopy(1 << 26, 5, C.byref(sec), C.byref(byt))
    print("[probe] copy %.3f ms" % (sec.value * 1e3), flush=True)
    print("[probe] generating the power-law matrix n=2e6", flush=True)
    ia, ja, a = driver.powerlaw_sym(2000000, seed=4)
    M = driver.Matrix(lib, ia, ja, a)
    try:
        for kind in (2, 4):
            inst = driver.Instance(lib, M, prec=capi.PEVSL_PREC_FP32, block_rowkernel=kind)
            try:
                rc = lib.pEVSL_BenchBlock(inst.p, 16, 5, C.byref(sec), C.byref(byt))
                assert rc == 0
                print("[probe] block kernel %d, width 16: %.1f us per step, %.1f us per column,"
                      " %.2f TB/s algorithmic" % (kind, sec.value * 1e6, sec.value * 1e6 / 16,
                                                  byt.value / sec.value / 1e12), flush=True)
            finally:
                inst.close()
    finally:
        M.free()
    print("[probe] done", flush=True)
    return 0


def main():
    if "--irregular" in sys.argv[1:]:
        return irregular()
    lib = capi.load()
    print("[probe] copy", flush=True)
    sec, byt = C.c_double(), C.c_double()
    lib.pEVSL_BenchCopy(1 << 26, 5, C.byref(sec), C.byref(byt))
    print("[probe] copy %.3f ms, %.1f GB/s" % (sec.value * 1e3, byt.value / sec.value / 1e9),
          flush=True)
    ia, ja, a = driver.laplacian3d(100, 100, 100)
    M = driver.Matrix(lib, ia, ja, a)
    try:
        print("[probe] filter64", flush=True)
        inst = driver.Instance(lib, M, lanes=capi.PEVSL_SPMV_STREAM)
        try:
            pol, rc = driver.find_pol(lib, [0.8397, 0.8865, 0.0, 12.0])
            assert rc == 0
            n = M.n
            x = np.random.default_rng(1).uniform(-1, 1, n)
            y = np.empty(n)
            X, Y = capi.Parvec(), capi.Parvec()
            lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(X), x.ctypes.data_as(capi.c_double_p))
            lib.pEVSL_ParvecCreateShell(n, n, 0, 0, C.byref(Y), y.ctypes.data_as(capi.c_double_p))
            rc = lib.pEVSL_ChebAv(inst.p, C.byref(pol), C.byref(X), C.byref(Y), None)
            assert rc == 0
            print("[probe] filter64 degree %d" % pol.deg, flush=True)
            lib.pEVSL_FreePol(C.byref(pol))
            print("[probe] cgs64", flush=True)
            lib.pEVSL_BenchReorth(inst.p, 1024, 3, C.byref(sec), C.byref(byt))
            print("[probe] cgs64 %.3f ms/pass" % (sec.value * 1e3), flush=True)
        finally:
            inst.close()
        print("[probe] mixed", flush=True)
        inst = driver.Instance(lib, M, prec=capi.PEVSL_PREC_FP32,
                               basis_prec=capi.PEVSL_PREC_FP32)
        try:
            tr = capi.TuneResult()
            lib.pEVSL_Autotune(inst.p, 5, C.byref(tr))
            lib.pEVSL_BenchReorth(inst.p, 1024, 3, C.byref(sec), C.byref(byt))
            print("[probe] cgs32 %.3f ms/pass; tuned width %d kernel %d" % (
                sec.value * 1e3, tr.block_width, tr.block_rowkernel), flush=True)
        finally:
            inst.close()
    finally:
        M.free()
    print("[probe] done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
