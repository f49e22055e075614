"""ctypes binding of libpevsl_amd.so (the C API declared in include/pevsl.h
and include/pevsl_amd.h).

The structures mirror include/pevsl_struct.h field by field.  Loading fails
loudly when the shared library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpevsl_amd.so")

MPI_COMM_WORLD = 0
PEVSL_NOT_IMPLEMENT = 11
PEVSL_ERR_DEVICE_OOM = -7
PEVSL_PREC_FP64 = 0
PEVSL_PREC_FP32 = 1
PEVSL_RAND_LIBC = 0
PEVSL_RAND_RADEMACHER = 1
PEVSL_RAND_UNIFORM = 2
PEVSL_SPMV_STREAM = 128
PEVSL_SLICER_MDEG = 20
PEVSL_SLICER_NVEC = 8
PEVSL_SLICER_NBINS = 512
PEVSL_SLICER_NFREQ = 24
PEVSL_SLICER_NIN = 42 + 4 * PEVSL_SLICER_NFREQ

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)
c_float_p = C.POINTER(C.c_float)


class Csr(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("ia", c_int_p),
                ("ja", c_int_p), ("a", c_double_p)]


class Parcsr(C.Structure):
    _fields_ = [("comm", C.c_int), ("comm_handle", C.c_void_p),
                ("nrow_global", C.c_int), ("ncol_global", C.c_int),
                ("nrow_local", C.c_int), ("ncol_local", C.c_int),
                ("row_starts", c_int_p), ("col_starts", c_int_p),
                ("first_row", C.c_int), ("first_col", C.c_int),
                ("diag", C.POINTER(Csr)), ("offd", C.POINTER(Csr)),
                ("col_map_offd", c_int_p), ("dev", C.c_void_p)]


class Parvec(C.Structure):
    _fields_ = [("comm", C.c_int), ("n_global", C.c_int), ("n_local", C.c_int),
                ("n_first", C.c_int), ("data", c_double_p)]


class Parvecs(C.Structure):
    _fields_ = [("comm", C.c_int), ("n_global", C.c_int), ("n_local", C.c_int),
                ("n_first", C.c_int), ("n_vecs", C.c_int), ("ld", C.c_int),
                ("data", c_double_p)]


class Polparams(C.Structure):
    _fields_ = [("max_deg", C.c_int), ("min_deg", C.c_int), ("damping", C.c_int),
                ("thresh_ext", C.c_double), ("thresh_int", C.c_double),
                ("intvtol", C.c_double), ("type", C.c_int), ("mu", c_double_p),
                ("cc", C.c_double), ("dd", C.c_double), ("gam", C.c_double),
                ("bar", C.c_double), ("deg", C.c_int)]


class Stat(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "t_setBsv", "t_setASigBsv", "t_eigbounds", "t_dos", "t_iter", "t_mvA",
        "t_mvB", "t_svB", "t_svLT", "t_svASigB", "t_reorth", "t_eig", "t_blas",
        "t_ritz", "t_polAv", "t_ratAv", "t_sth")] + [(n, C.c_size_t) for n in (
        "n_mvA", "n_mvB", "n_svB", "n_svLT", "n_svASigB", "n_polAv", "n_ratAv",
        "alloced", "alloced_total", "alloced_max")]


class Data(C.Structure):
    _fields_ = [("comm", C.c_int), ("N", C.c_int), ("n", C.c_int),
                ("nfirst", C.c_int), ("ifGenEv", C.c_int), ("Amv", C.c_void_p),
                ("Bmv", C.c_void_p), ("Bsol", C.c_void_p), ("LTsol", C.c_void_p),
                ("stats", C.POINTER(Stat)), ("nev_computed", C.c_int),
                ("eval_computed", c_double_p), ("evec_computed", C.c_void_p),
                ("sigma_mult", C.c_double), ("dev", C.c_void_p)]


class LanInfo(C.Structure):
    _fields_ = [("its", C.c_int), ("kdim", C.c_int), ("nbreak", C.c_int),
                ("nconv", C.c_int), ("nev", C.c_int), ("deg", C.c_int),
                ("nmatvec", C.c_longlong), ("tr", C.c_double),
                ("t_total", C.c_double), ("tol_accept", C.c_double)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_longlong),
                ("seconds", C.c_double), ("bytes", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("lanes", C.c_int), ("block_width", C.c_int),
                ("t_lanes", C.c_double * 6), ("bytes_lanes", C.c_double),
                ("t_block", C.c_double * 6), ("bytes_block", C.c_double * 6),
                ("t_stream", C.c_double), ("t_blockrow", C.c_double * 6),
                ("block_rowkernel", C.c_int), ("t_blockclen", C.c_double * 6),
                ("bytes_clen", C.c_double * 6), ("t_blockbal", C.c_double * 6),
                ("t_blockbalclen", C.c_double * 6), ("t_blockseg", C.c_double * 6)]


class StepBench(C.Structure):
    _fields_ = [("nslices", C.c_int), ("k", C.c_int), ("reps", C.c_int),
                ("sec_step", C.c_double), ("wall_step", C.c_double),
                ("sec_filter", C.c_double), ("sec_reorth", C.c_double),
                ("sec_cgs", C.c_double), ("filter_bytes", C.c_double),
                ("cgs_bytes", C.c_double), ("filter_launches", C.c_longlong),
                ("eager_group_sec", C.c_double), ("graph_group_sec", C.c_double),
                ("graph_nodes", C.c_int), ("graph_group_width", C.c_int)]


class KernelStats(C.Structure):
    _fields_ = [("t_filter", C.c_double), ("t_reorth", C.c_double),
                ("t_eig", C.c_double), ("t_ritz", C.c_double),
                ("t_kpm", C.c_double), ("filter_bytes", C.c_double),
                ("reorth_bytes", C.c_double), ("filter_steps", C.c_longlong),
                ("filter_launches", C.c_longlong), ("t_cgs", C.c_double),
                ("cgs_bytes", C.c_double), ("cgs_passes", C.c_longlong)]


DataP = C.POINTER(Data)

# name -> (restype, argtypes); the exact set of symbols declared in include/*.h
SIGNATURES = {
    # chebpol.c
    "pEVSL_SetPolDef": (None, [C.POINTER(Polparams)]),
    "pEVSL_FindPol": (C.c_int, [c_double_p, C.POINTER(Polparams)]),
    "pEVSL_FreePol": (None, [C.POINTER(Polparams)]),
    # kpmdos.c / landos.c / lantrbnd.c
    "pEVSL_Kpmdos": (C.c_int, [DataP, C.c_int, C.c_int, C.c_int, c_double_p,
                               C.c_int, C.c_int, C.c_int, c_double_p, c_double_p]),
    "pEVSL_LanDosG": (C.c_int, [DataP, C.c_int, C.c_int, C.c_int, c_double_p,
                                c_double_p, c_double_p, c_double_p, C.c_int,
                                C.c_int, C.c_int]),
    "pEVSL_SpslicerLan": (None, [c_double_p, c_double_p, C.c_int, C.c_int,
                                 c_double_p]),
    "pEVSL_LanTrbounds": (C.c_int, [DataP, C.c_int, C.c_int, C.c_double,
                                    C.POINTER(Parvec), C.c_int, c_double_p,
                                    c_double_p, C.c_void_p]),
    # pevsl.c
    "pEVSL_Start": (C.c_int, [C.c_int, C.POINTER(DataP)]),
    "pEVSL_Finish": (C.c_int, [DataP]),
    "pEVSL_SetProbSizes": (C.c_int, [DataP, C.c_int, C.c_int, C.c_int]),
    "pEVSL_SetAParcsr": (C.c_int, [DataP, C.POINTER(Parcsr)]),
    "pEVSL_SetBParcsr": (C.c_int, [DataP, C.POINTER(Parcsr)]),
    "pEVSL_SetAMatvec": (C.c_int, [DataP, C.c_void_p, C.c_void_p]),
    "pEVSL_SetBMatvec": (C.c_int, [DataP, C.c_void_p, C.c_void_p]),
    "pEVSL_SetBSol": (C.c_int, [DataP, C.c_void_p, C.c_void_p]),
    "pEVSL_SetLTSol": (C.c_int, [DataP, C.c_void_p, C.c_void_p]),
    "pEVSL_SetStdEig": (C.c_int, [DataP]),
    "pEVSL_SetGenEig": (C.c_int, [DataP]),
    "pEVSL_SetSigmaMult": (C.c_int, [DataP, C.c_double]),
    # parcsr.c / parcsrmv.c
    "pEVSL_ParcsrCreate": (C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p,
                                     C.POINTER(Parcsr), C.c_int]),
    "pEVSL_ParcsrSetup": (C.c_int, [C.POINTER(Csr), C.POINTER(Parcsr)]),
    "pEVSL_ParcsrFree": (None, [C.POINTER(Parcsr)]),
    "pEVSL_ParcsrGetLocalMat": (C.c_int, [C.POINTER(Parcsr), C.c_int, C.c_void_p,
                                          C.POINTER(Csr), C.c_char]),
    "pEVSL_ParcsrNnz": (C.c_int, [C.POINTER(Parcsr)]),
    "pEVSL_ParcsrLocalNnz": (C.c_int, [C.POINTER(Parcsr)]),
    "pEVSL_ParcsrMatvec": (None, [C.POINTER(Parcsr), C.POINTER(Parvec),
                                  C.POINTER(Parvec)]),
    # parvec.c
    "pEVSL_ParvecCreate": (None, [C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(Parvec)]),
    "pEVSL_ParvecCreateShell": (None, [C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(Parvec), c_double_p]),
    "pEVSL_ParvecDupl": (None, [C.POINTER(Parvec), C.POINTER(Parvec)]),
    "pEVSL_ParvecFree": (None, [C.POINTER(Parvec)]),
    "pEVSL_ParvecRand": (None, [C.POINTER(Parvec)]),
    "pEVSL_ParvecDot": (None, [C.POINTER(Parvec), C.POINTER(Parvec), c_double_p]),
    "pEVSL_ParvecNrm2": (None, [C.POINTER(Parvec), c_double_p]),
    "pEVSL_ParvecCopy": (None, [C.POINTER(Parvec), C.POINTER(Parvec)]),
    "pEVSL_ParvecSum": (None, [C.POINTER(Parvec), c_double_p]),
    "pEVSL_ParvecScal": (None, [C.POINTER(Parvec), C.c_double]),
    "pEVSL_ParvecSetScalar": (None, [C.POINTER(Parvec), C.c_double]),
    "pEVSL_ParvecSetZero": (None, [C.POINTER(Parvec)]),
    "pEVSL_ParvecAxpy": (None, [C.c_double, C.POINTER(Parvec), C.POINTER(Parvec)]),
    "pEVSL_ParvecSameSize": (C.c_int, [C.POINTER(Parvec), C.POINTER(Parvec)]),
    "pEVSL_ParvecWrite": (C.c_int, [C.POINTER(Parvec), C.c_char_p]),
    # parvecs.c
    "pEVSL_ParvecsCreateShell": (None, [C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.POINTER(Parvecs),
                                        c_double_p]),
    "pEVSL_ParvecsCreate": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.POINTER(Parvecs)]),
    "pEVSL_ParvecsDuplParvec": (None, [C.POINTER(Parvec), C.c_int, C.c_int,
                                       C.POINTER(Parvecs)]),
    "pEVSL_ParvecsFree": (None, [C.POINTER(Parvecs)]),
    "pEVSL_ParvecsGetParvecShell": (None, [C.POINTER(Parvecs), C.c_int,
                                           C.POINTER(Parvec)]),
    "pEVSL_ParvecsGemv": (None, [C.c_double, C.POINTER(Parvecs), C.c_int,
                                 c_double_p, C.c_double, C.POINTER(Parvec)]),
    "pEVSL_ParvecsGemtvWithWspace": (None, [C.c_double, C.POINTER(Parvecs),
                                            C.c_int, C.POINTER(Parvec),
                                            C.c_double, c_double_p, c_double_p]),
    "pEVSL_ParvecsGemtv": (None, [C.c_double, C.POINTER(Parvecs), C.c_int,
                                  C.POINTER(Parvec), C.c_double, c_double_p]),
    # spmat.c
    "pEVSL_CsrResize": (None, [C.c_int, C.c_int, C.c_int, C.POINTER(Csr)]),
    "pEVSL_FreeCsr": (None, [C.POINTER(Csr)]),
    "pEVSL_CooResize": (None, [C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "pEVSL_FreeCoo": (None, [C.c_void_p]),
    "pEVSL_CooToCsr": (C.c_int, [C.c_int, C.c_void_p, C.POINTER(Csr)]),
    "pEVSL_CsrToCoo": (C.c_int, [C.POINTER(Csr), C.c_int, C.c_void_p]),
    "pEVSL_MatvecGen": (C.c_int, [C.c_double, C.POINTER(Csr), c_double_p,
                                  C.c_double, c_double_p]),
    "pEVSL_Matvec": (C.c_int, [C.POINTER(Csr), c_double_p, c_double_p]),
    "pEVSL_SortRow": (None, [C.POINTER(Csr)]),
    # utils.c / vector.c
    "pEVSL_SortInt": (None, [c_int_p, C.c_int]),
    "pEVSL_SortDouble": (None, [C.c_int, c_double_p, c_int_p]),
    "pEVSL_LinSpace": (None, [C.c_double, C.c_double, C.c_int, c_double_p]),
    "pEVSL_Part1d": (None, [C.c_int, C.c_int, c_int_p, c_int_p, c_int_p, C.c_int]),
    "pEVSL_BinarySearch": (C.c_int, [c_int_p, C.c_int, C.c_int]),
    "pEVSL_Vecset": (None, [C.c_int, C.c_double, c_double_p]),
    "pEVSL_Wtime": (C.c_double, []),
    "linspace": (None, [C.c_double, C.c_double, C.c_int, c_double_p]),
    "vecset": (None, [C.c_int, C.c_double, c_double_p]),
    "sort_double": (None, [C.c_int, c_double_p, c_int_p]),
    # spslicer.c / stats.c / cheblanNr.c
    "pEVSL_SpslicerKpm": (C.c_int, [c_double_p, c_double_p, C.c_int, c_double_p,
                                    C.c_int, C.c_int]),
    "pEVSL_StatsReset": (None, [DataP]),
    "pEVSL_StatsPrint": (None, [DataP, C.c_void_p]),
    "pEVSL_ChebLanNr": (C.c_int, [DataP, c_double_p, C.c_int, C.c_double,
                                  C.POINTER(Parvec), C.POINTER(Polparams), c_int_p,
                                  C.POINTER(c_double_p), C.POINTER(C.POINTER(Parvecs)),
                                  C.POINTER(c_double_p), C.c_void_p]),
    # pevsl_amd.h extensions
    "pEVSL_SetFilterPrecision": (C.c_int, [DataP, C.c_int]),
    "pEVSL_SetMixedTol": (C.c_int, [DataP, C.c_double]),
    "pEVSL_SetRitzRefine": (C.c_int, [DataP, C.c_int]),
    "pEVSL_SetBasisPrecision": (C.c_int, [DataP, C.c_int]),
    "pEVSL_SetRandMode": (C.c_int, [DataP, C.c_int, C.c_ulonglong]),
    "pEVSL_SetKeepVectors": (C.c_int, [DataP, C.c_int]),
    "pEVSL_SetBlockKernel": (C.c_int, [DataP, C.c_int]),
    "pEVSL_RandSeed": (None, [C.c_uint]),
    "pEVSL_BenchBlock": (C.c_int, [DataP, C.c_int, C.c_int, c_double_p, c_double_p]),
    "pEVSL_BenchBatchStep": (C.c_int, [DataP, C.c_int, c_double_p, C.POINTER(Polparams),
                                       C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(StepBench)]),
    "pEVSL_ChebIterSetup": (C.c_int, [C.c_double, C.c_double, C.c_int, C.POINTER(Parcsr),
                                      C.POINTER(C.c_void_p)]),
    "pEVSL_ChebIterSolv1": (None, [c_double_p, c_double_p, C.c_void_p]),
    "pEVSL_ChebIterSolv2": (None, [c_double_p, c_double_p, C.c_void_p]),
    "pEVSL_ChebIterFree": (None, [C.c_void_p]),
    "pEVSL_ChebIterGetRes": (c_double_p, [C.c_void_p]),
    "pEVSL_ChebIterStatsPrint": (None, [C.c_void_p, C.c_void_p]),
    # Fortran 90 interface (include/pevsl_f90.h): all arguments by reference
    "pevsl_start_f90_": (None, [C.c_void_p] * 2),
    "pevsl_finish_f90_": (None, [C.c_void_p] * 1),
    "pevsl_seta_parcsr_f90_": (None, [C.c_void_p] * 2),
    "pevsl_setb_parcsr_f90_": (None, [C.c_void_p] * 2),
    "pevsl_setprobsizes_f90_": (None, [C.c_void_p] * 4),
    "pevsl_setamv_f90_": (None, [C.c_void_p] * 3),
    "pevsl_setbmv_f90_": (None, [C.c_void_p] * 3),
    "pevsl_setbsol_f90_": (None, [C.c_void_p] * 3),
    "pevsl_set_stdeig_f90_": (None, [C.c_void_p] * 1),
    "pevsl_set_geneig_f90_": (None, [C.c_void_p] * 1),
    "pevsl_parcsrcreate_f90_": (None, [C.c_void_p] * 9),
    "pevsl_parcsr_free_f90_": (None, [C.c_void_p] * 1),
    "pevsl_parcsrmatvec_f90_": (None, [C.c_void_p] * 3),
    "pevsl_bsv_f90_": (None, [C.c_void_p] * 3),
    "pevsl_lanbounds_f90_": (None, [C.c_void_p] * 6),
    "pevsl_findpol_f90_": (None, [C.c_void_p] * 4),
    "pevsl_freepol_f90_": (None, [C.c_void_p] * 1),
    "pevsl_cheblannr_f90_": (None, [C.c_void_p] * 5),
    "pevsl_get_nev_f90_": (None, [C.c_void_p] * 2),
    "pevsl_copy_result_f90_": (None, [C.c_void_p] * 4),
    "pevsl_setup_chebiter_f90_": (None, [C.c_void_p] * 5),
    "pevsl_chebiter_f90_": (None, [C.c_void_p] * 4),
    "pevsl_chebiterstatsprint_f90_": (None, [C.c_void_p] * 1),
    "pevsl_free_chebiterb_f90_": (None, [C.c_void_p] * 1),
    "pevsl_setbsol_chebiter_f90_": (None, [C.c_void_p] * 3),
    "pevsl_setup_lspolsqrt_f90_": (None, [C.c_void_p] * 6),
    "pevsl_lspolsol_f90_": (None, [C.c_void_p] * 3),
    "pevsl_free_lspol_f90_": (None, [C.c_void_p] * 1),
    "pevsl_setltsol_lspol_f90_": (None, [C.c_void_p] * 2),
    "pevsl_kpmdos_ecnt_f90_": (None, [C.c_void_p] * 7),
    "pevsl_landos_ecnt_f90_": (None, [C.c_void_p] * 6),
    "pevsl_set_sigmamult_f90_": (None, [C.c_void_p] * 2),
    "pEVSL_SetupLSPolSqrt": (None, [C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(Parcsr), C.POINTER(C.c_void_p)]),
    "pEVSL_SetupLSPolRec": (None, [C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.POINTER(Parcsr), C.POINTER(C.c_void_p)]),
    "pEVSL_SetupLSPol": (None, [C.c_int, C.c_double, C.c_double, C.c_double,
                                C.POINTER(Parcsr), C.c_void_p, C.POINTER(C.c_void_p)]),
    "pEVSL_LSPolSol": (None, [c_double_p, c_double_p, C.c_void_p]),
    "pEVSL_LSPolFree": (None, [C.c_void_p]),
    "pEVSL_LSPolInfo": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), c_double_p, c_double_p,
                                  c_double_p]),
    "pEVSL_SetSpmvTuning": (C.c_int, [DataP, C.c_int, C.c_int]),
    "pEVSL_ChebLanNrBatch": (C.c_int, [DataP, C.c_int, c_double_p, c_int_p,
                                       C.c_double, C.POINTER(Parvec),
                                       C.POINTER(Polparams), c_int_p,
                                       C.POINTER(c_double_p),
                                       C.POINTER(C.POINTER(Parvecs)),
                                       C.POINTER(c_double_p), C.c_void_p]),
    "pEVSL_ChebAv": (C.c_int, [DataP, C.POINTER(Polparams), C.POINTER(Parvec),
                               C.POINTER(Parvec), C.POINTER(Parvec)]),
    "pEVSL_GetLanInfo": (C.c_int, [DataP, C.c_int, C.POINTER(LanInfo)]),
    "pEVSL_GetKernelStats": (C.c_int, [DataP, C.POINTER(KernelStats)]),
    "pEVSL_GetKernelTimes": (C.c_int, [DataP, C.c_int, C.POINTER(KernelTime),
                                       C.POINTER(C.c_int)]),
    "pEVSL_DeviceMatvec": (C.c_int, [DataP, C.c_void_p, C.c_void_p]),
    "pEVSL_ChebAvNoAlloc": (C.c_int, [DataP, C.POINTER(Polparams), C.c_void_p, C.c_void_p]),
    "pEVSL_SetBatchWave": (C.c_int, [DataP, C.c_int]),
    "pEVSL_SetGraphFilter": (C.c_int, [DataP, C.c_int]),
    "pEVSL_SetGramKernel": (C.c_int, [DataP, C.c_int]),
    "pEVSL_SetColumnOffsets16": (C.c_int, [DataP, C.c_int]),
    "pEVSL_ColumnOffsets16": (C.c_int, [DataP]),
    "pEVSL_BenchGram": (C.c_int, [DataP, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pEVSL_CGS_DGKS": (C.c_int, [DataP, C.c_int, C.c_int, C.POINTER(Parvecs), C.POINTER(Parvec),
                                 C.POINTER(C.c_double), C.c_void_p]),
    "pEVSL_ChebAvNoAllocMaxRows": (C.c_size_t, []),
    "pEVSL_DeviceInfo": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_size_t)]),
    "pEVSL_SymmTridEigvals": (C.c_int, [C.c_int, c_double_p, c_double_p, C.c_double,
                                        c_double_p, c_int_p, c_double_p]),
    "pEVSL_Autotune": (C.c_int, [DataP, C.c_int, C.POINTER(TuneResult)]),
    "pEVSL_BenchReorth": (C.c_int, [DataP, C.c_int, C.c_int, c_double_p, c_double_p]),
    "pEVSL_BenchCopy": (C.c_int, [C.c_size_t, C.c_int, c_double_p, c_double_p]),
    # learned slicer
    "pEVSL_MlpCreate": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_ulonglong, C.POINTER(C.c_void_p)]),
    "pEVSL_MlpFree": (C.c_int, [C.c_void_p]),
    "pEVSL_MlpShape": (C.c_int, [C.c_void_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                 C.POINTER(C.c_size_t)]),
    "pEVSL_MlpSetWeights": (C.c_int, [C.c_void_p, c_float_p, C.c_size_t]),
    "pEVSL_MlpGetWeights": (C.c_int, [C.c_void_p, c_float_p, C.c_size_t]),
    "pEVSL_MlpForward": (C.c_int, [C.c_void_p, c_float_p, c_float_p, C.c_int, c_float_p]),
    "pEVSL_MlpLossGrad": (C.c_int, [C.c_void_p, c_float_p, c_float_p, c_float_p, C.c_int,
                                    c_float_p, c_float_p]),
    "pEVSL_MlpTrainStep": (C.c_int, [C.c_void_p, c_float_p, c_float_p, c_float_p, C.c_int,
                                     C.c_float, c_float_p]),
    "pEVSL_MlpTrain": (C.c_int, [C.c_void_p, c_float_p, c_float_p, c_float_p, C.c_int,
                                 C.c_int, C.c_int, C.c_float, C.c_ulonglong, c_float_p,
                                 C.c_int]),
    "pEVSL_MlpSave": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pEVSL_MlpLoad": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "pEVSL_KpmBinMasses": (C.c_int, [c_double_p, C.c_int, c_double_p, C.c_int, c_double_p,
                                     c_double_p]),
    "pEVSL_SlicerFromBins": (C.c_int, [c_double_p, C.c_int, c_double_p, C.c_int,
                                       c_double_p]),
    "pEVSL_SlicerFeaturesFromMu": (C.c_int, [c_double_p, C.c_int, C.c_int, c_double_p,
                                             c_float_p, c_float_p]),
    "pEVSL_SlicerFeatures": (C.c_int, [DataP, c_double_p, c_float_p, c_float_p]),
    "pEVSL_SpslicerMlp": (C.c_int, [DataP, C.c_void_p, c_double_p, C.c_int, c_double_p,
                                    c_double_p]),
    "pEVSL_SlicerJudge": (C.c_int, [c_double_p, C.c_int, c_double_p, C.c_double, C.c_int,
                                    c_double_p, c_double_p, c_double_p]),
}

_lib = None


def load(path=None):
    """Load libpevsl_amd.so and attach the signatures; raises if it is
    missing or lacks a declared symbol."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            "libpevsl_amd.so not found at %s: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`make -C pevsl_amd`" % p)
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)  # AttributeError if a symbol is missing
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib
