/*
 * pevsl.h -- C API of the MI355X spectrum-slicing library.
 *
 * Same function names, argument meaning, ownership and return codes as the
 * reference's public header INC/pevsl.h (line numbers cited per group), so
 * that a driver such as TESTS/Lap/LapPLanN.c builds against this header and
 * links against libpevsl_amd.so instead of libpevsl.a.  The compute-heavy
 * entry points (ParcsrMatvec, Kpmdos, LanDosG, ChebLanNr) run hand-written
 * HIP kernels on the GPU; host arrays cross the boundary exactly where the
 * reference passes host arrays.
 *
 * Library extensions (batched slices, precision, device buffers) are in
 * pevsl_amd.h.
 */
#ifndef PEVSL_H
#define PEVSL_H

#include <stdio.h>

#include "pevsl_def.h"
#include "pevsl_struct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- chebpol.c (INC/pevsl.h:12-14) ---------------------------------------
 * SetPolDef: defaults (max_deg 5000, min_deg 2, damping 2, thresh_ext 0.2,
 *            thresh_int 0.8, deg 0, intvtol 1e-9)       SRC/chebpol.c:10-19
 * FindPol:   filter design for intv[0..1] inside [intv[2], intv[3]];
 *            malloc's pol->mu; -1 on an invalid interval. SRC/chebpol.c:359
 * FreePol:   frees pol->mu.                              SRC/chebpol.c:491 */
/* INC/pevsl.h:12 */
void pEVSL_SetPolDef(pevsl_Polparams *pol);
/* INC/pevsl.h:13 */
int pEVSL_FindPol(double *intv, pevsl_Polparams *pol);
/* INC/pevsl.h:14 */
void pEVSL_FreePol(pevsl_Polparams *pol);

/* ---- kpmdos.c (INC/pevsl.h:17-18) ----------------------------------------
 * Chebyshev moments mu[0..Mdeg] (caller-allocated) of the spectral density
 * and the eigenvalue-count estimate *ecnt for [intv[0], intv[1]].  All nvec
 * random vectors advance together as columns of one block SpMM on the GPU;
 * with an operator installed by pEVSL_SetAMatvec (host callback) they run
 * one after the other, one callback per degree, as the reference's loop.
 * ngroups/groupid/gl_comm are accepted for source compatibility; with one
 * process ngroups must be 1.                           SRC/kpmdos.c:36-197 */
/* INC/pevsl.h:17 */
int pEVSL_Kpmdos(pevsl_Data *pevsl, int Mdeg, int damping, int nvec,
                 double *intv, int ngroups, int groupid, MPI_Comm gl_comm,
                 double *mu, double *ecnt);

/* ---- landos.c (INC/pevsl.h:21-25) ----------------------------------------
 * Lanczos DOS: nvec random vectors, msteps full-reorthogonalized Lanczos
 * steps each (batched on the GPU), Gaussian-blurred Ritz values on npts
 * points xdos/ydos (caller-allocated), count estimate *neig.
 *                                                       SRC/landos.c:40-314
 * SpslicerLan: equal-share cuts of the integrated DOS.  SRC/landos.c:331 */
/* INC/pevsl.h:21 */
int pEVSL_LanDosG(pevsl_Data *pevsl, const int nvec, int msteps, int npts,
                  double *xdos, double *ydos, double *neig, double *intv,
                  int ngroups, int groupid, MPI_Comm gl_comm);
/* INC/pevsl.h:25 */
void pEVSL_SpslicerLan(double *xi, double *yi, int n_int, int npts,
                       double *sli);

/* ---- lantrbnd.c (INC/pevsl.h:28-30) --------------------------------------
 * Spectrum bounds by thick-restart Lanczos, standard or generalized
 * (B-inner products, after pEVSL_SetGenEig).            SRC/lantrbnd.c:31 */
/* INC/pevsl.h:28 */
int pEVSL_LanTrbounds(pevsl_Data *pevsl, int lanm, int maxit, double tol,
                      pevsl_Parvec *vinit, int bndtype, double *lammin,
                      double *lammax, FILE *fstats);

/* ---- lspolapprox.c (INC/pevsl.h:31-37) -----------------------------------
 * Least-squares Chebyshev approximation of B^{-1/2} on [lmin, lmax] (degree
 * <= max_deg, relative error sum < tol) for the L^T solve of a generalized
 * problem, its solve (host arrays; installed with pEVSL_SetBSol /
 * pEVSL_SetLTSol it runs on the device with no host copies) and its
 * release.  The fit is host work; solves need B set up with
 * pEVSL_ParcsrSetup (its HBM copy).
 *                                       SRC/lspolapprox.c:63-146, 221-321 */
/* INC/pevsl.h:32 */
void pEVSL_SetupLSPolSqrt(int max_deg, double tol, double lmin, double lmax,
                          pevsl_Parcsr *B, void **vdata);
/* INC/pevsl.h:35 */
void pEVSL_LSPolSol(double *db, double *dx, void *data);
/* INC/pevsl.h:37 */
void pEVSL_LSPolFree(void *vdata);

/* ---- pevsl.c (INC/pevsl.h:40-51) -----------------------------------------
 * Start seeds the library's rand() stream (glibc's algorithm) with the rank
 * (0) as SRC/pevsl.c:36 seeds libc rand(), and creates the HIP stream and
 * device context.  SetAParcsr / SetBParcsr use the matrix's HBM copy
 * (SRC/pevsl.c:81-121).  SetAMatvec / SetBMatvec install host callbacks
 * (every product then crosses PCIe).  SetBSol / SetLTSol install the solves
 * of a generalized problem (host callbacks, or pEVSL_LSPolSol, which runs on
 * the device); SetGenEig switches Kpmdos, LanTrbounds, ChebAv and ChebLanNr
 * to A x = lambda B x (SRC/pevsl.c:123-210), and LanDosG as well. */
/* INC/pevsl.h:40 */
int pEVSL_Start(MPI_Comm comm, pevsl_Data **data);
/* INC/pevsl.h:41 */
int pEVSL_Finish(pevsl_Data *pevsl_data);
/* INC/pevsl.h:42 */
int pEVSL_SetProbSizes(pevsl_Data *pevsl_data, int N, int n, int nfirst);
/* INC/pevsl.h:43 */
int pEVSL_SetAParcsr(pevsl_Data *pevsl_data, pevsl_Parcsr *A);
/* INC/pevsl.h:44 */
int pEVSL_SetBParcsr(pevsl_Data *pevsl_data, pevsl_Parcsr *B);
/* INC/pevsl.h:45 */
int pEVSL_SetAMatvec(pevsl_Data *pevsl_data, MVFunc func, void *data);
/* INC/pevsl.h:46 */
int pEVSL_SetBMatvec(pevsl_Data *pevsl_data, MVFunc func, void *data);
/* INC/pevsl.h:47 */
int pEVSL_SetBSol(pevsl_Data *pevsl_data, SVFunc func, void *data);
/* INC/pevsl.h:48 */
int pEVSL_SetLTSol(pevsl_Data *pevsl_data, SVFunc func, void *data);
/* INC/pevsl.h:49 */
int pEVSL_SetStdEig(pevsl_Data *pevsl_data);
/* INC/pevsl.h:50 */
int pEVSL_SetGenEig(pevsl_Data *pevsl_data);
/* INC/pevsl.h:51 */
int pEVSL_SetSigmaMult(pevsl_Data *pevsl_data, double mult);

/* ---- parcsr.c (INC/pevsl.h:54-61) ----------------------------------------
 * One rank: row_starts/col_starts must be NULL or {0, n}; Setup keeps the
 * whole matrix in diag (offd has zero columns).  Setup returns -1 for a CSR
 * of another row count and -2 for a malformed one (ia[0] != 0, ia
 * decreasing, or a column outside [0, ncol)), before allocating anything.
 *                                                      SRC/parcsr.c:24-304 */
/* INC/pevsl.h:54 */
int pEVSL_ParcsrCreate(int nrow, int ncol, int *row_starts, int *col_starts,
                       pevsl_Parcsr *A, MPI_Comm comm);
/* INC/pevsl.h:56 */
int pEVSL_ParcsrSetup(pevsl_Csr *Ai, pevsl_Parcsr *A);
/* INC/pevsl.h:57 */
void pEVSL_ParcsrFree(pevsl_Parcsr *A);
/* INC/pevsl.h:58 */
int pEVSL_ParcsrGetLocalMat(pevsl_Parcsr *A, int idx, pevsl_Coo *coo,
                            pevsl_Csr *csr, char stype);
/* INC/pevsl.h:60 */
int pEVSL_ParcsrNnz(pevsl_Parcsr *A);
/* INC/pevsl.h:61 */
int pEVSL_ParcsrLocalNnz(pevsl_Parcsr *A);

/* ---- parcsrmv.c (INC/pevsl.h:64) -----------------------------------------
 * y = A x with host vectors; the product runs on the GPU.
 *                                                   SRC/parcsrmv.c:84-93 */
/* INC/pevsl.h:64 */
void pEVSL_ParcsrMatvec(pevsl_Parcsr *A, pevsl_Parvec *x, pevsl_Parvec *y);

/* ---- parvec.c (INC/pevsl.h:67-83): host vector utilities ----------------- */
/* INC/pevsl.h:67 */
void pEVSL_ParvecCreate(int nglobal, int nlocal, int nfirst, MPI_Comm comm,
                        pevsl_Parvec *x);
/* INC/pevsl.h:68 */
void pEVSL_ParvecCreateShell(int nglobal, int nlocal, int nfirst,
                             MPI_Comm comm, pevsl_Parvec *x, double *data);
/* INC/pevsl.h:70 */
void pEVSL_ParvecDupl(pevsl_Parvec *x, pevsl_Parvec *y);
/* INC/pevsl.h:71 */
void pEVSL_ParvecFree(pevsl_Parvec *x);
/* INC/pevsl.h:72 */
void pEVSL_ParvecRand(pevsl_Parvec *x);
/* INC/pevsl.h:73 */
void pEVSL_ParvecDot(pevsl_Parvec *x, pevsl_Parvec *y, double *t);
/* INC/pevsl.h:74 */
void pEVSL_ParvecNrm2(pevsl_Parvec *x, double *t);
/* INC/pevsl.h:75 */
void pEVSL_ParvecCopy(pevsl_Parvec *x, pevsl_Parvec *y);
/* INC/pevsl.h:76 */
void pEVSL_ParvecSum(pevsl_Parvec *x, double *t);
/* INC/pevsl.h:77 */
void pEVSL_ParvecScal(pevsl_Parvec *x, double t);
/* INC/pevsl.h:79 */
void pEVSL_ParvecSetScalar(pevsl_Parvec *x, double t);
/* INC/pevsl.h:80 */
void pEVSL_ParvecSetZero(pevsl_Parvec *x);
/* INC/pevsl.h:81 */
void pEVSL_ParvecAxpy(double a, pevsl_Parvec *x, pevsl_Parvec *y);
/* INC/pevsl.h:82 */
int pEVSL_ParvecSameSize(pevsl_Parvec *x, pevsl_Parvec *y);
/* INC/pevsl.h:83 */
int pEVSL_ParvecWrite(pevsl_Parvec *x, const char *fn);

/* ---- parvecs.c (INC/pevsl.h:86-105): host multi-vectors ------------------ */
/* INC/pevsl.h:86 */
void pEVSL_ParvecsCreateShell(int nglobal, int nvecs, int ld, int nlocal,
                              int nfirst, MPI_Comm comm, pevsl_Parvecs *x,
                              double *data);
/* INC/pevsl.h:89 */
void pEVSL_ParvecsCreate(int nglobal, int nvecs, int ld, int nlocal,
                         int nfirst, MPI_Comm comm, pevsl_Parvecs *x);
/* INC/pevsl.h:92 */
void pEVSL_ParvecsDuplParvec(pevsl_Parvec *x, int nvecs, int ld,
                             pevsl_Parvecs *y);
/* INC/pevsl.h:94 */
void pEVSL_ParvecsFree(pevsl_Parvecs *x);
/* INC/pevsl.h:96 */
void pEVSL_ParvecsGetParvecShell(pevsl_Parvecs *X, int i, pevsl_Parvec *x);
/* INC/pevsl.h:98 */
void pEVSL_ParvecsGemv(double alp, pevsl_Parvecs *A, int nvecs, double *x,
                       double bet, pevsl_Parvec *y);
/* INC/pevsl.h:101 */
void pEVSL_ParvecsGemtvWithWspace(double alp, pevsl_Parvecs *A, int nvecs,
                                  pevsl_Parvec *x, double bet, double *y,
                                  double *w);
/* INC/pevsl.h:104 */
void pEVSL_ParvecsGemtv(double alp, pevsl_Parvecs *A, int nvecs,
                        pevsl_Parvec *x, double bet, double *y);

/* ---- spmat.c (INC/pevsl.h:108-116): host sparse utilities ---------------
 * CooToCsr returns -1, and leaves csr untouched, when an index lies outside
 * the declared size. */
/* INC/pevsl.h:108 */
void pEVSL_CsrResize(int nrow, int ncol, int nnz, pevsl_Csr *csr);
/* INC/pevsl.h:109 */
void pEVSL_FreeCsr(pevsl_Csr *csr);
/* INC/pevsl.h:110 */
void pEVSL_CooResize(int nrow, int ncol, int nnz, pevsl_Coo *coo);
/* INC/pevsl.h:111 */
void pEVSL_FreeCoo(pevsl_Coo *coo);
/* INC/pevsl.h:112 */
int pEVSL_CooToCsr(int cooidx, pevsl_Coo *coo, pevsl_Csr *csr);
/* INC/pevsl.h:113 */
int pEVSL_CsrToCoo(pevsl_Csr *csr, int cooidx, pevsl_Coo *coo);
/* INC/pevsl.h:114 */
int pEVSL_MatvecGen(double alp, pevsl_Csr *A, double *x, double bet,
                    double *y);
/* INC/pevsl.h:115 */
int pEVSL_Matvec(pevsl_Csr *A, double *x, double *y);
/* INC/pevsl.h:116 */
void pEVSL_SortRow(pevsl_Csr *A);

/* ---- utils.c / vector.c (INC/pevsl.h:119-130) --------------------------- */
/* INC/pevsl.h:119 */
void pEVSL_SortInt(int *x, int n);
/* INC/pevsl.h:120 */
void pEVSL_SortDouble(int n, double *v, int *ind);
/* INC/pevsl.h:121 */
void pEVSL_LinSpace(double a, double b, int num, double *arr);
/* INC/pevsl.h:122 */
void pEVSL_Part1d(int len, int pnum, int *idx, int *j1, int *j2, int job);
/* INC/pevsl.h:123 */
int pEVSL_BinarySearch(int *x, int n, int key);
/* INC/pevsl.h:124 */
void pEVSL_Vecset(int n, double t, double *v);
/* INC/pevsl.h:125 */
double pEVSL_Wtime(void);
/* INC/pevsl.h:128 */
void linspace(double a, double b, int num, double *arr);
/* INC/pevsl.h:129 */
void vecset(int n, double t, double *v);
/* INC/pevsl.h:130 */
void sort_double(int n, double *v, int *ind);

/* ---- spslicer.c (INC/pevsl.h:133) ----------------------------------------
 * Equal-count slices sli[0..n_int] (caller-allocated) from the KPM moments;
 * the DOS integral at the npts grid points is evaluated on the GPU.
 * Returns 0, or 1 / 2 / 3 as SRC/spslicer.c:127-136.  SRC/spslicer.c:62 */
/* INC/pevsl.h:133 */
int pEVSL_SpslicerKpm(double *sli, double *mu, int Mdeg, double *intv,
                      int n_int, int npts);

/* ---- stats.c (INC/pevsl.h:136-137) --------------------------------------- */
/* INC/pevsl.h:136 */
void pEVSL_StatsReset(pevsl_Data *pevsl);
/* INC/pevsl.h:137 */
void pEVSL_StatsPrint(pevsl_Data *pevsl, FILE *fstats);

/* ---- cheblanNr.c (INC/pevsl.h:140-142) -----------------------------------
 * Filtered Lanczos without restart on p((A-cc)/dd) with full
 * reorthogonalization.  Outputs malloc'd host arrays: *lamo, *reso (maxit
 * entries, the first *nevOut valid) and *Wo (n x nevOut Ritz vectors, or
 * NULL when vector output is switched off with pEVSL_SetKeepVectors).
 * Returns 0, -1 (|gam| > 1) or 2 (zero Ritz vector).  SRC/cheblanNr.c:54 */
/* INC/pevsl.h:140 */
int pEVSL_ChebLanNr(pevsl_Data *pevsl, double *intv, int maxit, double tol,
                    pevsl_Parvec *vinit, pevsl_Polparams *pol, int *nevOut,
                    double **lamo, pevsl_Parvecs **Wo, double **reso,
                    FILE *fstats);

#ifdef __cplusplus
}
#endif

#endif
