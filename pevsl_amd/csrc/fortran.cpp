// fortran.cpp -- the Fortran 90 interface of pEVSL (FORTRAN/pevsl_f90.c):
// one `<name>_f90_` entry point per reference routine (the mangling of
// PEVSL_FORT, INC/pevsl_def.h:108), every argument by reference, objects
// (instance, matrices, polynomials, solver data) passed to Fortran as
// uintptr_t handles.  Each routine forwards to the C ABI; ChebLanNr keeps
// its results in the instance (nev_computed, eval_computed, evec_computed)
// for pevsl_get_nev / pevsl_copy_result, as the reference does.  No Fortran
// compiler is needed to build it; a Fortran caller declares the routines as
// external subroutines with INTEGER*8 handles.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "internal.hpp"

extern "C" {
#include "pevsl_f90.h"
}

namespace {

pevsl_Data *inst(uintptr_t *h) { return reinterpret_cast<pevsl_Data *>(*h); }
pevsl_Parcsr *mat(uintptr_t *h) { return reinterpret_cast<pevsl_Parcsr *>(*h); }

}  // namespace

extern "C" {

void pevsl_start_f90_(MPI_Fint *Fcomm, uintptr_t *pevslf90) {
  pevsl_Data *p = NULL;
  pEVSL_Start(MPI_Comm_f2c(*Fcomm), &p);
  *pevslf90 = reinterpret_cast<uintptr_t>(p);
}

void pevsl_finish_f90_(uintptr_t *pevslf90) { pEVSL_Finish(inst(pevslf90)); }

void pevsl_seta_parcsr_f90_(uintptr_t *pevslf90, uintptr_t *Af90) {
  pEVSL_SetAParcsr(inst(pevslf90), mat(Af90));
}

void pevsl_setb_parcsr_f90_(uintptr_t *pevslf90, uintptr_t *Bf90) {
  pEVSL_SetBParcsr(inst(pevslf90), mat(Bf90));
}

void pevsl_setprobsizes_f90_(uintptr_t *pevslf90, int *N, int *n, int *nfirst) {
  pEVSL_SetProbSizes(inst(pevslf90), *N, *n, *nfirst);
}

void pevsl_setamv_f90_(uintptr_t *pevslf90, void *func, void *data) {
  pEVSL_SetAMatvec(inst(pevslf90), reinterpret_cast<MVFunc>(func), data);
}

void pevsl_setbmv_f90_(uintptr_t *pevslf90, void *func, void *data) {
  pEVSL_SetBMatvec(inst(pevslf90), reinterpret_cast<MVFunc>(func), data);
}

void pevsl_setbsol_f90_(uintptr_t *pevslf90, void *func, void *data) {
  pEVSL_SetBSol(inst(pevslf90), reinterpret_cast<SVFunc>(func), data);
}

void pevsl_set_stdeig_f90_(uintptr_t *pevslf90) { pEVSL_SetStdEig(inst(pevslf90)); }

void pevsl_set_geneig_f90_(uintptr_t *pevslf90) { pEVSL_SetGenEig(inst(pevslf90)); }

void pevsl_parcsrcreate_f90_(int *nrow, int *ncol, int *row_starts, int *col_starts, int *ia,
                             int *ja, double *aa, MPI_Fint *Fcomm, uintptr_t *matf90) {
  pevsl_Parcsr *A;
  PEVSL_MALLOC(A, 1, pevsl_Parcsr);
  pEVSL_ParcsrCreate(*nrow, *ncol, row_starts, col_starts, A, MPI_Comm_f2c(*Fcomm));
  pevsl_Csr loc;
  loc.nrows = A->nrow_local;
  loc.ncols = A->ncol_global;
  loc.ia = ia;
  loc.ja = ja;
  loc.a = aa;
  pEVSL_ParcsrSetup(&loc, A);
  *matf90 = reinterpret_cast<uintptr_t>(A);
}

void pevsl_parcsr_free_f90_(uintptr_t *Af90) {
  pevsl_Parcsr *A = mat(Af90);
  pEVSL_ParcsrFree(A);
  PEVSL_FREE(A);
}

void pevsl_parcsrmatvec_f90_(double *x, double *y, uintptr_t *matf90) {
  pevsl_Parcsr *A = mat(matf90);
  pevsl_Parvec X, Y;
  pEVSL_ParvecCreateShell(A->ncol_global, A->ncol_local, A->first_col, A->comm, &X, x);
  pEVSL_ParvecCreateShell(A->nrow_global, A->nrow_local, A->first_row, A->comm, &Y, y);
  pEVSL_ParcsrMatvec(A, &X, &Y);
}

void pevsl_bsv_f90_(uintptr_t *pevslf90, double *b, double *x) {
  pevsl_Data *p = inst(pevslf90);
  p->Bsol->func(b, x, p->Bsol->data);
}

void pevsl_lanbounds_f90_(uintptr_t *pevslf90, int *mlan, int *nsteps, double *tol,
                          double *lmin, double *lmax) {
  (void)tol;  // the reference's wrapper fixes the tolerance at 1e-8
  pevsl_Data *p = inst(pevslf90);
  pevsl_Parvec vinit;
  pEVSL_ParvecCreate(p->N, p->n, p->nfirst, p->comm, &vinit);
  pEVSL_ParvecRand(&vinit);
  pEVSL_LanTrbounds(p, *mlan, *nsteps, 1e-8, &vinit, 1, lmin, lmax, NULL);
  pEVSL_ParvecFree(&vinit);
}

void pevsl_findpol_f90_(double *xintv, double *thresh_int, double *thresh_ext,
                        uintptr_t *polf90) {
  pevsl_Polparams *pol;
  PEVSL_MALLOC(pol, 1, pevsl_Polparams);
  pEVSL_SetPolDef(pol);
  pol->damping = 2;
  pol->thresh_int = *thresh_int;
  pol->thresh_ext = *thresh_ext;
  pol->max_deg = 50000;
  pEVSL_FindPol(xintv, pol);
  fprintf(stdout, " polynomial deg %d, bar %e gam %e\n", pol->deg, pol->bar, pol->gam);
  *polf90 = reinterpret_cast<uintptr_t>(pol);
}

void pevsl_freepol_f90_(uintptr_t *polf90) {
  pevsl_Polparams *pol = reinterpret_cast<pevsl_Polparams *>(*polf90);
  pEVSL_FreePol(pol);
  PEVSL_FREE(pol);
}

void pevsl_cheblannr_f90_(uintptr_t *pevslf90, double *xintv, int *max_its, double *tol,
                          uintptr_t *polf90) {
  pevsl_Data *p = inst(pevslf90);
  pevsl_Polparams *pol = reinterpret_cast<pevsl_Polparams *>(*polf90);
  pEVSL_StatsReset(p);
  pevsl_Parvec vinit;
  pEVSL_ParvecCreate(p->N, p->n, p->nfirst, p->comm, &vinit);
  pEVSL_ParvecRand(&vinit);
  int nev = 0;
  double *lam = NULL, *res = NULL;
  pevsl_Parvecs *Y = NULL;
  const int ierr =
      pEVSL_ChebLanNr(p, xintv, *max_its, *tol, &vinit, pol, &nev, &lam, &Y, &res, stdout);
  if (ierr) printf("ChebLanNr error %d\n", ierr);
  pEVSL_ParvecFree(&vinit);
  pEVSL_StatsPrint(p, stdout);
  if (res) PEVSL_FREE(res);
  p->nev_computed = nev;
  p->eval_computed = lam;
  p->evec_computed = Y;
}

void pevsl_get_nev_f90_(uintptr_t *pevslf90, int *nev) { *nev = inst(pevslf90)->nev_computed; }

void pevsl_copy_result_f90_(uintptr_t *pevslf90, double *val, double *vec, int *ld) {
  pevsl_Data *p = inst(pevslf90);
  const int nev = p->nev_computed;
  if (nev > 0) memcpy(val, p->eval_computed, sizeof(double) * (size_t)nev);
  if (p->evec_computed) {
    for (int i = 0; i < nev; i++)
      memcpy(vec + (size_t)i * (*ld), p->evec_computed->data + (size_t)i * p->evec_computed->ld,
             sizeof(double) * (size_t)p->n);
  }
  p->nev_computed = 0;
  PEVSL_FREE(p->eval_computed);
  if (p->evec_computed) {
    pEVSL_ParvecsFree(p->evec_computed);
    PEVSL_FREE(p->evec_computed);
  }
}

void pevsl_setup_chebiter_f90_(double *lmin, double *lmax, int *deg, uintptr_t *parcsrf90,
                               uintptr_t *chebf90) {
  void *cheb = NULL;
  pEVSL_ChebIterSetup(*lmin, *lmax, *deg, mat(parcsrf90), &cheb);
  *chebf90 = reinterpret_cast<uintptr_t>(cheb);
}

void pevsl_chebiter_f90_(int *type, double *b, double *x, uintptr_t *chebf90) {
  void *cheb = reinterpret_cast<void *>(*chebf90);
  if (*type == 1)
    pEVSL_ChebIterSolv1(b, x, cheb);
  else
    pEVSL_ChebIterSolv2(b, x, cheb);
}

void pevsl_chebiterstatsprint_f90_(uintptr_t *chebf90) {
  pEVSL_ChebIterStatsPrint(reinterpret_cast<void *>(*chebf90), stdout);
}

void pevsl_free_chebiterb_f90_(uintptr_t *chebf90) {
  pEVSL_ChebIterFree(reinterpret_cast<void *>(*chebf90));
}

void pevsl_setbsol_chebiter_f90_(uintptr_t *pevslf90, int *type, uintptr_t *chebf90) {
  void *cheb = reinterpret_cast<void *>(*chebf90);
  pEVSL_SetBSol(inst(pevslf90), *type == 1 ? pEVSL_ChebIterSolv1 : pEVSL_ChebIterSolv2, cheb);
}

void pevsl_setup_lspolsqrt_f90_(double *lmin, double *lmax, int *maxdeg, double *tol,
                                uintptr_t *parcsrf90, uintptr_t *lspolf90) {
  void *lspol = NULL;
  pEVSL_SetupLSPolSqrt(*maxdeg, *tol, *lmin, *lmax, mat(parcsrf90), &lspol);
  *lspolf90 = reinterpret_cast<uintptr_t>(lspol);
}

void pevsl_lspolsol_f90_(double *b, double *x, uintptr_t *lspolf90) {
  pEVSL_LSPolSol(b, x, reinterpret_cast<void *>(*lspolf90));
}

void pevsl_free_lspol_f90_(uintptr_t *lspolf90) {
  pEVSL_LSPolFree(reinterpret_cast<void *>(*lspolf90));
}

void pevsl_setltsol_lspol_f90_(uintptr_t *pevslf90, uintptr_t *lspolf90) {
  pEVSL_SetLTSol(inst(pevslf90), pEVSL_LSPolSol, reinterpret_cast<void *>(*lspolf90));
}

void pevsl_kpmdos_ecnt_f90_(uintptr_t *pevslf90, int *Mdeg, int *damping, int *nvec,
                            double *intv, double *mu, double *ecnt) {
  pEVSL_Kpmdos(inst(pevslf90), *Mdeg, *damping, *nvec, intv, 1, 0, MPI_COMM_NULL, mu, ecnt);
}

void pevsl_landos_ecnt_f90_(uintptr_t *pevslf90, int *nvec, int *msteps, int *npts,
                            double *intv, double *ecnt) {
  double *xdos, *ydos;
  PEVSL_MALLOC(xdos, *npts, double);
  PEVSL_MALLOC(ydos, *npts, double);
  pEVSL_LanDosG(inst(pevslf90), *nvec, *msteps, *npts, xdos, ydos, ecnt, intv, 1, 0,
                MPI_COMM_NULL);
  PEVSL_FREE(xdos);
  PEVSL_FREE(ydos);
}

void pevsl_set_sigmamult_f90_(uintptr_t *pevslf90, double *mult) {
  pEVSL_SetSigmaMult(inst(pevslf90), *mult);
}

}  // extern "C"
