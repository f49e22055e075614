/* pevsl_f90.h -- the Fortran 90 interface (FORTRAN/pevsl_f90.c): routines
 * named <name>_f90_ (PEVSL_FORT, INC/pevsl_def.h:108), all arguments by
 * reference, objects as uintptr_t handles (declare INTEGER*8 on the Fortran
 * side).  pevsl_cheblannr keeps its results in the instance for
 * pevsl_get_nev / pevsl_copy_result (eigenvalues, and eigenvectors with
 * leading dimension *ld).  Declared here for C callers and tests; a Fortran
 * program declares them as external subroutines. */
#ifndef PEVSL_F90_H
#define PEVSL_F90_H

#include <stdint.h>

#include "pevsl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FORTRAN/pevsl_f90.c:16 */
void pevsl_start_f90_(MPI_Fint *Fcomm, uintptr_t *pevslf90);
/* FORTRAN/pevsl_f90.c:28 */
void pevsl_finish_f90_(uintptr_t *pevslf90);
/* FORTRAN/pevsl_f90.c:37 */
void pevsl_seta_parcsr_f90_(uintptr_t *pevslf90, uintptr_t *Af90);
/* FORTRAN/pevsl_f90.c:47 */
void pevsl_setb_parcsr_f90_(uintptr_t *pevslf90, uintptr_t *Bf90);
/* FORTRAN/pevsl_f90.c:63 */
void pevsl_setprobsizes_f90_(uintptr_t *pevslf90, int *N, int *n, int *nfirst);
/* FORTRAN/pevsl_f90.c:76 */
void pevsl_setamv_f90_(uintptr_t *pevslf90, void *func, void *data);
/* FORTRAN/pevsl_f90.c:89 */
void pevsl_setbmv_f90_(uintptr_t *pevslf90, void *func, void *data);
/* FORTRAN/pevsl_f90.c:98 */
void pevsl_setbsol_f90_(uintptr_t *pevslf90, void *func, void *data);
/* FORTRAN/pevsl_f90.c:107 */
void pevsl_set_stdeig_f90_(uintptr_t *pevslf90);
/* FORTRAN/pevsl_f90.c:116 */
void pevsl_set_geneig_f90_(uintptr_t *pevslf90);
/* FORTRAN/pevsl_f90.c:134 */
void pevsl_parcsrcreate_f90_(int *nrow, int *ncol, int *row_starts, int *col_starts, int *ia,
                             int *ja, double *aa, MPI_Fint *Fcomm, uintptr_t *matf90);
/* FORTRAN/pevsl_f90.c:159 */
void pevsl_parcsr_free_f90_(uintptr_t *Af90);
/* FORTRAN/pevsl_f90.c:168 */
void pevsl_parcsrmatvec_f90_(double *x, double *y, uintptr_t *matf90);
/* FORTRAN/pevsl_f90.c:176 */
void pevsl_bsv_f90_(uintptr_t *pevslf90, double *b, double *x);
/* FORTRAN/pevsl_f90.c:193 */
void pevsl_lanbounds_f90_(uintptr_t *pevslf90, int *mlan, int *nsteps, double *tol,
                          double *lmin, double *lmax);
/* FORTRAN/pevsl_f90.c:225 */
void pevsl_findpol_f90_(double *xintv, double *thresh_int, double *thresh_ext,
                        uintptr_t *polf90);
/* FORTRAN/pevsl_f90.c:243 */
void pevsl_freepol_f90_(uintptr_t *polf90);
/* FORTRAN/pevsl_f90.c:253 */
void pevsl_cheblannr_f90_(uintptr_t *pevslf90, double *xintv, int *max_its, double *tol,
                          uintptr_t *polf90);
/* FORTRAN/pevsl_f90.c:296 */
void pevsl_get_nev_f90_(uintptr_t *pevslf90, int *nev);
/* FORTRAN/pevsl_f90.c:311 */
void pevsl_copy_result_f90_(uintptr_t *pevslf90, double *val, double *vec, int *ld);
/* FORTRAN/pevsl_f90.c:333 */
void pevsl_setup_chebiter_f90_(double *lmin, double *lmax, int *deg, uintptr_t *parcsrf90,
                               uintptr_t *chebf90);
/* FORTRAN/pevsl_f90.c:344 */
void pevsl_chebiter_f90_(int *type, double *b, double *x, uintptr_t *chebf90);
/* FORTRAN/pevsl_f90.c:357 */
void pevsl_chebiterstatsprint_f90_(uintptr_t *chebf90);
/* FORTRAN/pevsl_f90.c:367 */
void pevsl_free_chebiterb_f90_(uintptr_t *chebf90);
/* FORTRAN/pevsl_f90.c:377 */
void pevsl_setbsol_chebiter_f90_(uintptr_t *pevslf90, int *type, uintptr_t *chebf90);
/* FORTRAN/pevsl_f90.c:391 */
void pevsl_setup_lspolsqrt_f90_(double *lmin, double *lmax, int *maxdeg, double *tol,
                                uintptr_t *parcsrf90, uintptr_t *lspolf90);
/* FORTRAN/pevsl_f90.c:402 */
void pevsl_lspolsol_f90_(double *b, double *x, uintptr_t *lspolf90);
/* FORTRAN/pevsl_f90.c:410 */
void pevsl_free_lspol_f90_(uintptr_t *lspolf90);
/* FORTRAN/pevsl_f90.c:420 */
void pevsl_setltsol_lspol_f90_(uintptr_t *pevslf90, uintptr_t *lspolf90);
/* FORTRAN/pevsl_f90.c:429 */
void pevsl_kpmdos_ecnt_f90_(uintptr_t *pevslf90, int *Mdeg, int *damping, int *nvec,
                            double *intv, double *mu, double *ecnt);
/* FORTRAN/pevsl_f90.c:437 */
void pevsl_landos_ecnt_f90_(uintptr_t *pevslf90, int *nvec, int *msteps, int *npts,
                            double *intv, double *ecnt);
/* FORTRAN/pevsl_f90.c:453 */
void pevsl_set_sigmamult_f90_(uintptr_t *pevslf90, double *mult);

#ifdef __cplusplus
}
#endif

#endif
