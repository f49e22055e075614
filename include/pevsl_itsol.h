/* pevsl_itsol.h -- Chebyshev iteration for the B-solves of a generalized
 * problem: the interface of EXTERNAL/ITERSOL/pevsl_itsol.h (chebiter.c:
 * 64-308), implemented on the device.  Setup takes bounds [lmin, lmax] of B's
 * spectrum and the number of steps; Solv1 is the Templates form, Solv2 the
 * form of Saad (2003), p. 399; both are SVFunc solves (host arrays) that can
 * be installed with pEVSL_SetBSol, where they run on device vectors.
 * GetRes returns NULL (no convergence history is kept, as with the
 * reference's default SAVE_CONV_HIST 0).  StatsPrint reports the device time
 * of the solves (products and vector updates) and the number of products. */
#ifndef PEVSL_ITSOL_H
#define PEVSL_ITSOL_H

#include <stdio.h>

#include "pevsl_struct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* EXTERNAL/ITERSOL/pevsl_itsol.h:15 */
int pEVSL_ChebIterSetup(double lmin, double lmax, int deg, pevsl_Parcsr *A, void **data);
/* EXTERNAL/ITERSOL/pevsl_itsol.h:18 */
void pEVSL_ChebIterSolv1(double *db, double *dx, void *data);
/* EXTERNAL/ITERSOL/pevsl_itsol.h:20 */
void pEVSL_ChebIterSolv2(double *db, double *dx, void *data);
/* EXTERNAL/ITERSOL/pevsl_itsol.h:22 */
void pEVSL_ChebIterFree(void *vdata);
/* EXTERNAL/ITERSOL/pevsl_itsol.h:24 */
double *pEVSL_ChebIterGetRes(void *data);
/* EXTERNAL/ITERSOL/pevsl_itsol.h:26 */
void pEVSL_ChebIterStatsPrint(void *data, FILE *fstats);

#ifdef __cplusplus
}
#endif

#endif
