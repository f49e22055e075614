/*
 * genlap_plan.c -- C driver of the generalized spectrum-slicing run, written
 * against include/ only and linked with libpevsl_amd.so.  It makes the
 * sequence of API calls of the reference driver TESTS/Gen_Lap/LapPLanN.c
 * (KPM branch, one process, one group): A = nx x ny x nz Laplacian, B =
 * ParcsrLaplace(n, 1, 1) = tridiag(-1, 4, -1), the L^T solve by
 * pEVSL_SetupLSPolSqrt(50, 1e-6, 2, 6), SetGenEig, LanTrbounds(50, 200,
 * 1e-8), Kpmdos(Mdeg 50, nvec 60), SpslicerKpm, then FindPol (damping 2,
 * thresholds 0.6/0.15, max_deg 3000) and ChebLanNr per slice.
 *
 * The reference's B-solve is a MUMPS factorization (an external library).
 * Here -bsol picks it:
 *   tridiag   exact host solve of the tridiagonal B (Thomas algorithm): a
 *             caller-supplied direct solver through pEVSL_SetBSol, the place
 *             MUMPS takes (host round trip per solve)
 *   lspol     the library's least-squares series of B^{-1} on [2, 6]
 *             (pEVSL_SetupLSPolRec(50, 1e-12, ...)), applied on the device
 *   chebiter  40 steps of Chebyshev iteration on [2, 6], on the device
 *
 * usage: genlap_plan [-nx 8 -ny 8 -nz 8 -a 1.5 -b 2.5 -nslices 4
 *                     -bsol tridiag|lspol|chebiter]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pevsl.h"
#include "pevsl_amd.h"
#include "pevsl_itsol.h"

static const char *args(int argc, char **argv, const char *k) {
  for (int i = 1; i < argc - 1; i++)
    if (argv[i][0] == '-' && !strcmp(argv[i] + 1, k)) return argv[i + 1];
  return NULL;
}
static int argi(int argc, char **argv, const char *k, int d) {
  const char *v = args(argc, argv, k);
  return v ? atoi(v) : d;
}
static double argd(int argc, char **argv, const char *k, double d) {
  const char *v = args(argc, argv, k);
  return v ? atof(v) : d;
}

/* 7-point (5-point for nz == 1) Laplacian, natural ordering, diagonal 6 (4
 * when nz == 1) as TESTS/COMMON/lapl.c builds it */
static void lap_csr(int nx, int ny, int nz, pevsl_Csr *A) {
  int n = nx * ny * nz, nnz = 0;
  pEVSL_CsrResize(n, n, 7 * n, A);
  A->ia[0] = 0;
  for (int r = 0; r < n; r++) {
    int k = r / (nx * ny), i = (r - k * nx * ny) / nx, j = r - k * nx * ny - i * nx;
    int cols[7], m = 0;
    if (k > 0) cols[m++] = r - nx * ny;
    if (i > 0) cols[m++] = r - nx;
    if (j > 0) cols[m++] = r - 1;
    cols[m++] = r;
    if (j < nx - 1) cols[m++] = r + 1;
    if (i < ny - 1) cols[m++] = r + nx;
    if (k < nz - 1) cols[m++] = r + nx * ny;
    for (int q = 0; q < m; q++) {
      A->ja[nnz] = cols[q];
      A->a[nnz] = cols[q] == r ? (nz > 1 ? 6.0 : 4.0) : -1.0;
      nnz++;
    }
    A->ia[r + 1] = nnz;
  }
}

/* x = B^{-1} b for B = tridiag(-1, 4, -1) of order n (Thomas algorithm) */
typedef struct {
  int n;
  double *c;
} Thomas;
static void thomas_solve(double *b, double *x, void *data) {
  const Thomas *T = (const Thomas *)data;
  const int n = T->n;
  double *c = T->c;
  double m = 4.0;
  c[0] = -1.0 / m;
  x[0] = b[0] / m;
  for (int i = 1; i < n; i++) {
    m = 4.0 + c[i - 1];
    c[i] = -1.0 / m;
    x[i] = (b[i] + x[i - 1]) / m;
  }
  for (int i = n - 2; i >= 0; i--) x[i] -= c[i] * x[i + 1];
}

int main(int argc, char **argv) {
  const int nx = argi(argc, argv, "nx", 8), ny = argi(argc, argv, "ny", 8),
            nz = argi(argc, argv, "nz", 8);
  const double a = argd(argc, argv, "a", 1.5), b = argd(argc, argv, "b", 2.5);
  int nslices = argi(argc, argv, "nslices", 4);
  const char *bsol = args(argc, argv, "bsol");
  if (!bsol) bsol = "tridiag";
  if (nslices < 1) nslices = 1;
  const int n = nx * ny * nz;
  printf("Laplacian A : %d x %d x %d, n = %d\nLaplacian B : %d x 1, n = %d\n", nx, ny, nz, n, n,
         n);
  printf("Interval: [%20.15f, %20.15f]  -- %d slices, B-solve %s\n", a, b, nslices, bsol);
  pevsl_Csr Aloc, Bloc;
  lap_csr(nx, ny, nz, &Aloc);
  lap_csr(n, 1, 1, &Bloc);
  pevsl_Parcsr A, B;
  pEVSL_ParcsrCreate(n, n, NULL, NULL, &A, MPI_COMM_WORLD);
  pEVSL_ParcsrSetup(&Aloc, &A);
  pEVSL_ParcsrCreate(n, n, NULL, NULL, &B, MPI_COMM_WORLD);
  pEVSL_ParcsrSetup(&Bloc, &B);
  pEVSL_FreeCsr(&Aloc);
  pEVSL_FreeCsr(&Bloc);

  pevsl_Data *pevsl;
  pEVSL_Start(MPI_COMM_WORLD, &pevsl);
  pEVSL_SetAParcsr(pevsl, &A);
  pEVSL_SetBParcsr(pevsl, &B);
  Thomas th = {n, (double *)malloc(sizeof(double) * n)};
  void *bdata = NULL, *cheb = NULL, *Bpol = NULL;
  if (!strcmp(bsol, "lspol")) {
    pEVSL_SetupLSPolRec(50, 1e-12, 2.0, 6.0, &B, &bdata);
    pEVSL_SetBSol(pevsl, pEVSL_LSPolSol, bdata);
  } else if (!strcmp(bsol, "chebiter")) {
    pEVSL_ChebIterSetup(2.0, 6.0, 40, &B, &cheb);
    pEVSL_SetBSol(pevsl, pEVSL_ChebIterSolv2, cheb);
  } else {
    pEVSL_SetBSol(pevsl, thomas_solve, &th);
  }
  pEVSL_SetupLSPolSqrt(50, 1e-6, 2.0, 6.0, &B, &Bpol);
  pEVSL_SetLTSol(pevsl, pEVSL_LSPolSol, Bpol);
  pEVSL_SetGenEig(pevsl);

  pevsl_Parvec vinit;
  pEVSL_ParvecCreate(A.ncol_global, A.ncol_local, A.first_col, MPI_COMM_WORLD, &vinit);
  pEVSL_ParvecRand(&vinit);
  double lmin, lmax, ecount, t0 = pEVSL_Wtime();
  int ierr = pEVSL_LanTrbounds(pevsl, 50, 200, 1e-8, &vinit, 1, &lmin, &lmax, NULL);
  if (ierr) { printf("LanTrbounds error %d\n", ierr); return 1; }
  printf("Step 0: Eigenvalue bounds for B^{-1}*A: [%.15e, %.15e] (%.3f s)\n", lmin, lmax,
         pEVSL_Wtime() - t0);
  double xintv[4] = {a, b, lmin, lmax};
  const int Mdeg = 50, nvec = 60;
  double *mu = (double *)malloc((Mdeg + 1) * sizeof(double));
  ierr = pEVSL_Kpmdos(pevsl, Mdeg, 1, nvec, xintv, 1, 0, MPI_COMM_NULL, mu, &ecount);
  if (ierr) { printf("kpmdos error %d\n", ierr); return 1; }
  printf(" KPMDOS: estimated eig count in interval: %f \n", ecount);
  double *sli = (double *)malloc((nslices + 1) * sizeof(double));
  ierr = pEVSL_SpslicerKpm(sli, mu, Mdeg, xintv, nslices, (int)(10 * ecount));
  if (ierr) { printf("spslicer error %d\n", ierr); return 1; }
  const int ev_int = (int)(1 + ecount / ((double)nslices));
  int nev = ev_int + 2, mlan = 5 * nev > 300 ? 5 * nev : 300, total = 0;
  if (mlan > n) mlan = n;
  t0 = pEVSL_Wtime();
  for (int sl = 0; sl < nslices; sl++) {
    double iv[4] = {sli[sl], sli[sl + 1], lmin, lmax};
    pevsl_Polparams pol;
    pEVSL_SetPolDef(&pol);
    pol.damping = 2;
    pol.thresh_int = 0.6;
    pol.thresh_ext = 0.15;
    pol.max_deg = 3000;
    pEVSL_FindPol(iv, &pol);
    int nev2;
    double *lam, *res;
    pevsl_Parvecs *Y;
    ierr = pEVSL_ChebLanNr(pevsl, iv, mlan, 1e-8, &vinit, &pol, &nev2, &lam, &Y, &res, NULL);
    if (ierr) { printf("ChebLanNr error %d\n", ierr); return 1; }
    int *ind = (int *)malloc((nev2 > 0 ? nev2 : 1) * sizeof(int));
    sort_double(nev2, lam, ind);
    double rmax = 0.0;
    for (int i = 0; i < nev2; i++) rmax = res[i] > rmax ? res[i] : rmax;
    printf("slice %2d [%.10f, %.10f] deg %4d: found %4d, max res %.2e", sl, iv[0], iv[1],
           pol.deg, nev2, rmax);
    if (nev2) printf("  first %.15e last %.15e", lam[0], lam[nev2 - 1]);
    printf("\n");
    total += nev2;
    free(lam);
    free(res);
    free(ind);
    if (Y) {
      pEVSL_ParvecsFree(Y);
      free(Y);
    }
    pEVSL_FreePol(&pol);
  }
  printf("total %d eigenvalues in [%g, %g], %.3f s\n", total, a, b, pEVSL_Wtime() - t0);
  free(sli);
  free(mu);
  pEVSL_ParvecFree(&vinit);
  pEVSL_Finish(pevsl);
  if (bdata) pEVSL_LSPolFree(bdata);
  if (cheb) pEVSL_ChebIterFree(cheb);
  pEVSL_LSPolFree(Bpol);
  free(th.c);
  pEVSL_ParcsrFree(&A);
  pEVSL_ParcsrFree(&B);
  return 0;
}
