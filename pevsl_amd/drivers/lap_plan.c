/*
 * lap_plan.c -- C driver of the Laplacian spectrum-slicing run written
 * against include/pevsl.h only (plus optional pevsl_amd.h switches), linked
 * with libpevsl_amd.so.  It makes the same sequence of API calls as the
 * reference driver TESTS/Lap/LapPLanN.c:118-296 (ParcsrCreate/Setup,
 * Start, SetAParcsr, Kpmdos, SpslicerKpm, ParvecRand, SetPolDef/FindPol,
 * ChebLanNr, sort_double, FreePol, ParvecsFree, Finish) with one process
 * and one group, and prints per slice the computed and exact counts.
 *
 * With -mlp <weights> the slices come from the learned slicer
 * (pEVSL_SpslicerMlp); the full KPM run then judges them
 * (pEVSL_SlicerJudge) and the driver falls back to pEVSL_SpslicerKpm when
 * the largest full-KPM slice count exceeds the mean by more than 15%.
 *
 * With -bounds 1 the spectrum bounds [lmin, lmax] come from
 * pEVSL_LanTrbounds (thick-restart Lanczos, as TESTS/Gen_Lap/LapPLanN.c:138
 * and TESTS/MM/MMPLanN.c:111 do) instead of the Laplacian's hand-set
 * [0, 12] (TESTS/Lap/LapPLanN.c:86-87).
 *
 * usage: lap_plan [-nx 20 -ny 20 -nz 20 -a 0.6 -b 1.2 -nslices 5
 *                  -mixed 0 -batch 0 -mlp weights.bin -bounds 0]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pevsl.h"
#include "pevsl_amd.h"

static int argi(int argc, char **argv, const char *k, int d) {
  for (int i = 1; i < argc - 1; i++)
    if (argv[i][0] == '-' && !strcmp(argv[i] + 1, k)) return atoi(argv[i + 1]);
  return d;
}
static const char *args(int argc, char **argv, const char *k) {
  for (int i = 1; i < argc - 1; i++)
    if (argv[i][0] == '-' && !strcmp(argv[i] + 1, k)) return argv[i + 1];
  return NULL;
}
static double argd(int argc, char **argv, const char *k, double d) {
  for (int i = 1; i < argc - 1; i++)
    if (argv[i][0] == '-' && !strcmp(argv[i] + 1, k)) return atof(argv[i + 1]);
  return d;
}

/* local rows of the 7-point Laplacian, natural ordering */
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

static int exact_count(int nx, int ny, int nz, double a, double b) {
  int c = 0;
  for (int i = 1; i <= nx; i++)
    for (int j = 1; j <= ny; j++)
      for (int k = 1; k <= nz; k++) {
        double tx = sin(i * 0.5 * M_PI / (nx + 1.0)), ty = sin(j * 0.5 * M_PI / (ny + 1.0));
        double tz = nz == 1 ? 0.0 : sin(k * 0.5 * M_PI / (nz + 1.0));
        double ev = 4 * (tx * tx + ty * ty + tz * tz);
        if (ev >= a - 2.220446049250313e-16 && ev <= b + 2.220446049250313e-16) c++;
      }
  return c;
}

int main(int argc, char **argv) {
  int nx = argi(argc, argv, "nx", 20), ny = argi(argc, argv, "ny", 20),
      nz = argi(argc, argv, "nz", 20), nslices = argi(argc, argv, "nslices", 5);
  int mixed = argi(argc, argv, "mixed", 0), batch = argi(argc, argv, "batch", 0);
  if (nslices < 1 || nslices > 100000) nslices = 1;
  double a = argd(argc, argv, "a", 0.6), b = argd(argc, argv, "b", 1.2);
  double lmin = 0.0, lmax = nz == 1 ? 8.0 : 12.0, tol = 1e-8, ecount;
  double xintv[4] = {a, b, lmin, lmax};
  int n = nx * ny * nz, Mdeg = 300, nvec = 60, ierr;

  pevsl_Csr Aloc;
  lap_csr(nx, ny, nz, &Aloc);
  pevsl_Parcsr A;
  pEVSL_ParcsrCreate(n, n, NULL, NULL, &A, MPI_COMM_WORLD);
  pEVSL_ParcsrSetup(&Aloc, &A);
  pEVSL_FreeCsr(&Aloc);

  pevsl_Data *pevsl;
  pEVSL_Start(MPI_COMM_WORLD, &pevsl);
  pEVSL_SetAParcsr(pevsl, &A);
  if (mixed) pEVSL_SetFilterPrecision(pevsl, PEVSL_PREC_FP32);

  if (argi(argc, argv, "bounds", 0)) {
    pevsl_Parvec vb;
    pEVSL_ParvecCreate(A.ncol_global, A.ncol_local, A.first_col, MPI_COMM_WORLD, &vb);
    pEVSL_ParvecRand(&vb);
    double t0b = pEVSL_Wtime();
    ierr = pEVSL_LanTrbounds(pevsl, 50, 3000, 1e-8, &vb, 1, &lmin, &lmax, NULL);
    pEVSL_ParvecFree(&vb);
    if (ierr) { printf("LanTrbounds error %d\n", ierr); return 1; }
    printf("LanTrbounds: [%.10f, %.10f] (%.3f s)\n", lmin, lmax, pEVSL_Wtime() - t0b);
    xintv[2] = lmin;
    xintv[3] = lmax;
  }
  double *mu = (double *)malloc((Mdeg + 1) * sizeof(double));
  double *sli = (double *)malloc((nslices + 1) * sizeof(double));
  const char *mlp_path = args(argc, argv, "mlp");
  double t0 = pEVSL_Wtime();
  if (mlp_path) {
    pevsl_Mlp *net;
    double ecount_mlp;
    if (pEVSL_MlpLoad(mlp_path, &net)) { printf("cannot load %s\n", mlp_path); return 1; }
    ierr = pEVSL_SpslicerMlp(pevsl, net, xintv, nslices, sli, &ecount_mlp);
    pEVSL_MlpFree(net);
    if (ierr) { printf("SpslicerMlp error %d\n", ierr); return 1; }
    printf("learned slicer: count %.1f (%.3f s)\n", ecount_mlp, pEVSL_Wtime() - t0);
  }
  ierr = pEVSL_Kpmdos(pevsl, Mdeg, 1, nvec, xintv, 1, 0, MPI_COMM_NULL, mu, &ecount);
  if (ierr) { printf("kpmdos error %d\n", ierr); return 1; }
  int npts = 10 * ecount;
  int use_kpm = 1;
  if (mlp_path) {
    double *cnt = (double *)malloc(nslices * sizeof(double)), imb;
    pEVSL_SlicerJudge(mu, Mdeg, xintv, (double)n, nslices, sli, cnt, &imb);
    use_kpm = imb > 0.15;
    printf("judge (full KPM): largest slice %.1f%% above the mean -> %s\n", 100 * imb,
           use_kpm ? "fall back to SpslicerKpm" : "accepted");
    free(cnt);
  }
  if (use_kpm) {
    ierr = pEVSL_SpslicerKpm(sli, mu, Mdeg, xintv, nslices, npts);
    if (ierr) { printf("spslicer error %d\n", ierr); return 1; }
  }
  printf("KPM count %.4f, slices:", ecount);
  for (int j = 0; j <= nslices; j++) printf(" %.6f", sli[j]);
  printf("  (%.3f s)\n", pEVSL_Wtime() - t0);
  int ev_int = (int)(1 + ecount / ((double)nslices));

  pevsl_Parvec vinit;
  pEVSL_ParvecCreate(A.ncol_global, A.ncol_local, A.first_col, MPI_COMM_WORLD, &vinit);
  pEVSL_ParvecRand(&vinit);

  int nev = ev_int + 2, mlan = nev * 5 > 300 ? nev * 5 : 300, total = 0, bad = 0;
  if (mlan > n) mlan = n;
  pevsl_Polparams *pols = (pevsl_Polparams *)calloc(nslices, sizeof(pevsl_Polparams));
  for (int sl = 0; sl < nslices; sl++) {
    double iv[4] = {sli[sl], sli[sl + 1], lmin, lmax};
    pEVSL_SetPolDef(&pols[sl]);
    pols[sl].damping = 2;
    pols[sl].thresh_int = 0.8;
    pols[sl].thresh_ext = 0.2;
    pols[sl].max_deg = 3000;
    pEVSL_FindPol(iv, &pols[sl]);
  }
  pEVSL_SetKeepVectors(pevsl, 0);
  t0 = pEVSL_Wtime();
  if (batch) {
    double *intvs = (double *)malloc(4 * nslices * sizeof(double));
    int *maxits = (int *)malloc(nslices * sizeof(int)), *nevs = (int *)malloc(nslices * sizeof(int));
    double **lams = (double **)malloc(nslices * sizeof(double *));
    double **ress = (double **)malloc(nslices * sizeof(double *));
    pevsl_Parvecs **Ws = (pevsl_Parvecs **)malloc(nslices * sizeof(pevsl_Parvecs *));
    for (int sl = 0; sl < nslices; sl++) {
      intvs[4 * sl] = sli[sl];
      intvs[4 * sl + 1] = sli[sl + 1];
      intvs[4 * sl + 2] = lmin;
      intvs[4 * sl + 3] = lmax;
      maxits[sl] = mlan;
    }
    ierr = pEVSL_ChebLanNrBatch(pevsl, nslices, intvs, maxits, tol, &vinit, pols, nevs,
                                lams, Ws, ress, NULL);
    if (ierr) { printf("ChebLanNrBatch error %d\n", ierr); return 1; }
    for (int sl = 0; sl < nslices; sl++) {
      int ex = exact_count(nx, ny, nz, sli[sl], sli[sl + 1]);
      printf("slice %2d [%.6f, %.6f] deg %4d: found %5d exact %5d\n", sl, sli[sl],
             sli[sl + 1], pols[sl].deg, nevs[sl], ex);
      total += nevs[sl];
      bad += nevs[sl] != ex;
      free(lams[sl]);
      free(ress[sl]);
    }
    free(intvs); free(maxits); free(nevs); free(lams); free(ress); free(Ws);
  } else {
    for (int sl = 0; sl < nslices; sl++) {
      double iv[4] = {sli[sl], sli[sl + 1], lmin, lmax};
      int nev2;
      double *lam, *res;
      pevsl_Parvecs *Y;
      ierr = pEVSL_ChebLanNr(pevsl, iv, mlan, tol, &vinit, &pols[sl], &nev2, &lam, &Y,
                             &res, NULL);
      if (ierr) { printf("ChebLanNr error %d\n", ierr); return 1; }
      int *ind = (int *)malloc((nev2 > 0 ? nev2 : 1) * sizeof(int));
      sort_double(nev2, lam, ind);
      int ex = exact_count(nx, ny, nz, sli[sl], sli[sl + 1]);
      printf("slice %2d [%.6f, %.6f] deg %4d: found %5d exact %5d  lam[0] %.15f\n", sl,
             sli[sl], sli[sl + 1], pols[sl].deg, nev2, ex, nev2 ? lam[0] : 0.0);
      total += nev2;
      bad += nev2 != ex;
      free(lam);
      free(res);
      free(ind);
      if (Y) {
        pEVSL_ParvecsFree(Y);
        free(Y);
      }
    }
  }
  printf("total %d eigenvalues, %d slice(s) with a count different from the exact one, "
         "%.3f s\n", total, bad, pEVSL_Wtime() - t0);
  for (int sl = 0; sl < nslices; sl++) pEVSL_FreePol(&pols[sl]);
  free(pols);
  free(sli);
  free(mu);
  pEVSL_ParvecFree(&vinit);
  pEVSL_ParcsrFree(&A);
  pEVSL_Finish(pevsl);
  return bad ? 3 : 0;
}
