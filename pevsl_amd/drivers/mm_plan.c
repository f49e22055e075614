/*
 * mm_plan.c -- C driver of the Matrix Market spectrum-slicing run, written
 * against include/pevsl.h only and linked with libpevsl_amd.so.  It makes
 * the sequence of API calls of the reference driver TESTS/MM/MMPLanN.c (KPM
 * branch, one process, one group): read the matrix (the rules of
 * TESTS/COMMON/parmm.c: MM1/MM0 indexing, a symmetric file's triangle
 * mirrored), pEVSL_ParcsrCreate/Setup, Start, SetAParcsr, ParvecRand,
 * LanTrbounds(50, 200, 1e-8), Kpmdos(Mdeg 50, nvec 60), SpslicerKpm,
 * FindPol (damping 2, thresholds 0.6/0.15, max_deg 3000) and ChebLanNr per
 * slice, and prints each slice's count.
 *
 * usage: mm_plan <file.mtx> <a> <b> <nslices> [idxin = 1]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pevsl.h"

/* COO triplets of a sparse real/integer coordinate Matrix Market file;
 * returns 0 on success, 1-4 for an unreadable or unsupported file, 5 for an
 * index outside the declared size */
static int read_mm(const char *fn, int idxin, pevsl_Coo *coo) {
  idxin = idxin > 0; /* any positive value means 1-based, as parmm.c:17 */
  FILE *f = fopen(fn, "r");
  if (!f) return 1;
  char line[1024], obj[64], fmt[64], field[64], symm[64];
  if (!fgets(line, sizeof line, f) ||
      sscanf(line, "%%%%MatrixMarket %63s %63s %63s %63s", obj, fmt, field, symm) != 4 ||
      strcmp(fmt, "coordinate") != 0 ||
      (strcmp(field, "real") != 0 && strcmp(field, "integer") != 0)) {
    fclose(f);
    return 2;
  }
  const int sym = strcmp(symm, "symmetric") == 0;
  do {
    if (!fgets(line, sizeof line, f)) { fclose(f); return 3; }
  } while (line[0] == '%');
  int nrow, ncol, nnz;
  if (sscanf(line, "%d %d %d", &nrow, &ncol, &nnz) != 3) { fclose(f); return 3; }
  pEVSL_CooResize(nrow, ncol, sym ? 2 * nnz : nnz, coo);
  int m = 0;
  for (int k = 0; k < nnz; k++) {
    double r, c, v;
    if (!fgets(line, sizeof line, f) || sscanf(line, "%lf %lf %lf", &r, &c, &v) != 3) {
      fclose(f);
      return 4;
    }
    const int ri = (int)r - idxin, ci = (int)c - idxin;
    if (ri < 0 || ri >= nrow || ci < 0 || ci >= ncol) {
      fclose(f);
      return 5; /* index outside the declared size */
    }
    coo->ir[m] = ri;
    coo->jc[m] = ci;
    coo->vv[m++] = v;
    if (sym && ri != ci) {
      coo->ir[m] = ci;
      coo->jc[m] = ri;
      coo->vv[m++] = v;
    }
  }
  coo->nnz = m;
  fclose(f);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 5) {
    printf("usage: %s <file.mtx> <a> <b> <nslices> [idxin]\n", argv[0]);
    return 1;
  }
  const double a = atof(argv[2]), b = atof(argv[3]);
  int nslices = atoi(argv[4]);
  const int idxin = argc > 5 ? atoi(argv[5]) : 1;
  if (nslices < 1 || nslices > 100000) nslices = 1;
  pevsl_Coo coo;
  int ierr = read_mm(argv[1], idxin, &coo);
  if (ierr) { printf("cannot read %s (%d)\n", argv[1], ierr); return 1; }
  pevsl_Csr Aloc;
  ierr = pEVSL_CooToCsr(0, &coo, &Aloc);
  pEVSL_FreeCoo(&coo);
  if (ierr) { printf("CooToCsr error %d\n", ierr); return 1; }
  const int n = Aloc.nrows;
  pevsl_Parcsr A;
  pEVSL_ParcsrCreate(n, n, NULL, NULL, &A, MPI_COMM_WORLD);
  pEVSL_ParcsrSetup(&Aloc, &A);
  pEVSL_FreeCsr(&Aloc);

  pevsl_Data *pevsl;
  pEVSL_Start(MPI_COMM_WORLD, &pevsl);
  pEVSL_SetAParcsr(pevsl, &A);
  pevsl_Parvec vinit;
  pEVSL_ParvecCreate(A.ncol_global, A.ncol_local, A.first_col, MPI_COMM_WORLD, &vinit);
  pEVSL_ParvecRand(&vinit);
  double lmin, lmax, ecount, t0 = pEVSL_Wtime();
  ierr = pEVSL_LanTrbounds(pevsl, 50, 200, 1e-8, &vinit, 1, &lmin, &lmax, NULL);
  if (ierr) { printf("LanTrbounds error %d\n", ierr); return 1; }
  printf("n %d, bounds [%.15e, %.15e] (%.3f s)\n", n, lmin, lmax, pEVSL_Wtime() - t0);
  double xintv[4] = {a, b, lmin, lmax};
  const int Mdeg = 50, nvec = 60;
  double *mu = (double *)malloc((Mdeg + 1) * sizeof(double));
  ierr = pEVSL_Kpmdos(pevsl, Mdeg, 1, nvec, xintv, 1, 0, MPI_COMM_NULL, mu, &ecount);
  if (ierr) { printf("kpmdos error %d\n", ierr); return 1; }
  double *sli = (double *)malloc((nslices + 1) * sizeof(double));
  ierr = pEVSL_SpslicerKpm(sli, mu, Mdeg, xintv, nslices, (int)(10 * ecount));
  if (ierr) { printf("spslicer error %d\n", ierr); return 1; }
  printf("KPM count %.4f\n", ecount);
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
    printf("slice %2d [%.10f, %.10f] deg %4d: found %4d", sl, iv[0], iv[1], pol.deg, nev2);
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
  pEVSL_ParcsrFree(&A);
  pEVSL_Finish(pevsl);
  return 0;
}
