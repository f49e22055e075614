/*
 * lap_oracle.c -- CPU oracle driver for the Laplacian spectrum-slicing run.
 * TEST INFRASTRUCTURE ONLY (result checks and CPU baseline; see oracle.h).
 *
 * Follows the flow of TESTS/Lap/LapPLanN.c:15-323 for one process and one
 * group: srand(0) (SRC/pevsl.c:36 with rank 0), KPM DOS (Mdeg 300, Jackson,
 * nvec 60) or Lanczos DOS (dostype 1), equal-count slicing, one random vinit
 * shared by all slices, then per slice FindPol (damping 2, thresh 0.8/0.2,
 * max_deg 3000) and filtered Lanczos with mlan = min(max(5 nev, 300), n),
 * tol 1e-8.  Prints one JSON object.
 *
 * Options: -nx -ny -nz -a -b -nslices -dostype -nvec -Mdeg -seed
 *          -slice s (run only slice s; -1 = all)
 *          -maxsteps K (cap Lanczos steps, for bounded CPU-baseline samples)
 *          -dostype 2: no DOS estimate; one slice [a, b] (nslices forced to
 *                      1), nev taken as -nev (timing samples)
 *          -reorthk K: also time three CGS-DGKS passes against K basis
 *                      columns ("t_cgs_pass" in the output)
 * The output's "threads" is the OpenMP thread count the long loops of the
 * oracle ran on (OMP_NUM_THREADS); results do not depend on it.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <omp.h>

#include "oracle.h"

static double wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int argi(int argc, char **argv, const char *k, int d) {
  for (int i = 1; i < argc - 1; i++)
    if (argv[i][0] == '-' && !strcmp(argv[i] + 1, k)) return atoi(argv[i + 1]);
  return d;
}
static double argd(int argc, char **argv, const char *k, double d) {
  for (int i = 1; i < argc - 1; i++)
    if (argv[i][0] == '-' && !strcmp(argv[i] + 1, k)) return atof(argv[i + 1]);
  return d;
}

static void print_arr(const char *name, const double *x, int n, int last) {
  printf("\"%s\": [", name);
  for (int i = 0; i < n; i++) printf("%s%.17g", i ? ", " : "", x[i]);
  printf("]%s", last ? "" : ", ");
}

int main(int argc, char **argv) {
  int nx = argi(argc, argv, "nx", 20), ny = argi(argc, argv, "ny", 20),
      nz = argi(argc, argv, "nz", 20);
  double a = argd(argc, argv, "a", 0.6), b = argd(argc, argv, "b", 1.2);
  int nslices = argi(argc, argv, "nslices", 5);
  int dostype = argi(argc, argv, "dostype", 0);
  int nvec = argi(argc, argv, "nvec", 60), Mdeg = argi(argc, argv, "Mdeg", 300);
  int seed = argi(argc, argv, "seed", 0);
  int only = argi(argc, argv, "slice", -1);
  int maxsteps = argi(argc, argv, "maxsteps", 0);
  int reorthk = argi(argc, argv, "reorthk", 0);
  if (dostype == 2) nslices = 1;
  double lmin = 0.0, lmax = nz == 1 ? 8.0 : 12.0, tol = 1e-8;
  int n = nx * ny * nz;
  double xintv[4] = {a, b, lmin, lmax};

  orc_csr A;
  orc_lap3d(nx, ny, nz, &A);
  orc_srand((unsigned)seed);

  double t0 = wtime(), ecount = 0;
  double *mu = (double *)calloc((size_t)Mdeg + 1, sizeof(double));
  double *sli = (double *)calloc((size_t)nslices + 1, sizeof(double));
  int ierr;
  if (dostype == 0) {
    ierr = orc_kpmdos(&A, Mdeg, 1, nvec, xintv, mu, &ecount);
    if (ierr) { fprintf(stderr, "kpmdos error %d\n", ierr); return 1; }
    int npts = (int)(10 * ecount);
    ierr = orc_spslicer_kpm(sli, mu, Mdeg, xintv, nslices, npts);
    if (ierr) { fprintf(stderr, "spslicer error %d\n", ierr); return 1; }
  } else if (dostype == 2) {
    sli[0] = a;
    sli[1] = b;
    ecount = argi(argc, argv, "nev", 100);
  } else {
    int msteps = 40, npts = 200;
    double *xd = (double *)calloc(npts, sizeof(double));
    double *yd = (double *)calloc(npts, sizeof(double));
    orc_landos(&A, nvec, msteps, npts, xd, yd, &ecount, xintv, 1.0);
    orc_spslicer_lan(xd, yd, nslices, npts, sli);
    free(xd);
    free(yd);
  }
  double t_dos = wtime() - t0;
  int ev_int = (int)(1 + ecount / ((double)nslices));
  double *vinit = (double *)malloc(sizeof(double) * (size_t)n);
  orc_rand_vec(n, vinit);

  printf("{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"n\": %d, \"nnz\": %d, ", nx, ny,
         nz, n, A.ia[n]);
  printf("\"a\": %.17g, \"b\": %.17g, \"nslices\": %d, \"dostype\": %d, ", a, b,
         nslices, dostype);
  printf("\"ecount\": %.17g, \"t_dos\": %.6f, ", ecount, t_dos);
  if (dostype == 0) print_arr("mu", mu, Mdeg + 1, 0);
  print_arr("sli", sli, nslices + 1, 0);
  printf("\"slices\": [");
  int first = 1;
  for (int sl = 0; sl < nslices; sl++) {
    if (only >= 0 && sl != only) continue;
    int nev = ev_int + 2;
    int mlan = 5 * nev > 300 ? 5 * nev : 300;
    if (mlan > n) mlan = n;
    if (maxsteps > 0 && mlan > maxsteps) mlan = maxsteps;
    double intv[4] = {sli[sl], sli[sl + 1], lmin, lmax};
    orc_pol pol;
    orc_set_pol_def(&pol);
    pol.damping = 2;
    pol.thresh_int = 0.8;
    pol.thresh_ext = 0.2;
    pol.max_deg = 3000;
    double tp = wtime();
    orc_find_pol(intv, &pol);
    tp = wtime() - tp;
    double *lam = (double *)malloc(sizeof(double) * (size_t)mlan);
    double *res = (double *)malloc(sizeof(double) * (size_t)mlan);
    int nev2 = 0;
    orc_lan_info info;
    double tl = wtime();
    ierr = orc_cheblan_nr(&A, intv, mlan, tol, vinit, &pol, &nev2, lam, res,
                          NULL, &info);
    tl = wtime() - tl;
    int *ind = (int *)malloc(sizeof(int) * (size_t)(nev2 > 0 ? nev2 : 1));
    orc_sort_double(nev2, lam, ind);
    double *rs = (double *)malloc(sizeof(double) * (size_t)(nev2 > 0 ? nev2 : 1));
    for (int i = 0; i < nev2; i++) rs[i] = res[ind[i]];
    int nex = orc_exact_eig_lap3(nx, ny, nz, sli[sl], sli[sl + 1], NULL, 0);
    printf("%s{\"slice\": %d, \"intv\": [%.17g, %.17g], \"ierr\": %d, ",
           first ? "" : ", ", sl, sli[sl], sli[sl + 1], ierr);
    printf("\"deg\": %d, \"type\": %d, \"bar\": %.17g, \"gam\": %.17g, ", pol.deg,
           pol.type, pol.bar, pol.gam);
    print_arr("pol_mu", pol.mu, pol.deg + 1, 0);
    printf("\"mlan\": %d, \"its\": %d, \"kdim\": %d, \"nbreak\": %d, "
           "\"nconv\": %d, \"nmatvec\": %d, \"tr\": %.17g, ",
           mlan, info.its, info.kdim, info.nbreak, info.nconv, info.nmatvec,
           info.tr_last);
    printf("\"nev\": %d, \"nev_exact\": %d, \"t_findpol\": %.6f, "
           "\"t_lan\": %.6f, ",
           nev2, nex, tp, tl);
    print_arr("lam", lam, nev2, 0);
    print_arr("res", rs, nev2, 1);
    printf("}");
    fflush(stdout);
    first = 0;
    free(lam);
    free(res);
    free(rs);
    free(ind);
    orc_free_pol(&pol);
  }
  printf("]");
  if (reorthk > 0) {
    double *V = (double *)malloc(sizeof(double) * (size_t)n * reorthk);
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *w = (double *)malloc(sizeof(double) * (size_t)reorthk);
    orc_rand_vec(n, v);
    for (size_t i = 0; i < (size_t)n * reorthk; i++) V[i] = 1e-3 * v[i % n] + 1e-6;
    double nrm, tc = wtime();
    for (int r = 0; r < 3; r++) {
      orc_rand_vec(n, v);
      orc_cgs_dgks(n, reorthk, 1, V, n, v, &nrm, w);
    }
    printf(", \"reorthk\": %d, \"t_cgs_pass\": %.6f", reorthk, (wtime() - tc) / 3);
    free(V);
    free(v);
    free(w);
  }
  printf(", \"threads\": %d, \"t_total\": %.6f}\n", omp_get_max_threads(), wtime() - t0);
  free(mu);
  free(sli);
  free(vinit);
  orc_csr_free(&A);
  return 0;
}
