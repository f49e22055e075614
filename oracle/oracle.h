/*
 * oracle.h -- CPU oracle for the pEVSL spectrum-slicing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64 restatement of the
 * reference algorithm (pEVSL) that checks the results of the HIP
 * implementation in pevsl_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links,
 * calls or falls back to it.  Its long loops may run on OpenMP threads;
 * results are bitwise independent of the thread count (oracle.c).
 *
 * Each function cites the reference file:line whose behaviour it restates.
 * The reference itself cannot be built here (it needs MPI and BLAS/LAPACK),
 * so each function is pinned independently of this restatement by tests/:
 *   orc_lap3d / orc_exact_eig_lap3 : dense eigvalsh of the generated matrix
 *                                    equals the closed form
 *                                    (TESTS/COMMON/lapl.c:139-169).
 *   orc_csrmv                      : x_i = sin(i) recipe of
 *                                    TESTS0/Matvec/test1.c:70-94 vs scipy.
 *   orc_tridiag_eig                : scipy.linalg.eigh_tridiagonal (LAPACK).
 *   orc_dampcf / orc_chebxPltd     : the textbook Jackson and sigma factors;
 *                                    numpy's Chebyshev evaluation.
 *   orc_kpmdos                     : exact moments of a diagonal matrix by
 *                                    the spectral sum; the C1 count 118
 *                                    within the stochastic error.
 *   orc_intchx / orc_spslicer_kpm  : numerical quadrature of the DOS; equal
 *                                    shares; monotone slices.
 *   orc_find_pol                   : the README degrees (README.md:111-115)
 *                                    and the invariants p(gam)=1,
 *                                    p(a)=p(b)=bar.
 *   orc_cheb_av                    : p(A)v on a diagonal matrix == scalar
 *                                    evaluation of p at each diagonal entry.
 *   orc_cgs_dgks                   : orthogonality; the exact projection.
 *   orc_cheblan_nr                 : the README counts per slice, the closed
 *                                    form to 1e-12, eigenvector residuals.
 *   orc_landos / orc_spslicer_lan  : full-length Lanczos equals the spectral
 *                                    measure of the start vector; Simpson
 *                                    exact on cubics.
 *   orc_sym_eig / orc_lan_trbounds : numpy.linalg.eigvalsh; closed-form
 *                                    extremes of Laplacians enclosed.
 */
#ifndef PEVSL_ORACLE_H
#define PEVSL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* 0-based CSR with int32 indices (INC/pevsl_struct.h:23-26) */
typedef struct {
  int nrows, ncols;
  int *ia, *ja;
  double *a;
} orc_csr;

/* filter polynomial (INC/pevsl_struct.h:143-173) */
typedef struct {
  int max_deg, min_deg, damping;
  double thresh_ext, thresh_int, intvtol;
  int type;
  double *mu; /* owned, max_deg+1 */
  double cc, dd, gam, bar;
  int deg;
} orc_pol;

/* diagnostics of one filtered-Lanczos run (not in the reference ABI) */
typedef struct {
  int kdim;        /* dimension of T at the final eigen-solve            */
  int its;         /* Lanczos steps taken (k at exit + 1, or maxit)       */
  int nbreak;      /* lucky breakdowns (fresh filtered random vectors)    */
  int nconv;       /* Ritz values of p(A) >= bar at the last test          */
  int nmatvec;     /* products with A                                    */
  double tr_last;  /* restricted trace at the last test                  */
} orc_lan_info;

/* matrices / vectors */
int  orc_lap3d(int nx, int ny, int nz, orc_csr *A);
void orc_csr_free(orc_csr *A);
void orc_csrmv(const orc_csr *A, const double *x, double *y);
int  orc_exact_eig_lap3(int nx, int ny, int nz, double a, double b,
                        double *out, int cap);
void orc_srand(unsigned seed);
void orc_rand_vec(int n, double *x);
void orc_linspace(double a, double b, int num, double *arr);
void orc_sort_double(int n, double *v, int *ind);

/* dense helpers */
int  orc_tridiag_eig(int n, const double *d, const double *e, double *w,
                     double *z);
void orc_cgs_dgks(int n, int k, int i_max, const double *V, int ld, double *v,
                  double *nrmv, double *work);
/* dense symmetric eigenpairs (A, Z column-major n x n; w ascending) */
int  orc_sym_eig(int n, const double *A, double *w, double *Z);
/* thick-restart Lanczos spectrum bounds (SRC/lantrbnd.c:31-408) */
int  orc_lan_trbounds(const orc_csr *A, int lanm, int maxit, double tol, const double *vinit,
                      int bndtype, double *lammin, double *lammax, int *its);

/* KPM DOS + slicing */
int  orc_dampcf(int m, int damping, double *jac);
int  orc_kpmdos(const orc_csr *A, int Mdeg, int damping, int nvec,
                const double *intv, double *mu, double *ecnt);
void orc_intchx(int Mdeg, const double *mu, int npts, const double *xi,
                double *yi);
int  orc_spslicer_kpm(double *sli, const double *mu, int Mdeg,
                      const double *intv, int n_int, int npts);

/* filter design / application */
void orc_set_pol_def(orc_pol *pol);
int  orc_find_pol(const double *intv, orc_pol *pol);
void orc_free_pol(orc_pol *pol);
int  orc_chebxPltd(int m, const double *mu, int npts, const double *xi,
                   double *yi);
int  orc_cheb_av(const orc_csr *A, const orc_pol *pol, const double *v,
                 double *y, double *w3);

/* filtered Lanczos, no restart */
int  orc_cheblan_nr(const orc_csr *A, const double *intv, int maxit,
                    double tol, const double *vinit, const orc_pol *pol,
                    int *nev, double *lam, double *res, double *Y,
                    orc_lan_info *info);

/* Lanczos DOS */
void orc_simpson(const double *xi, double *yi, int npts);
int  orc_landos(const orc_csr *A, int nvec, int msteps, int npts, double *xdos,
                double *ydos, double *neig, const double *intv,
                double sigma_mult);
void orc_spslicer_lan(const double *xi, double *yi, int n_int, int npts,
                      double *sli);

#ifdef __cplusplus
}
#endif
#endif
