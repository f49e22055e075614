/*
 * oracle.c -- CPU oracle (FP64) for the pEVSL spectrum-slicing
 * path.  TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py, to check results; never by the product library.
 *
 * Every routine restates the reference algorithm at the cited file:line.
 * BLAS level-1/2 calls of the reference are written out as plain loops in the
 * order of the netlib reference BLAS (sequential sums, column-by-column
 * GEMV), so results agree with the reference to rounding of the summation
 * order.  The tridiagonal eigensolver (LAPACK DSTEV in the reference,
 * SRC/miscla.c:29-66) is an implicit-shift QL iteration (Bowdler, Martin,
 * Reinsch & Wilkinson, Numer. Math. 11 (1968) 293-306), pinned against
 * scipy.linalg.eigh_tridiagonal in tests/.
 *
 * Built with OpenMP, the O(n) and O(n k) loops (CSR product, axpy, scal, the
 * CGS dot products and update) split over threads for n >= ORC_PAR_MIN in a
 * way that leaves every element's sequence of operations as in the
 * sequential loops: rows or elements are divided among threads, each dot
 * product stays one sequential sum, and the CGS update keeps its column
 * order per element.  Results are therefore bitwise the same for any
 * thread count.  Other dot products (norms, Lanczos alpha) stay sequential.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_ORTH_TOL 1e-14 /* INC/pevsl_def.h:18 orthTol */
#define ORC_NGS_MAX 2      /* INC/pevsl_def.h:112 NGS_MAX */
#define ORC_PAR_MIN 65536  /* vectors at least this long use threads */

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
    abort();
  }
  return p;
}

static double dot(int n, const double *x, const double *y) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}
static void axpy(int n, double a, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (n >= ORC_PAR_MIN)
  for (int i = 0; i < n; i++) y[i] += a * x[i];
}
static void scal(int n, double a, double *x) {
#pragma omp parallel for schedule(static) if (n >= ORC_PAR_MIN)
  for (int i = 0; i < n; i++) x[i] *= a;
}

/* ------------------------------------------------------------------------
 * Matrices and vectors
 * --------------------------------------------------------------------- */

/* 7-point (5-point if nz == 1) Dirichlet Laplacian in natural order
 * idx = k*nx*ny + i*nx + j, j fastest; entries per row in the order
 * -z, -y, -x, diag, +x, +y, +z.  TESTS/COMMON/lapl.c:4-70 (+ the COO->CSR
 * conversion SRC/spmat.c:56-86, which keeps the per-row insertion order). */
int orc_lap3d(int nx, int ny, int nz, orc_csr *A) {
  int n = nx * ny * nz;
  int per = nz > 1 ? 7 : 5;
  A->nrows = A->ncols = n;
  A->ia = (int *)xmalloc(sizeof(int) * (size_t)(n + 1));
  A->ja = (int *)xmalloc(sizeof(int) * (size_t)per * n);
  A->a = (double *)xmalloc(sizeof(double) * (size_t)per * n);
  const double off = -1.0, dg = nz > 1 ? 6.0 : 4.0;
  int nnz = 0;
  A->ia[0] = 0;
  for (int r = 0; r < n; r++) {
    int k = r / (nx * ny);
    int i = (r - k * nx * ny) / nx;
    int j = r - k * nx * ny - i * nx;
#define PUT(c, v) do { A->ja[nnz] = (c); A->a[nnz] = (v); nnz++; } while (0)
    if (k > 0) PUT(r - nx * ny, off);
    if (i > 0) PUT(r - nx, off);
    if (j > 0) PUT(r - 1, off);
    PUT(r, dg);
    if (j < nx - 1) PUT(r + 1, off);
    if (i < ny - 1) PUT(r + nx, off);
    if (k < nz - 1) PUT(r + nx * ny, off);
#undef PUT
    A->ia[r + 1] = nnz;
  }
  return 0;
}

void orc_csr_free(orc_csr *A) {
  free(A->ia);
  free(A->ja);
  free(A->a);
  A->ia = A->ja = NULL;
  A->a = NULL;
}

/* y = A x, one sequential sum per row: SRC/spmat.c:184-206 (dcsrmv 'N'). */
void orc_csrmv(const orc_csr *A, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (A->nrows >= ORC_PAR_MIN)
  for (int i = 0; i < A->nrows; i++) {
    double r = 0.0;
    for (int j = A->ia[i]; j < A->ia[i + 1]; j++) r += A->a[j] * x[A->ja[j]];
    y[i] = r;
  }
}

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* Closed-form eigenvalues of the Laplacian inside [a-eps, b+eps], sorted.
 * TESTS/COMMON/lapl.c:139-169.  Returns the count (writes at most cap). */
int orc_exact_eig_lap3(int nx, int ny, int nz, double a, double b,
                       double *out, int cap) {
  double thx = 0.5 * M_PI / (nx + 1.0);
  double thy = 0.5 * M_PI / (ny + 1.0);
  double thz = 0.5 * M_PI / (nz + 1.0);
  int cnt = 0;
  for (int i = 1; i <= nx; i++) {
    double tx = sin(i * thx);
    for (int j = 1; j <= ny; j++) {
      double ty = sin(j * thy);
      for (int k = 1; k <= nz; k++) {
        double tz = nz == 1 ? 0.0 : sin(k * thz);
        double ev = 4 * (tx * tx + ty * ty + tz * tz);
        if (ev >= a - DBL_EPSILON && ev <= b + DBL_EPSILON) {
          if (out && cnt < cap) out[cnt] = ev;
          cnt++;
        }
      }
    }
  }
  if (out) qsort(out, (size_t)(cnt < cap ? cnt : cap), sizeof(double), cmp_double);
  return cnt;
}

/* The reference seeds libc rand() with the world rank (SRC/pevsl.c:36) and
 * draws uniform [-1,1] values as (rand() - RAND_MAX/2)/(RAND_MAX/2)
 * (SRC/parvec.c:65-72).  The same libc generator is used here. */
void orc_srand(unsigned seed) { srand(seed); }

void orc_rand_vec(int n, double *x) {
  double t = ((double)RAND_MAX) / 2.0;
  for (int i = 0; i < n; i++) x[i] = (rand() - t) / t;
}

/* SRC/vector.c:15-25 */
void orc_linspace(double a, double b, int num, double *arr) {
  double h = num == 1 ? 0 : (b - a) / (num - 1);
  arr[0] = a;
  arr[num - 1] = b;
  for (int i = 1; i < num - 1; i++) arr[i] = a + i * h;
}

typedef struct { int i; double d; } dbl_idx;
static int cmp_dbl_idx(const void *a, const void *b) {
  double x = ((const dbl_idx *)a)->d, y = ((const dbl_idx *)b)->d;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* SRC/vector.c:85-105 */
void orc_sort_double(int n, double *v, int *ind) {
  if (!ind) {
    qsort(v, (size_t)n, sizeof(double), cmp_double);
    return;
  }
  dbl_idx *t = (dbl_idx *)xmalloc(sizeof(dbl_idx) * (size_t)n);
  for (int i = 0; i < n; i++) { t[i].d = v[i]; t[i].i = i; }
  qsort(t, (size_t)n, sizeof(dbl_idx), cmp_dbl_idx);
  for (int i = 0; i < n; i++) { v[i] = t[i].d; ind[i] = t[i].i; }
  free(t);
}

/* ------------------------------------------------------------------------
 * Dense helpers
 * --------------------------------------------------------------------- */

/* All eigenvalues (ascending, in w) and optionally eigenvectors (z: n x n,
 * column-major, column i belongs to w[i]) of the symmetric tridiagonal
 * matrix with diagonal d[0..n-1] and sub-diagonal e[0..n-2].
 * Replaces DSTEV (SRC/miscla.c:29-66) by implicit-shift QL. */
int orc_tridiag_eig(int n, const double *d, const double *e, double *w,
                    double *z) {
  if (n <= 0) return 0;
  double *f = (double *)xmalloc(sizeof(double) * (size_t)n);
  memcpy(w, d, sizeof(double) * (size_t)n);
  for (int i = 0; i < n - 1; i++) f[i] = e[i];
  f[n - 1] = 0.0;
  if (z) {
    memset(z, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; i++) z[(size_t)i * n + i] = 1.0;
  }
  int info = 0;
  for (int l = 0; l < n; l++) {
    int iter = 0, m;
    for (;;) {
      for (m = l; m < n - 1; m++) {
        double dd = fabs(w[m]) + fabs(w[m + 1]);
        if (fabs(f[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) { info = l + 1; break; }
      /* Wilkinson-type shift from the leading 2x2 block */
      double g = (w[l + 1] - w[l]) / (2.0 * f[l]);
      double r = hypot(g, 1.0);
      g = w[m] - w[l] + f[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      int early = 0;
      for (i = m - 1; i >= l; i--) {
        double ff = s * f[i], b = c * f[i];
        r = hypot(ff, g);
        f[i + 1] = r;
        if (r == 0.0) {
          w[i + 1] -= p;
          f[m] = 0.0;
          early = 1;
          break;
        }
        s = ff / r;
        c = g / r;
        g = w[i + 1] - p;
        r = (w[i] - g) * s + 2.0 * c * b;
        p = s * r;
        w[i + 1] = g + p;
        g = c * r - b;
        if (z) {
          double *zi = z + (size_t)i * n, *zi1 = z + (size_t)(i + 1) * n;
          for (int k = 0; k < n; k++) {
            double t = zi1[k];
            zi1[k] = s * zi[k] + c * t;
            zi[k] = c * zi[k] - s * t;
          }
        }
      }
      if (early) continue;
      w[l] -= p;
      f[l] = g;
      f[m] = 0.0;
    }
  }
  /* selection sort keeps eigenvector columns attached */
  for (int i = 0; i < n - 1; i++) {
    int kmin = i;
    for (int j = i + 1; j < n; j++) if (w[j] < w[kmin]) kmin = j;
    if (kmin != i) {
      double t = w[i]; w[i] = w[kmin]; w[kmin] = t;
      if (z) {
        double *a = z + (size_t)i * n, *b = z + (size_t)kmin * n;
        for (int k = 0; k < n; k++) { t = a[k]; a[k] = b[k]; b[k] = t; }
      }
    }
  }
  free(f);
  return info;
}

/* Classical Gram-Schmidt with the DGKS refinement test, at most i_max passes,
 * eta = 1/sqrt(2): SRC/miscla.c:211-240 with the GEMV pair of
 * SRC/parvecs.c:93-136.  V is n x k column-major with leading dimension ld;
 * work holds k doubles. */
void orc_cgs_dgks(int n, int k, int i_max, const double *V, int ld, double *v,
                  double *nrmv, double *work) {
  const double eta = 1.0 / sqrt(2.0);
  double old_nrm = sqrt(dot(n, v, v)), new_nrm = old_nrm;
  for (int it = 0; it < i_max; it++) {
#pragma omp parallel for schedule(static) if (n >= ORC_PAR_MIN && k > 1)
    for (int j = 0; j < k; j++) work[j] = dot(n, V + (size_t)j * ld, v);
    /* v -= V work, column by column for every element (netlib DGEMV 'N') */
#pragma omp parallel for schedule(static) if (n >= ORC_PAR_MIN)
    for (int i0 = 0; i0 < n; i0 += 1024) {
      const int i1 = i0 + 1024 < n ? i0 + 1024 : n;
      for (int j = 0; j < k; j++) {
        const double a = -work[j], *x = V + (size_t)j * ld;
        for (int i = i0; i < i1; i++) v[i] += a * x[i];
      }
    }
    new_nrm = sqrt(dot(n, v, v));
    if (new_nrm > eta * old_nrm) break;
    old_nrm = new_nrm;
  }
  if (nrmv) *nrmv = new_nrm;
}

/* ------------------------------------------------------------------------
 * KPM density of states and slicing
 * --------------------------------------------------------------------- */

/* Damping factors, jac[0] = 1/2: SRC/chebpol.c:34-61. */
int orc_dampcf(int m, int damping, double *jac) {
  double thetJ = 0, thetL = 0, a1 = 0, a2 = 0, dm = (double)m;
  if (damping == 1) {
    thetJ = M_PI / (dm + 2);
    a1 = 1 / (dm + 2);
    a2 = sin(thetJ);
  }
  if (damping == 2) thetL = M_PI / (dm + 1);
  jac[0] = 0.5;
  for (int k = 1; k <= m; k++) {
    if (damping == 0) {
      jac[k] = 1.0;
    } else if (damping == 1) {
      int k1 = k + 1;
      jac[k] = a1 * sin(k1 * thetJ) / a2 + (1 - k1 * a1) * cos(k * thetJ);
    } else {
      jac[k] = sin(k * thetL) / (k * thetL);
    }
  }
  return 0;
}

static int check_intv(const double *intv) {
  double a = intv[0], b = intv[1], lmin = intv[2], lmax = intv[3];
  if (a >= b) return -1;
  if (a >= lmax || b <= lmin) return -2;
  return 0;
}

/* Chebyshev moments of the spectral density by stochastic trace estimation
 * with unit-norm random vectors and the eigenvalue-count estimate for
 * [intv[0], intv[1]]: SRC/kpmdos.c:36-197 (single group, standard problem). */
int orc_kpmdos(const orc_csr *A, int Mdeg, int damping, int nvec,
               const double *intv, double *mu, double *ecnt) {
  int n = A->nrows;
  if (check_intv(intv) < 0) return -1;
  double aa = intv[0] > intv[2] ? intv[0] : intv[2];
  double bb = intv[1] < intv[3] ? intv[1] : intv[3];
  double ctr = (intv[3] + intv[2]) / 2.0, wid = (intv[3] - intv[2]) / 2.0;
  double t = (aa - ctr) / wid;
  if (t < -1.0 + DBL_EPSILON) t = -1.0 + DBL_EPSILON;
  double beta1 = acos(t);
  t = (bb - ctr) / wid;
  if (t > 1.0 - DBL_EPSILON) t = 1.0 - DBL_EPSILON;
  double beta2 = acos(t);
  double *jac = (double *)xmalloc(sizeof(double) * (size_t)(Mdeg + 1));
  orc_dampcf(Mdeg, damping, jac);
  jac[0] = 1.0;
  memset(mu, 0, sizeof(double) * (size_t)(Mdeg + 1));
  double tcnt = 0.0;
  double *buf = (double *)xmalloc(sizeof(double) * (size_t)n * 4);
  double *v = buf, *vkm1 = buf + n, *vk = buf + 2 * (size_t)n,
         *vkp1 = buf + 3 * (size_t)n;
  for (int m = 0; m < nvec; m++) {
    orc_rand_vec(n, v);
    t = sqrt(dot(n, v, v));
    scal(n, 1.0 / t, v);
    memcpy(vk, v, sizeof(double) * (size_t)n);
    memset(vkm1, 0, sizeof(double) * (size_t)n);
    mu[0] += jac[0];
    tcnt -= jac[0] * (beta2 - beta1);
    for (int k = 0; k < Mdeg; k++) {
      orc_csrmv(A, vk, vkp1);
      double sc = (k == 0 ? 1.0 : 2.0) / wid;
      axpy(n, -ctr, vk, vkp1);
      scal(n, sc, vkp1);
      axpy(n, -1.0, vkm1, vkp1);
      double *tmp = vkm1; vkm1 = vk; vk = vkp1; vkp1 = tmp;
      int k1 = k + 1;
      t = dot(n, vk, v) * 2.0 * jac[k1];
      mu[k1] += t;
      tcnt -= t * (sin(k1 * beta2) - sin(k1 * beta1)) / k1;
    }
  }
  t = 1.0 / (((double)nvec) * M_PI);
  for (int k = 0; k <= Mdeg; k++) mu[k] *= t;
  tcnt *= t * ((double)n);
  *ecnt = tcnt;
  free(buf);
  free(jac);
  return 0;
}

/* yi[j] = integral from xi[0] to xi[j] of sum_k mu_k T_k(t)/sqrt(1-t^2),
 * in closed form: SRC/spslicer.c:12-38. */
void orc_intchx(int Mdeg, const double *mu, int npts, const double *xi,
                double *yi) {
  if (npts <= 0) return;
  double *th = (double *)xmalloc(sizeof(double) * (size_t)npts);
  for (int j = 0; j < npts; j++) th[j] = acos(xi[j]);
  double th0 = th[0];
  for (int j = 0; j < npts; j++) yi[j] = mu[0] * (th0 - th[j]);
  for (int k = 1; k <= Mdeg; k++) {
    double val0 = sin(k * th0) / k;
    for (int j = 0; j < npts; j++) yi[j] += mu[k] * (val0 - sin(k * th[j]) / k);
  }
  free(th);
}

/* Equal-count slicing of [intv[0], intv[1]] from the KPM moments:
 * SRC/spslicer.c:62-156. */
int orc_spslicer_kpm(double *sli, const double *mu, int Mdeg,
                     const double *intv, int n_int, int npts) {
  if (check_intv(intv) < 0) return -1;
  double aa = intv[0] > intv[2] ? intv[0] : intv[2];
  double bb = intv[1] < intv[3] ? intv[1] : intv[3];
  memset(sli, 0, sizeof(double) * (size_t)(n_int + 1));
  if (n_int == 1) {
    sli[0] = intv[0];
    sli[1] = intv[1];
    return 0;
  }
  double ctr = (intv[3] + intv[2]) / 2, wid = (intv[3] - intv[2]) / 2;
  double aL = (aa - ctr) / wid, bL = (bb - ctr) / wid;
  if (aL < -1.0) aL = -1.0;
  if (bL > 1.0) bL = 1.0;
  if (npts < 2 * n_int + 1) npts = 2 * n_int + 1;
  double *xi = (double *)xmalloc(sizeof(double) * (size_t)npts);
  double *yi = (double *)xmalloc(sizeof(double) * (size_t)npts);
  orc_linspace(aL, bL, npts, xi);
  orc_intchx(Mdeg, mu, npts, xi, yi);
  double target = yi[npts - 1] / (double)n_int;
  int ls = 0, ii = 0, err = 0;
  sli[0] = intv[0];
  while (++ls < n_int) {
    while (ii < npts && yi[ii] < target) ii++;
    if (ii == npts) break;
    /* SRC/spslicer.c:116-118 reads yi[ii-1] even at ii = 0 (a target <= 0,
     * when the DOS integral over the interval is not positive); the grid's
     * first point is then the cut */
    if (ii > 0 && (target - yi[ii - 1]) < (yi[ii] - target)) ii--;
    sli[ls] = ctr + wid * xi[ii];
    target = yi[ii] + (yi[npts - 1] - yi[ii]) / (n_int - ls);
  }
  sli[n_int] = intv[1];
  if (ls != n_int) err = 1;
  for (int i = 1; i <= n_int; i++) {
    if (sli[i] < sli[i - 1]) { err += 2; break; }
  }
  free(xi);
  free(yi);
  return err;
}

/* ------------------------------------------------------------------------
 * Polynomial filter design and application
 * --------------------------------------------------------------------- */

/* SRC/chebpol.c:10-19 */
void orc_set_pol_def(orc_pol *pol) {
  pol->max_deg = 5000;
  pol->min_deg = 2;
  pol->damping = 2;
  pol->thresh_ext = 0.20;
  pol->thresh_int = 0.8;
  pol->mu = NULL;
  pol->deg = 0;
  pol->intvtol = 1e-9;
  pol->type = 0;
  pol->cc = pol->dd = pol->gam = pol->bar = 0.0;
}

void orc_free_pol(orc_pol *pol) {
  free(pol->mu);
  pol->mu = NULL;
}

/* SRC/chebpol.c:73-79 */
static double dif_eval(const double *v, int m, double thc, const double *jac) {
  double f = 0.0;
  for (int j = 0; j <= m; j++) f += v[j] * cos(j * thc) * jac[j];
  return f;
}

/* yi = p(xi) for p = sum mu_k T_k, three-term recurrence: SRC/chebpol.c:100-131 */
int orc_chebxPltd(int m, const double *mu, int npts, const double *xi,
                  double *yi) {
  double *b = (double *)xmalloc(sizeof(double) * (size_t)npts * 3);
  double *vkm1 = b, *vk = b + npts, *vkp1 = b + 2 * (size_t)npts;
  for (int j = 0; j < npts; j++) { vkm1[j] = 0.0; vk[j] = 1.0; yi[j] = mu[0]; }
  for (int k = 0; k < m; k++) {
    double sc = k == 0 ? 1.0 : 2.0;
    for (int j = 0; j < npts; j++) vkp1[j] = sc * xi[j] * vk[j] - vkm1[j];
    double *tmp = vkm1; vkm1 = vk; vk = vkp1; vkp1 = tmp;
    for (int j = 0; j < npts; j++) yi[j] += mu[k + 1] * vk[j];
  }
  free(b);
  return 0;
}

/* End-interval filter: a scaled Chebyshev polynomial written in the
 * T_k basis, degree grown until its value at the inner edge drops below
 * thresh_ext: SRC/chebpol.c:166-255. */
static void chext(orc_pol *pol, double aIn, double bIn) {
  int max_deg = pol->max_deg;
  double thresh = pol->thresh_ext;
  double *mu = pol->mu;
  double del = 0.1 * sqrt((bIn - aIn) * 0.5);
  double eps = 1e-13;
  double a, b, x, gam;
  if (pol->deg > 0) {
    thresh = -1.0;
    max_deg = pol->deg;
  }
  if (bIn >= (1.0 - eps)) {
    x = aIn; b = x - del; a = -1.0; gam = 1.0;
  } else {
    x = bIn; a = x + del; b = 1.0; gam = -1.0;
  }
  double e = (b - a) * 0.5, c = (b + a) * 0.5;
  double sigma1 = e / (gam - c), sigma = sigma1;
  int m1 = max_deg + 1;
  double *t0 = (double *)calloc((size_t)3 * m1, sizeof(double));
  double *t1 = t0 + m1, *tnew = t1 + m1;
  t0[0] = 1.0;
  t1[0] = -c * (sigma1 / e);
  t1[1] = sigma1 / e;
  double g0 = 1.0, g1 = (x - c) * sigma1 / e;
  int mbest = 0;
  if (g1 < thresh) {
    mbest = 1;
  } else {
    for (int k = 2; k <= max_deg; k++) {
      double sigma_new = 1.0 / (2.0 / sigma1 - sigma);
      double s1 = sigma_new / e, s2 = 2 * s1 * c, s3 = sigma * sigma_new;
      for (int j = 2; j <= k; j++) tnew[j - 1] = s1 * (t1[j] + t1[j - 2]);
      tnew[1] = tnew[1] + s1 * t1[0];
      tnew[0] = s1 * t1[1];
      tnew[k] = s1 * t1[k - 1];
      for (int j = 0; j <= k; j++) tnew[j] = tnew[j] - s2 * t1[j] - s3 * t0[j];
      for (int j = 0; j <= k; j++) { t0[j] = t1[j]; t1[j] = tnew[j]; }
      double gnew = 2 * (x - c) * s1 * g1 - s3 * g0;
      g0 = g1;
      g1 = gnew;
      mbest = k;
      if (g1 < thresh) break;
      sigma = sigma_new;
    }
  }
  memcpy(mu, t1, sizeof(double) * (size_t)(mbest + 1));
  pol->deg = mbest;
  pol->bar = g1;
  pol->gam = gam;
  free(t0);
}

/* Newton/bisection balancing of the delta-function centre so that the
 * damped expansion has p(a) == p(b): SRC/chebpol.c:284-332. */
static int rootchb(int m, const double *v, const double *jac, double tha,
                   double thb, double *mu, double *thcOut) {
  const int MaxIterBalan = 30;
  double tolBal = fabs(tha - thb) * 1.e-13;
  double thc = 0.5 * (tha + thb);
  double fb = dif_eval(v, m, thb, jac), fa = dif_eval(v, m, tha, jac);
  if ((fa > 0) || (fb < 0)) return 1;
  for (int it = 0; it <= MaxIterBalan; it++) {
    double fval = dif_eval(v, m, thc, jac);
    double d = 0.0;
    for (int j = 1; j <= m; j++) d += jac[j] * j * sin(j * thc) * v[j];
    double thN = thc + fval / d;
    if ((fabs(fval) < tolBal) || fabs(thc - thN) < DBL_EPSILON * fabs(thc)) break;
    if (fval > 0) {
      if ((thN < thb) || (thN > tha)) thN = 0.5 * (thc + tha);
      thb = thc;
      thc = thN;
    } else {
      if ((thN < thb) || (thN > tha)) thN = 0.5 * (thc + thb);
      tha = thc;
      thc = thN;
    }
  }
  for (int j = 0; j <= m; j++) mu[j] = cos(j * thc) * jac[j];
  *thcOut = thc;
  return 0;
}

/* Filter design for one slice: SRC/chebpol.c:359-484. */
int orc_find_pol(const double *intv, orc_pol *pol) {
  int max_deg = pol->max_deg, damping = pol->damping, min_deg;
  if (check_intv(intv) < 0) return -1;
  double *mu = (double *)xmalloc(sizeof(double) * (size_t)(max_deg + 1));
  /* the reference leaves mu uninitialised beyond the accepted degree */
  memset(mu, 0, sizeof(double) * (size_t)(max_deg + 1));
  pol->mu = mu;
  double *v = (double *)xmalloc(sizeof(double) * (size_t)(max_deg + 1));
  double *jac = (double *)xmalloc(sizeof(double) * (size_t)(max_deg + 1));
  double IntTol = pol->intvtol;
  double aa = intv[0] > intv[2] ? intv[0] : intv[2];
  double bb = intv[1] < intv[3] ? intv[1] : intv[3];
  double lmin = intv[2], lmax = intv[3];
  double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  pol->cc = cc;
  pol->dd = dd;
  aa = (aa - cc) / dd;
  bb = (bb - cc) / dd;
  if (aa < -1.0) aa = -1.0;
  if (bb > 1.0) bb = 1.0;
  double thb = acos(bb), tha = acos(aa), thc = 0.0;
  double itv[2] = {0, 0}, vals[2] = {0, 0}, t = 0.0, gam = 0.0;
  int nitv;
  if (aa - IntTol <= -1.0) {
    thc = tha;
    nitv = 1;
    aa = -1.0;
    gam = -1.0;
    pol->type = 1;
  } else if (bb + IntTol >= 1.0) {
    thc = thb;
    nitv = 1;
    bb = 1;
    gam = 1;
    pol->type = 2;
  } else {
    itv[0] = aa;
    itv[1] = bb;
    nitv = 2;
    pol->type = 0;
  }
  (void)thc;
  if (nitv == 1) {
    chext(pol, aa, bb);
  } else {
    /* SRC/chebpol.c:434 reads `2 + (int) 0.5/(bb-aa)`: the cast binds to
     * the literal 0.5, so the starting degree is always 2. */
    min_deg = 2 + (int)(((int)0.5) / (bb - aa));
    double thresh = pol->thresh_int;
    if (pol->deg > 0) {
      thresh = -1;
      max_deg = pol->deg;
    }
    for (int j = 0; j < min_deg; j++) v[j] = cos(j * thb) - cos(j * tha);
    int m;
    for (m = min_deg; m < max_deg; m++) {
      orc_dampcf(m, damping, jac);
      v[m] = cos(m * thb) - cos(m * tha);
      if (rootchb(m, v, jac, tha, thb, mu, &thc)) continue;
      gam = cos(thc);
      orc_chebxPltd(m, mu, 1, &gam, &t);
      orc_chebxPltd(m, mu, nitv, itv, vals);
      if (vals[0] <= t * thresh && vals[1] <= t * thresh) {
        m++;
        break;
      }
    }
    int mbest = m - 1;
    for (int j = 0; j <= m && j <= pol->max_deg; j++) mu[j] /= t;
    pol->bar = (vals[0] < vals[1] ? vals[0] : vals[1]) / t;
    pol->gam = gam;
    pol->deg = mbest;
  }
  free(v);
  free(jac);
  return 0;
}

/* y = p((A - cc)/dd) v by the Chebyshev three-term recurrence:
 * SRC/chebpol.c:509-574.  w3 holds 3n doubles. */
int orc_cheb_av(const orc_csr *A, const orc_pol *pol, const double *v,
                double *y, double *w3) {
  int n = A->nrows, m = pol->deg;
  const double *mu = pol->mu;
  double cc = pol->cc, dd = pol->dd, t1 = 1.0 / dd, t2 = 2.0 / dd;
  double *vk = w3, *vkp1 = w3 + n, *vkm1 = w3 + 2 * (size_t)n;
  memcpy(vk, v, sizeof(double) * (size_t)n);
  memcpy(y, vk, sizeof(double) * (size_t)n);
  scal(n, mu[0], y);
  for (int k = 1; k <= m; k++) {
    double t = k == 1 ? t1 : t2;
    orc_csrmv(A, vk, vkp1);
    axpy(n, -cc, vk, vkp1);
    scal(n, t, vkp1);
    if (k > 1) axpy(n, -1.0, vkm1, vkp1);
    axpy(n, mu[k], vkp1, y);
    double *tmp = vkm1; vkm1 = vk; vk = vkp1; vkp1 = tmp;
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * Filtered Lanczos without restart: SRC/cheblanNr.c:54-448
 * (standard problem, FILTER_VINIT = 1, Ntest = 30, cycle = 20).
 * lam/res need maxit entries; Y (optional) n x maxit column-major receives
 * the accepted Ritz vectors in the order of lam.
 * --------------------------------------------------------------------- */
int orc_cheblan_nr(const orc_csr *A, const double *intv, int maxit,
                   double tol, const double *vinit, const orc_pol *pol,
                   int *nevOut, double *lam, double *res, double *Y,
                   orc_lan_info *info) {
  int n = A->nrows;
  const int Ntest = 30, cycle = 20;
  if (maxit > n) maxit = n;
  double bar = pol->bar, gamB = pol->gam;
  orc_lan_info inf;
  memset(&inf, 0, sizeof(inf));
  *nevOut = 0;
  if (check_intv(intv) < 0) {
    if (info) *info = inf;
    return 0;
  }
  double aa = intv[0], bb = intv[1];
  if (gamB > 1.0 || gamB < -1.0) return -1;
  size_t ldv = (size_t)n;
  double *V = (double *)xmalloc(sizeof(double) * ldv * (size_t)(maxit + 1));
  double *dT = (double *)xmalloc(sizeof(double) * (size_t)maxit);
  double *eT = (double *)xmalloc(sizeof(double) * (size_t)maxit);
  double *EvalT = (double *)xmalloc(sizeof(double) * (size_t)maxit);
  double *wk = (double *)xmalloc(sizeof(double) * 3 * (size_t)n);
  double *vrand = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *warr = (double *)xmalloc(sizeof(double) * 3 * (size_t)maxit);
  int mv_per_filter = pol->deg;

  /* V(:,0) = p(A) vinit / ||.|| */
  orc_cheb_av(A, pol, vinit, V, wk);
  inf.nmatvec += mv_per_filter;
  double t = sqrt(dot(n, V, V));
  scal(n, 1.0 / t, V);
  double wn = 0.0, tr0 = 0.0, tr1 = 0.0, alpha, beta = 0.0;
  int nwn = 0, kdim = 0, count = 0, nconv = 0, k;
  for (k = 0; k < maxit; k++) {
    double *zold = k > 0 ? V + ldv * (size_t)(k - 1) : NULL;
    double *z = V + ldv * (size_t)k;
    double *znew = V + ldv * (size_t)(k + 1);
    orc_cheb_av(A, pol, z, znew, wk);
    inf.nmatvec += mv_per_filter;
    if (k > 0) axpy(n, -beta, zold, znew);
    alpha = dot(n, z, znew);
    dT[k] = alpha;
    wn += fabs(alpha);
    axpy(n, -alpha, z, znew);
    orc_cgs_dgks(n, k + 1, ORC_NGS_MAX, V, n, znew, &beta, warr);
    wn += 2.0 * beta;
    nwn += 3;
    if (beta * nwn < ORC_ORTH_TOL * wn) {
      inf.nbreak++;
      orc_rand_vec(n, vrand);
      orc_cheb_av(A, pol, vrand, znew, wk);
      inf.nmatvec += mv_per_filter;
      orc_cgs_dgks(n, k + 1, ORC_NGS_MAX, V, n, znew, &beta, warr);
      scal(n, 1.0 / beta, znew);
      beta = 0.0;
    } else {
      scal(n, 1.0 / beta, znew);
    }
    eT[k] = beta;
    if ((k < Ntest || (k - Ntest) % cycle != 0) && k != maxit - 1) continue;
    kdim = k + 1;
    orc_tridiag_eig(kdim, dT, eT, EvalT, NULL);
    count = kdim;
    tr1 = 0;
    nconv = 0;
    for (int i = 0; i < count; i++) {
      if (EvalT[i] + 5 * DBL_EPSILON >= bar) {
        tr1 += EvalT[i];
        nconv++;
      }
    }
    if ((fabs(tr1 - tr0) < 2e-12) || (fabs(tr1) + fabs(tr0) < 1e-10)) break;
    tr0 = tr1;
  }
  inf.its = k < maxit ? k + 1 : maxit;
  inf.kdim = kdim;
  inf.nconv = nconv;
  inf.tr_last = tr1;

  double *EvecT = (double *)xmalloc(sizeof(double) * (size_t)kdim * kdim);
  orc_tridiag_eig(kdim, dT, eT, EvalT, EvecT);
  double *u = (double *)xmalloc(sizeof(double) * (size_t)n);
  int nev = 0, rc = 0;
  for (int i = 0; i < count; i++) {
    if (EvalT[i] < bar) continue;
    const double *y = EvecT + (size_t)i * kdim;
    /* u = V(:, 0:kdim-1) y, column-by-column as in DGEMV 'N' */
    memset(u, 0, sizeof(double) * (size_t)n);
    for (int j = 0; j < kdim; j++) axpy(n, y[j], V + ldv * (size_t)j, u);
    t = sqrt(dot(n, u, u));
    if (t == 0.0) { rc = 2; break; }
    scal(n, 1.0 / t, u);
    orc_csrmv(A, u, wk);
    inf.nmatvec++;
    t = dot(n, wk, u);
    if (t < aa - DBL_EPSILON || t > bb + DBL_EPSILON) continue;
    axpy(n, -t, u, wk);
    double res0 = sqrt(dot(n, wk, wk));
    if (res0 < tol) {
      lam[nev] = t;
      res[nev] = res0;
      if (Y) memcpy(Y + (size_t)nev * n, u, sizeof(double) * (size_t)n);
      nev++;
    }
  }
  *nevOut = nev;
  if (info) *info = inf;
  free(u);
  free(EvecT);
  free(V);
  free(dT);
  free(eT);
  free(EvalT);
  free(wk);
  free(vrand);
  free(warr);
  return rc;
}

/* ------------------------------------------------------------------------
 * Lanczos DOS: SRC/landos.c:40-314, SRC/simpson.c:20-45,
 * SRC/landos.c:331-355 (standard problem, single group).
 * --------------------------------------------------------------------- */

/* In-place cumulative integral by composite Simpson (trapezoid for the odd
 * points): SRC/simpson.c:20-45. */
void orc_simpson(const double *xi, double *yi, int npts) {
  double tc = 0.0, ti, tm, ysav;
  int i = 1;
  ysav = yi[0];
  yi[0] = 0;
  while (i < npts - 1) {
    ti = (xi[i] - xi[i - 1]) * (ysav + yi[i]) / 2.0;
    tm = tc + ti;
    ti = (xi[i + 1] - xi[i - 1]) * (ysav + 4 * yi[i] + yi[i + 1]) / 6.0;
    tc += ti;
    yi[i] = tm;
    ysav = yi[i + 1];
    yi[i + 1] = tc;
    i += 2;
  }
  if (i == npts - 1) {
    ti = (xi[i] - xi[i - 1]) * (yi[i] + ysav) / 2.0;
    yi[i] = tc + ti;
  }
}

int orc_landos(const orc_csr *A, int nvec, int msteps, int npts, double *xdos,
               double *ydos, double *neig, const double *intv,
               double sigma_mult) {
  int n = A->nrows;
  int maxit = msteps < n ? msteps : n;
  size_t ldv = (size_t)n;
  double *V = (double *)xmalloc(sizeof(double) * ldv * (size_t)(maxit + 1));
  double *vinit = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *dT = (double *)xmalloc(sizeof(double) * (size_t)maxit);
  double *eT = (double *)xmalloc(sizeof(double) * (size_t)maxit);
  double *EvalT = (double *)xmalloc(sizeof(double) * (size_t)maxit);
  double *EvecT = (double *)xmalloc(sizeof(double) * (size_t)maxit * maxit);
  double *gamma2 = (double *)xmalloc(sizeof(double) * (size_t)maxit);
  double *warr = (double *)xmalloc(sizeof(double) * 3 * (size_t)maxit);
  double *y = (double *)calloc((size_t)npts, sizeof(double));
  int *ind = (int *)xmalloc(sizeof(int) * (size_t)npts);
  const double lm = intv[2], lM = intv[3];
  const double aa = intv[0] > intv[2] ? intv[0] : intv[2];
  const double bb = intv[1] < intv[3] ? intv[1] : intv[3];
  const double kappa = 1.25;
  const int M = msteps < 30 ? msteps : 30;
  const double H = (lM - lm) / (M - 1);
  const double sigma = H / sqrt(8 * log(kappa)) * sigma_mult;
  const double sigma2 = 2 * sigma * sigma;
  const double width = sigma * sqrt(-2.0 * log(1e-08));
  orc_linspace(aa, bb, npts, xdos);
  for (int m = 0; m < nvec; m++) {
    orc_rand_vec(n, vinit);
    double t = 1.0 / sqrt(dot(n, vinit, vinit));
    memcpy(V, vinit, sizeof(double) * (size_t)n);
    scal(n, t, V);
    double wn = 0.0, alpha, beta = 0.0;
    int nwn = 0;
    for (int k = 0; k < maxit; k++) {
      double *zold = k > 0 ? V + ldv * (size_t)(k - 1) : NULL;
      double *z = V + ldv * (size_t)k, *znew = V + ldv * (size_t)(k + 1);
      orc_csrmv(A, z, znew);
      if (k > 0) axpy(n, -beta, zold, znew);
      alpha = dot(n, z, znew);
      dT[k] = alpha;
      wn += fabs(alpha);
      axpy(n, -alpha, z, znew);
      orc_cgs_dgks(n, k + 1, ORC_NGS_MAX, V, n, znew, &beta, warr);
      wn += 2.0 * beta;
      nwn += 3;
      if (beta * nwn < ORC_ORTH_TOL * wn) {
        orc_rand_vec(n, znew);
        orc_cgs_dgks(n, k + 1, ORC_NGS_MAX, V, n, znew, &beta, warr);
        scal(n, 1.0 / beta, znew);
        beta = 0.0;
      } else {
        scal(n, 1.0 / beta, znew);
      }
      eT[k] = beta;
    }
    orc_tridiag_eig(maxit, dT, eT, EvalT, EvecT);
    for (int i = 0; i < maxit; i++)
      gamma2[i] = EvecT[(size_t)i * maxit] * EvecT[(size_t)i * maxit];
    for (int i = 0; i < maxit; i++) {
      const double th = EvalT[i];
      int np = 0;
      for (int j = 0; j < npts; j++)
        if (fabs(xdos[j] - th) < width) ind[np++] = j;
      for (int j = 0; j < np; j++) {
        double dx = xdos[ind[j]] - th;
        y[ind[j]] += gamma2[i] * exp(-(dx * dx) / sigma2);
      }
    }
  }
  double scaling = 1.0 / (nvec * sqrt(sigma2 * M_PI));
  for (int j = 0; j < npts; j++) { y[j] *= scaling; ydos[j] = y[j]; }
  orc_simpson(xdos, y, npts);
  *neig = y[npts - 1] * n;
  free(V); free(vinit); free(dT); free(eT); free(EvalT); free(EvecT);
  free(gamma2); free(warr); free(y); free(ind);
  return 0;
}

/* SRC/landos.c:331-355 (integrates yi in place, then cuts). */
void orc_spslicer_lan(const double *xi, double *yi, int n_int, int npts,
                      double *sli) {
  orc_simpson(xi, yi, npts);
  double t = yi[0];
  double want = (yi[npts - 1] - yi[0]) / (double)n_int;
  int ls = 0;
  sli[0] = xi[0];
  for (int k = 1; k < npts; k++) {
    if (yi[k] - t >= want) {
      ls = ls + 1;
      if (ls <= n_int) sli[ls] = xi[k]; /* the reference writes unguarded */
      t = yi[k];
    }
  }
  sli[n_int] = xi[npts - 1];
}

/* ------------------------------------------------------------------------
 * Dense symmetric eigenproblem (oracle of SymEigenSolver, SRC/miscla.c,
 * LAPACK DSYEV in the reference): Householder reduction to tridiagonal form
 * Q^T A Q = T (Golub & Van Loan, Alg. 8.3.1), then the implicit QL solver
 * above on T, eigenvectors Z = Q Z_T.  A and Z are n x n column-major;
 * eigenvalues ascending.
 * --------------------------------------------------------------------- */
int orc_sym_eig(int n, const double *A, double *w, double *Z) {
  if (n <= 0) return 0;
  const size_t nn = (size_t)n * n;
  double *B = (double *)xmalloc(sizeof(double) * nn);
  double *Q = (double *)xmalloc(sizeof(double) * nn);
  double *u = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *p = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *d = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *e = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *zt = (double *)xmalloc(sizeof(double) * nn);
  memcpy(B, A, sizeof(double) * nn);
  memset(Q, 0, sizeof(double) * nn);
  for (int i = 0; i < n; i++) Q[(size_t)i * n + i] = 1.0;
#define EL(M, r, c) (M)[(size_t)(c) * n + (r)]
  for (int j = 0; j < n - 2; j++) {
    /* Householder vector u for B(j+1:n, j): (I - 2uu^T/u^Tu) x = -+|x| e1 */
    double nrm = 0.0;
    for (int r = j + 1; r < n; r++) nrm += EL(B, r, j) * EL(B, r, j);
    nrm = sqrt(nrm);
    if (nrm == 0.0) continue;
    const double x0 = EL(B, j + 1, j);
    const double alpha = x0 >= 0 ? -nrm : nrm;
    for (int r = 0; r < n; r++) u[r] = 0.0;
    for (int r = j + 1; r < n; r++) u[r] = EL(B, r, j);
    u[j + 1] -= alpha;
    double uu = 0.0;
    for (int r = j + 1; r < n; r++) uu += u[r] * u[r];
    if (uu == 0.0) continue;
    const double beta = 2.0 / uu;
    /* B <- H B H with H = I - beta u u^T: p = beta B u, K = beta/2 u^T p,
       q = p - K u, B -= u q^T + q u^T */
    for (int r = 0; r < n; r++) {
      double t = 0.0;
      for (int c = j + 1; c < n; c++) t += EL(B, r, c) * u[c];
      p[r] = beta * t;
    }
    double K = 0.0;
    for (int r = j + 1; r < n; r++) K += u[r] * p[r];
    K *= 0.5 * beta;
    for (int r = 0; r < n; r++) p[r] -= K * u[r];
    for (int c = 0; c < n; c++)
      for (int r = 0; r < n; r++) EL(B, r, c) -= u[r] * p[c] + p[r] * u[c];
    /* Q <- Q H */
    for (int r = 0; r < n; r++) {
      double t = 0.0;
      for (int c = j + 1; c < n; c++) t += EL(Q, r, c) * u[c];
      t *= beta;
      for (int c = j + 1; c < n; c++) EL(Q, r, c) -= t * u[c];
    }
  }
  for (int i = 0; i < n; i++) d[i] = EL(B, i, i);
  for (int i = 0; i < n - 1; i++) e[i] = EL(B, i + 1, i);
  const int info = orc_tridiag_eig(n, d, e, w, zt);
  if (Z) {
    for (int c = 0; c < n; c++)
      for (int r = 0; r < n; r++) {
        double t = 0.0;
        for (int k = 0; k < n; k++) t += EL(Q, r, k) * zt[(size_t)c * n + k];
        EL(Z, r, c) = t;
      }
  }
#undef EL
  free(B); free(Q); free(u); free(p); free(d); free(e); free(zt);
  return info;
}

/* ------------------------------------------------------------------------
 * Thick-restart Lanczos bounds of the spectrum: SRC/lantrbnd.c:31-408
 * (standard problem).  Restart dimension lanm, at most maxit products with
 * A, two Ritz pairs kept at each restart; bounds Ritz value -+ |spike|
 * (bndtype <= 1) or the Kato-Temple form (bndtype 2); stop when
 * t1 + t2 < tol (|lmin| + |lmax|).  vinit is not modified; lucky breakdowns
 * draw from the libc rand() stream.
 * --------------------------------------------------------------------- */
int orc_lan_trbounds(const orc_csr *A, int lanm, int maxit, double tol, const double *vinit,
                     int bndtype, double *lammin, double *lammax, int *its) {
  const int n = A->nrows;
  if (lanm > n) lanm = n;
  const int lanm1 = lanm + 1;
  if (lanm == n && maxit > n) maxit = n;
  const int min_inner_step = 5;
  const size_t ld = (size_t)n, T1 = (size_t)lanm1;
  double *V = (double *)xmalloc(sizeof(double) * ld * (size_t)lanm1);
  double *T = (double *)calloc(T1 * T1, sizeof(double));
  double *Tk = (double *)xmalloc(sizeof(double) * T1 * T1);
  double *Ev = (double *)xmalloc(sizeof(double) * T1 * T1);
  double *Rval = (double *)xmalloc(sizeof(double) * T1);
  double *R0 = (double *)xmalloc(sizeof(double) * ld);
  double *R1 = (double *)xmalloc(sizeof(double) * ld);
  double *work = (double *)xmalloc(sizeof(double) * 3 * T1);
  double s[3] = {0, 0, 0}, lmin = 0.0, lmax = 0.0, t1 = 0.0, t2 = 0.0;
  int it = 0, trlen = 0, rc = 0;
  memcpy(V, vinit, sizeof(double) * ld);
  double t = sqrt(dot(n, V, V));
  for (int i = 0; i < n; i++) V[i] /= t;
  while (it < maxit) {
    double wn = 0.0, beta = 0.0;
    int nwn = 0;
    int k = trlen > 0 ? trlen + 1 : 0;
    if (k + min_inner_step > lanm1) { rc = -1; break; }
    if (trlen > 0) {
      const int k1 = k - 1;
      double *v = V + (size_t)k1 * ld, *vnew = V + (size_t)(k1 + 1) * ld;
      orc_csrmv(A, v, vnew);
      for (int i = 0; i < trlen; i++) {
        T[(size_t)i * T1 + i] = Rval[i];
        wn += fabs(Rval[i]);
      }
      s[k1] = dot(n, v, vnew);
      for (int j = 0; j < k; j++) axpy(n, -s[j], V + (size_t)j * ld, vnew);
      for (int i = 0; i < k1; i++) {
        T[(size_t)trlen * T1 + i] = s[i];
        T[(size_t)i * T1 + trlen] = s[i];
        wn += 2.0 * fabs(s[i]);
      }
      T[(size_t)trlen * T1 + trlen] = s[k1];
      wn += fabs(s[k1]);
      beta = sqrt(dot(n, vnew, vnew));
      wn += 2.0 * beta;
      nwn += 3 * k;
      if (beta * nwn < 1e-14 * wn) {
        orc_rand_vec(n, vnew);
        orc_cgs_dgks(n, k, 2, V, (int)ld, vnew, &beta, work);
        for (int i = 0; i < n; i++) vnew[i] /= beta;
        beta = 0.0;
      } else {
        for (int i = 0; i < n; i++) vnew[i] /= beta;
      }
      T[(size_t)k1 * T1 + k] = beta;
      T[(size_t)k * T1 + k1] = beta;
    }
    while (k < lanm && it < maxit) {
      k++;
      double *v = V + (size_t)(k - 1) * ld, *vnew = V + (size_t)k * ld;
      const double *zold = k > 1 ? V + (size_t)(k - 2) * ld : NULL;
      orc_csrmv(A, v, vnew);
      it++;
      if (zold) axpy(n, -beta, zold, vnew);
      const double alpha = dot(n, v, vnew);
      T[(size_t)(k - 1) * T1 + (k - 1)] = alpha;
      wn += fabs(alpha);
      axpy(n, -alpha, v, vnew);
      orc_cgs_dgks(n, k, 2, V, (int)ld, vnew, &beta, work);
      wn += 2.0 * beta;
      nwn += 3;
      if (beta * nwn < 1e-14 * wn) {
        orc_rand_vec(n, vnew);
        orc_cgs_dgks(n, k, 2, V, (int)ld, vnew, &beta, work);
        for (int i = 0; i < n; i++) vnew[i] /= beta;
        beta = 0.0;
      } else {
        for (int i = 0; i < n; i++) vnew[i] /= beta;
      }
      T[(size_t)k * T1 + (k - 1)] = beta;
      T[(size_t)(k - 1) * T1 + k] = beta;
    }
    /* eigenpairs of T(0:k, 0:k) (column-major with leading dimension k) */
    for (int c = 0; c < k; c++)
      for (int r = 0; r < k; r++) Tk[(size_t)c * k + r] = T[(size_t)c * T1 + r];
    orc_sym_eig(k, Tk, Rval, Ev);
    s[0] = beta * Ev[(size_t)0 * k + (k - 1)];
    s[1] = beta * Ev[(size_t)(k - 1) * k + (k - 1)];
    if (bndtype <= 1) {
      t1 = fabs(s[0]);
      t2 = fabs(s[1]);
    } else {
      t1 = 2.0 * s[0] * s[0] / (Rval[1] - Rval[0]);
      t2 = 2.0 * s[1] * s[1] / (Rval[k - 1] - Rval[k - 2]);
    }
    lmin = Rval[0] - t1;
    lmax = Rval[k - 1] + t2;
    /* the two extreme Ritz vectors become V(:,0), V(:,1) */
    for (size_t i = 0; i < ld; i++) {
      double a0 = 0.0, a1 = 0.0;
      for (int j = 0; j < k; j++) {
        a0 += V[(size_t)j * ld + i] * Ev[(size_t)0 * k + j];
        a1 += V[(size_t)j * ld + i] * Ev[(size_t)(k - 1) * k + j];
      }
      R0[i] = a0;
      R1[i] = a1;
    }
    trlen = 2;
    memcpy(V, R0, sizeof(double) * ld);
    memcpy(V + ld, R1, sizeof(double) * ld);
    Rval[1] = Rval[k - 1];
    if (t1 + t2 < tol * (fabs(lmin) + fabs(lmax))) break;
    memset(T, 0, sizeof(double) * T1 * T1);
    memcpy(V + (size_t)trlen * ld, V + (size_t)k * ld, sizeof(double) * ld);
  }
  *lammin = lmin;
  *lammax = lmax;
  if (its) *its = it;
  free(V); free(T); free(Tk); free(Ev); free(Rval); free(R0); free(R1); free(work);
  return rc;
}
