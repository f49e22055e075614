// itsol.cpp -- Chebyshev iteration for the B-solves of a generalized problem
// (EXTERNAL/ITERSOL/chebiter.c:77-308): x ~ B^{-1} b after deg steps, given
// bounds [lb, ub] of B's spectrum, in the two published forms the reference
// offers -- the Templates form (Solv1, chebiter.c:121-200) and the form of
// Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., p. 399 (Solv2,
// chebiter.c:204-270).  Both run on the device: one product with B's HBM copy
// and a few fused vector updates per step.  Called with host arrays
// (pEVSL_ChebIterSolv1/2) the vectors are staged through HBM; installed as
// an instance's B-solve (pEVSL_SetBSol) they run on device vectors with no
// host copies.
#include <cmath>
#include <cstdio>

#include "internal.hpp"

namespace pevsl {

struct ChebIterData {
  double lb = 0.0, ub = 0.0;
  int deg = 0;
  int N = 0, n = 0, nfirst = 0;
  pevsl_Parcsr *Bp = nullptr;
  DBuf<double> wk;          // w, r, p + 2 n staging for host-array calls
  hipStream_t s = nullptr;  // stream of host-array calls
  unsigned long n_chebmv = 0;
  double t_chebmv = 0.0;
};

namespace {

void bmv(const ChebIterData &C, const double *x, double *y, hipStream_t s) {
  const DevCsr *B = dev_matrix_of(C.Bp);
  if (!B) fatal("ChebIter: B has no device copy (pEVSL_ParcsrSetup)", __FILE__, __LINE__);
  spmv_f64(*B, x, y, auto_lanes(B->n ? (double)B->nnz / B->n : 1.0), s);
}

// Templates form; x is left untouched when deg < 1 (chebiter.c:152-155)
void solv1(ChebIterData &C, const double *b, double *x, double *wk, hipStream_t s) {
  const int n = C.n, deg = C.deg;
  if (deg < 1) return;
  double *w = wk, *r = wk + n, *p = wk + 2 * (size_t)n;
  const double d = (C.ub + C.lb) * 0.5, c = (C.ub - C.lb) * 0.5;
  double alp = 2.0 / d;
  copy_f64((size_t)n, b, p, s);          // p = b
  copy_f64((size_t)n, p, x, s);          // x = alp p
  scale_vec(n, x, nullptr, alp, s);
  bmv(C, x, w, s);                       // r = b - B x
  copy_f64((size_t)n, b, r, s);
  axpy_vec(n, -1.0, w, r, s);
  C.n_chebmv++;
  for (int i = 1; i < deg; i++) {
    const double t = c * alp * 0.5, bet = t * t;
    alp = 1.0 / (d - bet);
    scale_vec(n, p, nullptr, bet, s);    // p = r + bet p
    axpy_vec(n, 1.0, r, p, s);
    axpy_vec(n, alp, p, x, s);           // x = x + alp p
    bmv(C, x, w, s);                     // r = b - B x
    copy_f64((size_t)n, b, r, s);
    axpy_vec(n, -1.0, w, r, s);
    C.n_chebmv++;
  }
}

// Saad's form with a zero initial guess
void solv2(ChebIterData &C, const double *b, double *x, double *wk, hipStream_t s) {
  const int n = C.n, deg = C.deg;
  double *w = wk, *r = wk + n, *d = wk + 2 * (size_t)n;
  const double theta = (C.ub + C.lb) * 0.5, delta = (C.ub - C.lb) * 0.5;
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  fill_zero_f64((size_t)n, x, s);
  copy_f64((size_t)n, b, r, s);
  copy_f64((size_t)n, r, d, s);          // d = r / theta
  scale_vec(n, d, nullptr, 1.0 / theta, s);
  for (int i = 0; i < deg; i++) {
    axpy_vec(n, 1.0, d, x, s);           // x = x + d
    bmv(C, d, w, s);                     // r = r - B d
    axpy_vec(n, -1.0, w, r, s);
    const double rho1 = 1.0 / (2.0 * sigma - rho);
    scale_vec(n, d, nullptr, rho1 * rho, s);  // d = rho1 rho d + 2 rho1/delta r
    axpy_vec(n, 2.0 * rho1 / delta, r, d, s);
    rho = rho1;
    C.n_chebmv++;
  }
}

void host_solve(ChebIterData &C, const double *db, double *dx, int type) {
  const int n = C.n;
  if (!C.s) PV_HIP(hipStreamCreateWithFlags(&C.s, hipStreamNonBlocking));
  double *w = C.wk.ensure((size_t)5 * n), *b = w + 3 * (size_t)n, *x = w + 4 * (size_t)n;
  PV_HIP(hipMemcpyAsync(b, db, sizeof(double) * n, hipMemcpyHostToDevice, C.s));
  PV_HIP(hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyHostToDevice, C.s));
  EvTimer ev;
  ev.start(C.s);
  if (type == 1)
    solv1(C, b, x, w, C.s);
  else
    solv2(C, b, x, w, C.s);
  ev.stop(C.s);
  C.t_chebmv += ev.seconds();
  PV_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyDeviceToHost, C.s));
  PV_HIP(hipStreamSynchronize(C.s));
}

}  // namespace

// Device application when f is one of the library's Chebyshev-iteration
// solves; false otherwise.  wk: 3 n doubles.
bool chebiter_device_solve(SVFunc f, void *data, const double *dx, double *dy, double *wk,
                           hipStream_t s) {
  if (f != pEVSL_ChebIterSolv1 && f != pEVSL_ChebIterSolv2) return false;
  ChebIterData &C = *static_cast<ChebIterData *>(data);
  EvTimer ev;
  ev.start(s);
  if (f == pEVSL_ChebIterSolv1)
    solv1(C, dx, dy, wk, s);
  else
    solv2(C, dx, dy, wk, s);
  ev.stop(s);
  C.t_chebmv += ev.seconds();
  return true;
}

int chebiter_size(void *data) { return static_cast<ChebIterData *>(data)->n; }

}  // namespace pevsl

using namespace pevsl;

extern "C" {

int pEVSL_ChebIterSetup(double lmin, double lmax, int deg, pevsl_Parcsr *A, void **data) {
  if (!A || !data) return -1;
  ChebIterData *C = new ChebIterData();
  C->deg = deg > 0 ? deg : 0;
  C->lb = lmin;
  C->ub = lmax;
  C->N = A->ncol_global;
  C->n = A->ncol_local;
  C->nfirst = A->first_col;
  C->Bp = A;
  *data = C;
  return 0;
}

void pEVSL_ChebIterSolv1(double *db, double *dx, void *data) {
  host_solve(*static_cast<ChebIterData *>(data), db, dx, 1);
}

void pEVSL_ChebIterSolv2(double *db, double *dx, void *data) {
  host_solve(*static_cast<ChebIterData *>(data), db, dx, 2);
}

double *pEVSL_ChebIterGetRes(void *data) {
  (void)data;  // the convergence history is not kept (SAVE_CONV_HIST 0)
  return NULL;
}

void pEVSL_ChebIterStatsPrint(void *data, FILE *fstats) {
  const ChebIterData &C = *static_cast<const ChebIterData *>(data);
  if (!fstats) return;
  fprintf(fstats, "= = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = =\n");
  if (C.n_chebmv)
    fprintf(fstats, "   Matvec in ChebIter        :  %f (%8ld, avg %f)\n", C.t_chebmv,
            (long)C.n_chebmv, C.t_chebmv / C.n_chebmv);
  fprintf(fstats, "= = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = =\n");
}

void pEVSL_ChebIterFree(void *vdata) {
  ChebIterData *C = static_cast<ChebIterData *>(vdata);
  if (!C) return;
  if (C->s) {
    (void)hipStreamSynchronize(C->s);
    (void)hipStreamDestroy(C->s);
  }
  delete C;
}

}  // extern "C"
