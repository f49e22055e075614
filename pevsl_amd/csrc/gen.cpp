// gen.cpp -- the generalized problem A x = lambda B x on the device:
//   * setters of B, of its solve and of the L^T solve (SRC/pevsl.c:105-210);
//   * least-squares polynomial solves, B^{-1} or B^{-1/2} approximated by a
//     Chebyshev series in B (SRC/lspolapprox.c:63-322);
//   * the ifGenEv forms of ChebAv (SRC/chebpol.c:543-547), of the KPM
//     moments (SRC/kpmdos.c:105-134) and of ChebLanNr (SRC/cheblanNr.c:
//     125-435).
//
// B enters as a device CSR (SetBParcsr) or a host callback (SetBMatvec).  The
// B- and L^T-solves are the reference's SVFunc host callbacks, called with
// host copies of the vectors -- except that the library's own solves
// (pEVSL_LSPolSol: the fused Chebyshev step kernel on B's device CSR;
// pEVSL_ChebIterSolv1/2: Chebyshev iteration) are recognised and applied on
// the device, with no host round trip.  The
// vectors, the two Lanczos bases V and Z = B V and the reorthogonalization
// (rocBLAS DGEMVs) stay in HBM; host work is the O(k) Lanczos scalars and
// the tridiagonal eigenvectors.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "internal.hpp"

namespace pevsl {

// pEVSL_SetupLSPol* result (opaque to callers as in the reference).
struct LSPolData {
  int N = 0, n = 0, nfirst = 0;
  int max_deg = 0, deg = 0;
  double intv[2] = {0.0, 0.0};
  double cc = 0.0, dd = 1.0, tol = 0.0;
  std::vector<double> mu;
  pevsl_Parcsr *Bp = nullptr;  // its HBM copy is looked up at the first solve
  DBuf<double> wk;          // 3 n recurrence vectors + 2 n staging for host calls
  hipStream_t s = nullptr;  // stream of host-array calls (pEVSL_LSPolSol)
};

namespace {

// y = sum_{k <= deg} mu_k T_k((B - cc)/dd) x on the device (the recurrence of
// pnav, SRC/lspolapprox.c:245-288, fused per degree into one pass).
void lspol_apply(const LSPolData &L, const double *dx, double *dy, double *wk, hipStream_t s) {
  const DevCsr *Bd = dev_matrix_of(L.Bp);
  if (!Bd) fatal("LSPol: B has no device copy (pEVSL_ParcsrSetup)", __FILE__, __LINE__);
  const DevCsr &B = *Bd;
  const int n = B.n, m = L.deg;
  const double *mu = L.mu.data();
  if (m == 0) {
    copy_f64((size_t)n, dx, dy, s);
    scale_vec(n, dy, nullptr, mu[0], s);
    return;
  }
  const int lanes = auto_lanes(n ? (double)B.nnz / n : 1.0);
  double *b[3] = {wk, wk + n, wk + 2 * (size_t)n};
  const double *vk = dx, *vkm1 = nullptr;
  int nb = 0;
  for (int k = 1; k <= m; k++) {
    double *vkp1 = b[nb];
    nb = (nb + 1) % 3;
    cheb_step_vec<double>(B, vk, vkm1, vkp1, dy, L.cc, k == 1 ? 1.0 / L.dd : 2.0 / L.dd, mu[0],
                          mu[k], k == 1, lanes, s);
    vkm1 = vk;
    vk = vkp1;
  }
}

// sum_{j <= m} mu_j T_j(x) by the three-term recurrence (chebxPltd,
// SRC/chebpol.c:100-131, at one point)
double cheb_series(int m, const double *mu, double x) {
  double tkm1 = 0.0, tk = 1.0, y = mu[0];
  for (int k = 0; k < m; k++) {
    const double tkp1 = (k == 0 ? 1.0 : 2.0) * x * tk - tkm1;
    tkm1 = tk;
    tk = tkp1;
    y += mu[k + 1] * tk;
  }
  return y;
}

// Chebyshev coefficients of f on [a, b] by Gauss-Chebyshev quadrature on
// 4 max_deg nodes, the degree growing until the sum of the relative errors
// at 50 max_deg equispaced points is below tol (SRC/lspolapprox.c:63-146).
void find_lspol(double (*f)(double), LSPolData &L) {
  const double a = L.intv[0], b = L.intv[1];
  L.cc = (a + b) / 2;
  L.dd = (b - a) / 2;
  const int maxdeg = L.max_deg, npts = 4 * maxdeg;
  std::vector<double> th((size_t)npts), g((size_t)npts);
  for (int i = 0; i < npts; i++) {
    th[(size_t)i] = (2 * i + 1) * (M_PI / (2 * npts));
    g[(size_t)i] = f(L.cc + L.dd * cos(th[(size_t)i]));
  }
  const int nrm_pts = 50 * maxdeg;
  std::vector<double> xs((size_t)nrm_pts), fx((size_t)nrm_pts);
  linspace(-1.0, 1.0, nrm_pts, xs.data());
  for (int i = 0; i < nrm_pts; i++) fx[(size_t)i] = f(L.cc + L.dd * xs[(size_t)i]);
  L.mu.assign((size_t)std::max(maxdeg, 1), 0.0);
  int k = 0;
  for (k = 0; k < maxdeg; k++) {
    double sum = 0.0;
    for (int i = 0; i < npts; i++) sum += cos(k * th[(size_t)i]) * g[(size_t)i];
    L.mu[(size_t)k] = (k == 0 ? 1.0 / npts : 2.0 / npts) * sum;
    double err = 0.0;
    for (int i = 0; i < nrm_pts; i++)
      err += fabs((cheb_series(k, L.mu.data(), xs[(size_t)i]) - fx[(size_t)i]) / fx[(size_t)i]);
    if (err < L.tol) {
      k++;
      break;
    }
  }
  L.deg = k - 1;
}

double rec(double a) { return 1.0 / a; }
double isqrt(double a) { return 1.0 / sqrt(a); }

// host matvec installed in Bmv->func by SetBParcsr, for callers that invoke
// the hook directly with host arrays
void parcsr_mv_host(double *x, double *y, void *data) {
  pevsl_Parcsr *B = (pevsl_Parcsr *)data;
  pevsl_Parvec X, Y;
  pEVSL_ParvecCreateShell(B->ncol_global, B->ncol_local, B->first_col, B->comm, &X, x);
  pEVSL_ParvecCreateShell(B->nrow_global, B->nrow_local, B->first_row, B->comm, &Y, y);
  pEVSL_ParcsrMatvec(B, &X, &Y);
}

template <typename F>
void host_round_trip(Ctx &c, int n, F f, void *data, const double *dx, double *dy) {
  c.hb_x.resize((size_t)n);
  c.hb_y.resize((size_t)n);
  PV_HIP(hipMemcpyAsync(c.hb_x.data(), dx, sizeof(double) * n, hipMemcpyDeviceToHost,
                        c.stream));
  PV_HIP(hipStreamSynchronize(c.stream));
  f(c.hb_x.data(), c.hb_y.data(), data);
  PV_HIP(hipMemcpyAsync(dy, c.hb_y.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                        c.stream));
}

void solve_with(Ctx &c, pevsl_Data *p, SVFunc f, void *data, const double *dx, double *dy) {
  if (f == pEVSL_LSPolSol) {
    const LSPolData &L = *static_cast<const LSPolData *>(data);
    if (L.n != p->n) fatal("LSPol data of another size", __FILE__, __LINE__);
    lspol_apply(L, dx, dy, c.gen_lsw.ensure((size_t)3 * p->n), c.stream);
    return;
  }
  if (f == pEVSL_ChebIterSolv1 || f == pEVSL_ChebIterSolv2) {
    if (chebiter_size(data) != p->n) fatal("ChebIter data of another size", __FILE__, __LINE__);
    chebiter_device_solve(f, data, dx, dy, c.gen_lsw.ensure((size_t)3 * p->n), c.stream);
    return;
  }
  host_round_trip(c, p->n, f, data, dx, dy);
}

// deterministic device dot product (two-stage, fixed order)
double ddot(Ctx &c, int n, const double *x, const double *y) {
  double *pt = c.gen_dot.ensure((size_t)dot_part_blocks(n) + 8);
  const int nb = dot_partials(n, x, y, pt, c.stream);
  reduce_partials(pt, nb, 1, pt + nb, c.stream);
  double r = 0.0;
  PV_HIP(hipMemcpyAsync(&r, pt + nb, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PV_HIP(hipStreamSynchronize(c.stream));
  return r;
}

}  // namespace

void apply_B(Ctx &c, pevsl_Data *p, const double *dx, double *dy) {
  if (c.B) {
    spmv_f64(*c.B, dx, dy, auto_lanes(c.B->n ? (double)c.B->nnz / c.B->n : 1.0), c.stream);
    return;
  }
  if (!p->Bmv || !p->Bmv->func) fatal("matrix B not set", __FILE__, __LINE__);
  host_round_trip(c, p->n, p->Bmv->func, p->Bmv->data, dx, dy);
}

void solve_B(Ctx &c, pevsl_Data *p, const double *dx, double *dy) {
  if (!p->Bsol || !p->Bsol->func) fatal("B solve not set", __FILE__, __LINE__);
  solve_with(c, p, p->Bsol->func, p->Bsol->data, dx, dy);
}

void solve_LT(Ctx &c, pevsl_Data *p, const double *dx, double *dy) {
  if (!p->LTsol || !p->LTsol->func) fatal("L^T solve not set", __FILE__, __LINE__);
  solve_with(c, p, p->LTsol->func, p->LTsol->data, dx, dy);
}

double gen_dot(Ctx &c, int n, const double *x, const double *y) { return ddot(c, n, x, y); }

// for i < ngs: w = T(:, 0:k-1)^T v; v = v - N(:, 0:k-1) w
// (CGS_DGKS2, SRC/miscla.c:247-262), T and N column-major with leading dim ld
void cgs2(Ctx &c, int n, int k, int ngs, const double *T, const double *N, size_t ld,
          double *v) {
  double *w = c.gen_cgs.ensure((size_t)std::max(k, 1));
  const double one = 1.0, zero = 0.0, mone = -1.0;
  for (int i = 0; i < ngs; i++) {
    PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_transpose, n, k, &one, T, (int)ld, v, 1,
                          &zero, w, 1));
    PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_none, n, k, &mone, N, (int)ld, w, 1, &one,
                          v, 1));
  }
}

// y = p(A B^{-1}) v in the reference's generalized ChebAv form: per degree
// w2 = B \ v_k, v_{k+1} = t (A w2 - cc v_k) - v_{k-1}, y += mu_k v_{k+1}.
void cheb_filter_gen(Ctx &c, pevsl_Data *p, const pevsl_Polparams *pol, const double *dv,
                     double *dy) {
  const int n = p->n, m = pol->deg;
  const double *mu = pol->mu;
  const double cc = pol->cc, dd = pol->dd;
  const hipStream_t s = c.stream;
  double *w = c.gen_fw.ensure((size_t)4 * n);
  double *vk = w, *vkp1 = w + n, *vkm1 = w + 2 * (size_t)n, *w2 = w + 3 * (size_t)n;
  copy_f64((size_t)n, dv, vk, s);
  copy_f64((size_t)n, vk, dy, s);
  scale_vec(n, dy, nullptr, mu[0], s);
  for (int k = 1; k <= m; k++) {
    const double t = k == 1 ? 1.0 / dd : 2.0 / dd;
    solve_B(c, p, vk, w2);
    apply_A(c, w2, vkp1);
    cheb_update(n, vk, k > 1 ? vkm1 : nullptr, vkp1, dy, cc, t, mu[k], s);
    double *tmp = vkm1;
    vkm1 = vk;
    vk = vkp1;
    vkp1 = tmp;
  }
  c.kstats.filter_steps += m;
}

// dots[m * Mdeg + k] = <v_m, T_{k+1}(B^{-1}A scaled) w_m> of the generalized
// KPM (SRC/kpmdos.c:105-156): w = L^{-T} v for a unit random v, v <- B w.
void kpm_moments_gen(Ctx &c, pevsl_Data *p, int Mdeg, int nvec, double ctr, double wid,
                     std::vector<double> &dots) {
  const int n = p->n;
  const hipStream_t s = c.stream;
  DBuf<double> buf, dd;
  double *v = buf.ensure((size_t)5 * n), *w = v + n, *vk = v + 2 * (size_t)n,
         *vkm1 = v + 3 * (size_t)n, *vkp1 = v + 4 * (size_t)n;
  double *pt = dd.ensure((size_t)Mdeg + dot_part_blocks(n) + 8);
  std::vector<double> hd((size_t)Mdeg);
  for (int m = 0; m < nvec; m++) {
    rand_vectors(c, n, 1, v, (size_t)n);
    scale_vec(n, v, nullptr, 1.0 / sqrt(ddot(c, n, v, v)), s);
    solve_LT(c, p, v, w);
    apply_B(c, p, w, v);
    copy_f64((size_t)n, w, vk, s);
    PV_HIP(hipMemsetAsync(vkm1, 0, sizeof(double) * n, s));
    for (int k = 0; k < Mdeg; k++) {
      apply_A(c, vk, w);
      solve_B(c, p, w, vkp1);
      const double scal = (k == 0 ? 1.0 : 2.0) / wid;
      cheb_update(n, vk, vkm1, vkp1, nullptr, ctr, scal, 0.0, s);
      double *t = vkm1;
      vkm1 = vk;
      vk = vkp1;
      vkp1 = t;
      const int nb = dot_partials(n, vk, v, pt + Mdeg, s);
      reduce_partials(pt + Mdeg, nb, 1, pt + k, s);
      c.kstats.filter_steps++;
    }
    PV_HIP(hipMemcpyAsync(hd.data(), pt, sizeof(double) * Mdeg, hipMemcpyDeviceToHost, s));
    PV_HIP(hipStreamSynchronize(s));
    std::copy(hd.begin(), hd.end(), dots.begin() + (long)m * Mdeg);
  }
}

// Tridiagonal matrices of the generalized Lanczos DOS (SRC/landos.c:129-275,
// ifGenEv): per random vector, v = L^{-T} r / |.|_B, maxit Lanczos steps with
// bases V and Z = B V, CGS_DGKS2 and B-solves; dT/eT[m * maxit + k].
void landos_gen(Ctx &c, pevsl_Data *p, int nvec, int maxit, std::vector<double> &dT,
                std::vector<double> &eT) {
  const int n = p->n;
  const hipStream_t s = c.stream;
  const size_t ld = ((size_t)n + 63) / 64 * 64;
  DBuf<double> Vb, Zb, rb;
  double *V = Vb.ensure(ld * (size_t)(maxit + 1)), *Z = Zb.ensure(ld * (size_t)(maxit + 1));
  double *r = rb.ensure(ld);
  auto vcol = [&](int j) { return V + (size_t)j * ld; };
  auto zcol = [&](int j) { return Z + (size_t)j * ld; };
  for (int m = 0; m < nvec; m++) {
    rand_vectors(c, n, 1, r, ld);
    solve_LT(c, p, r, vcol(0));
    apply_B(c, p, vcol(0), zcol(0));
    double t = 1.0 / sqrt(ddot(c, n, vcol(0), zcol(0)));
    scale_vec(n, zcol(0), nullptr, t, s);
    scale_vec(n, vcol(0), nullptr, t, s);
    double wn = 0.0, beta = 0.0;
    int nwn = 0;
    for (int k = 0; k < maxit; k++) {
      double *v = vcol(k), *z = zcol(k), *vnew = vcol(k + 1), *znew = zcol(k + 1);
      apply_A(c, v, znew);
      if (k > 0) axpy_vec(n, -beta, zcol(k - 1), znew, s);
      const double alpha = ddot(c, n, v, znew);
      dT[(size_t)m * maxit + k] = alpha;
      wn += fabs(alpha);
      axpy_vec(n, -alpha, z, znew, s);
      cgs2(c, n, k + 1, NGS_MAX, V, Z, ld, znew);
      solve_B(c, p, znew, vnew);
      beta = sqrt(ddot(c, n, vnew, znew));
      wn += 2.0 * beta;
      nwn += 3;
      if (beta * nwn < orthTol * wn) {
        rand_vectors(c, n, 1, vnew, ld);
        cgs2(c, n, k + 1, NGS_MAX, Z, V, ld, vnew);
        apply_B(c, p, vnew, znew);
        beta = sqrt(ddot(c, n, vnew, znew));
        t = 1.0 / beta;
        scale_vec(n, vnew, nullptr, t, s);
        scale_vec(n, znew, nullptr, t, s);
        beta = 0.0;
      } else {
        t = 1.0 / beta;
        scale_vec(n, vnew, nullptr, t, s);
        scale_vec(n, znew, nullptr, t, s);
      }
      eT[(size_t)m * maxit + k] = beta;
    }
  }
}

// Non-restarted filtered Lanczos for A x = lambda B x (SRC/cheblanNr.c:
// 54-448, ifGenEv branches): bases V and Z = B V in HBM, the filter
// p(A B^{-1}) applied to z, CGS_DGKS2 against (V, Z), B-inner products.
int cheblan_gen(Ctx &c, pevsl_Data *p, double *intv, int maxit, double tol,
                pevsl_Parvec *vinit, pevsl_Polparams *pol, int *nevOut, double **lamo,
                pevsl_Parvecs **Wo, double **reso, FILE *fstats) {
  const double tall = wtime();
  const int n = p->n, N = p->N;
  const hipStream_t s = c.stream;
  const int Ntest = 30, cycle = 20;
  maxit = std::min(N, maxit);
  const double bar = pol->bar, gamB = pol->gam;
  const double aa = intv[0], bb = intv[1];
  if (gamB > 1.0 || gamB < -1.0) {
    fprintf(stdout, "gamB error %.15e\n", gamB);
    return -1;
  }
  const size_t ld = ((size_t)n + 63) / 64 * 64;
  DBuf<double> Vb, Zb, wkb;
  double *V = Vb.ensure(ld * (size_t)(maxit + 1)), *Z = Zb.ensure(ld * (size_t)(maxit + 1));
  double *wk = wkb.ensure((size_t)3 * ld), *w2 = wk + ld, *vr = wk + 2 * ld;
  auto vcol = [&](int j) { return V + (size_t)j * ld; };
  auto zcol = [&](int j) { return Z + (size_t)j * ld; };
  std::vector<double> dT((size_t)maxit), eT((size_t)maxit);
  // V(:,0) = p(A B^{-1}) vinit (the reference filters the start vector)
  PV_HIP(hipMemcpyAsync(vr, vinit->data, sizeof(double) * n, hipMemcpyHostToDevice, s));
  cheb_filter_gen(c, p, pol, vr, vcol(0));
  // unit B-norm: z = B v / sqrt(v' B v), v /= sqrt(v' B v)
  apply_B(c, p, vcol(0), zcol(0));
  double t = 1.0 / sqrt(ddot(c, n, vcol(0), zcol(0)));
  scale_vec(n, zcol(0), nullptr, t, s);
  scale_vec(n, vcol(0), nullptr, t, s);
  double wn = 0.0, tr0 = 0.0, tr1 = 0.0, beta = 0.0;
  int nwn = 0, kdim = 0, k = 0, nconv = 0;
  for (k = 0; k < maxit; k++) {
    double *v = vcol(k), *z = zcol(k), *vnew = vcol(k + 1), *znew = zcol(k + 1);
    cheb_filter_gen(c, p, pol, z, znew);
    if (k > 0) axpy_vec(n, -beta, zcol(k - 1), znew, s);
    const double alpha = ddot(c, n, v, znew);
    dT[(size_t)k] = alpha;
    wn += fabs(alpha);
    axpy_vec(n, -alpha, z, znew, s);
    // znew -= Z(:,0:k) V(:,0:k)' znew, then vnew = B \ znew
    const double tr = wtime();
    cgs2(c, n, k + 1, NGS_MAX, V, Z, ld, znew);
    p->stats->t_reorth += wtime() - tr;
    solve_B(c, p, znew, vnew);
    beta = sqrt(ddot(c, n, vnew, znew));
    wn += 2.0 * beta;
    nwn += 3;
    if (beta * nwn < orthTol * wn) {
      if (fstats) fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", k, beta);
      // a new filtered random vector, B-orthogonalized against V
      rand_vectors(c, n, 1, vr, ld);
      cheb_filter_gen(c, p, pol, vr, vnew);
      cgs2(c, n, k + 1, NGS_MAX, Z, V, ld, vnew);
      apply_B(c, p, vnew, znew);
      beta = sqrt(ddot(c, n, vnew, znew));
      t = 1.0 / beta;
      scale_vec(n, vnew, nullptr, t, s);
      scale_vec(n, znew, nullptr, t, s);
      beta = 0.0;
    } else {
      t = 1.0 / beta;
      scale_vec(n, vnew, nullptr, t, s);
      scale_vec(n, znew, nullptr, t, s);
    }
    eT[(size_t)k] = beta;
    if ((k < Ntest || (k - Ntest) % cycle != 0) && k != maxit - 1) continue;
    // restricted trace of T(0:k, 0:k): eigenvalues >= bar (Sturm bisection)
    kdim = k + 1;
    double *td = c.gen_tri.ensure((size_t)3 * kdim + 16);
    PV_HIP(hipMemcpyAsync(td, dT.data(), sizeof(double) * kdim, hipMemcpyHostToDevice, s));
    if (kdim > 1)
      PV_HIP(hipMemcpyAsync(td + kdim, eT.data(), sizeof(double) * (kdim - 1),
                            hipMemcpyHostToDevice, s));
    const double te = wtime();
    tridiag_trace_above(kdim, td, td + kdim, bar - 5 * DBL_EPSILON, td + 2 * (size_t)kdim,
                        td + 3 * (size_t)kdim, s);
    double o[2];
    PV_HIP(hipMemcpyAsync(o, td + 3 * (size_t)kdim, sizeof(double) * 2, hipMemcpyDeviceToHost,
                          s));
    PV_HIP(hipStreamSynchronize(s));
    p->stats->t_eig += wtime() - te;
    tr1 = o[0];
    nconv = (int)o[1];
    if (fstats) fprintf(fstats, "  k %4d:   nconv %4d  tr1 %21.15e\n", k, nconv, tr1);
    if (fabs(tr1 - tr0) < 2e-12 || fabs(tr1) + fabs(tr0) < 1e-10) break;
    tr0 = tr1;
  }
  if (k >= maxit)
    fprintf(fstats ? fstats : stdout,
            "The max number of iterations [%d] has been reached. The eigenvalues computed may "
            "not have converged\n",
            maxit);
  // eigenpairs of T(0:kdim-1) on the device (divide and conquer)
  const double tt = wtime();
  DBuf<double> D, E, C;
  DBuf<int> info;
  D.ensure((size_t)kdim);
  E.ensure((size_t)std::max(kdim, 1));
  C.ensure((size_t)kdim * kdim);
  info.ensure(1);
  PV_HIP(hipMemcpyAsync(D.p, dT.data(), sizeof(double) * kdim, hipMemcpyHostToDevice, s));
  if (kdim > 1)
    PV_HIP(hipMemcpyAsync(E.p, eT.data(), sizeof(double) * (kdim - 1), hipMemcpyHostToDevice,
                          s));
  PV_BLAS(rocsolver_dstedc(c.blas, rocblas_evect_tridiagonal, kdim, D.p, E.p, C.p, kdim,
                           info.p));
  std::vector<double> ev((size_t)kdim);
  PV_HIP(hipMemcpyAsync(ev.data(), D.p, sizeof(double) * kdim, hipMemcpyDeviceToHost, s));
  PV_HIP(hipStreamSynchronize(s));
  // Ritz pairs with rho(lambda) >= bar, B-normalized, accepted in [a, b]
  // with ||A u - t B u|| < tol (SRC/cheblanNr.c:352-413)
  std::vector<double> Lam, Res, Vecs;
  double *u = vr;
  const double one = 1.0, zero = 0.0;
  for (int i = 0; i < kdim; i++) {
    if (ev[(size_t)i] < bar) continue;
    PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_none, n, kdim, &one, V, (int)ld,
                          C.p + (size_t)i * kdim, 1, &zero, u, 1));
    apply_B(c, p, u, w2);
    t = sqrt(ddot(c, n, u, w2));
    if (t == 0.0) return 2;
    t = 1.0 / t;
    scale_vec(n, u, nullptr, t, s);
    scale_vec(n, w2, nullptr, t, s);
    apply_A(c, u, wk);
    t = ddot(c, n, wk, u);
    if (t < aa - DBL_EPSILON || t > bb + DBL_EPSILON) continue;
    axpy_vec(n, -t, w2, wk, s);
    const double res0 = sqrt(ddot(c, n, wk, wk));
    if (res0 < tol) {
      Lam.push_back(t);
      Res.push_back(res0);
      if (Wo && c.keep_vectors) {
        Vecs.resize(Vecs.size() + (size_t)n);
        PV_HIP(hipMemcpyAsync(Vecs.data() + Vecs.size() - n, u, sizeof(double) * n,
                              hipMemcpyDeviceToHost, s));
        PV_HIP(hipStreamSynchronize(s));
      }
    }
  }
  p->stats->t_ritz += wtime() - tt;
  const int nev = (int)Lam.size();
  *nevOut = nev;
  const size_t cap = (size_t)std::max(maxit, std::max(nev, 1));
  double *lam, *res;
  PEVSL_MALLOC(lam, cap, double);
  PEVSL_MALLOC(res, cap, double);
  for (int i = 0; i < nev; i++) {
    lam[i] = Lam[(size_t)i];
    res[i] = Res[(size_t)i];
  }
  *lamo = lam;
  *reso = res;
  if (Wo) {
    if (c.keep_vectors) {
      pevsl_Parvecs *W;
      PEVSL_MALLOC(W, 1, pevsl_Parvecs);
      pEVSL_ParvecsDuplParvec(vinit, std::max(nev, 1), vinit->n_local, W);
      W->n_vecs = nev;
      if (nev) memcpy(W->data, Vecs.data(), sizeof(double) * (size_t)nev * n);
      *Wo = W;
    } else {
      *Wo = NULL;
    }
  }
  pevsl_LanInfo li{};
  li.its = k < maxit ? k + 1 : maxit;
  li.kdim = kdim;
  li.nconv = nconv;
  li.nev = nev;
  li.deg = pol->deg;
  li.tr = tr1;
  li.t_total = wtime() - tall;
  c.info.push_back(li);
  p->stats->t_iter = wtime() - tall;
  return 0;
}

}  // namespace pevsl

using namespace pevsl;

extern "C" {

int pEVSL_SetBParcsr(pevsl_Data *p, pevsl_Parcsr *B) {
  const DevCsr *M = dev_matrix_of(B);
  if (!p || !M) return -1;
  p->N = B->ncol_global;
  p->n = B->ncol_local;
  p->nfirst = B->first_col;
  if (!p->Bmv) PEVSL_CALLOC(p->Bmv, 1, pevsl_Matvec);
  p->Bmv->func = parcsr_mv_host;
  p->Bmv->data = (void *)B;
  ctx_of(p).B = M;
  return 0;
}

int pEVSL_SetBMatvec(pevsl_Data *p, MVFunc func, void *data) {
  if (!p) return -1;
  if (!p->Bmv) PEVSL_CALLOC(p->Bmv, 1, pevsl_Matvec);
  p->Bmv->func = func;
  p->Bmv->data = data;
  ctx_of(p).B = nullptr;
  return 0;
}

int pEVSL_SetBSol(pevsl_Data *p, SVFunc func, void *data) {
  if (!p) return -1;
  if (!p->Bsol) PEVSL_CALLOC(p->Bsol, 1, pevsl_Bsol);
  p->Bsol->func = func;
  p->Bsol->data = data;
  return 0;
}

int pEVSL_SetLTSol(pevsl_Data *p, SVFunc func, void *data) {
  if (!p) return -1;
  if (!p->LTsol) PEVSL_CALLOC(p->LTsol, 1, pevsl_Ltsol);
  p->LTsol->func = func;
  p->LTsol->data = data;
  return 0;
}

int pEVSL_SetGenEig(pevsl_Data *p) {
  if (!p) return -1;
  p->ifGenEv = 1;
  return 0;
}

void pEVSL_SetupLSPol(int max_deg, double tol, double lmin, double lmax, pevsl_Parcsr *B,
                      double (*ffun)(double), void **vdata) {
  LSPolData *L = new LSPolData();
  L->N = B->ncol_global;
  L->n = B->ncol_local;
  L->nfirst = B->first_col;
  L->max_deg = max_deg;
  L->intv[0] = lmin;
  L->intv[1] = lmax;
  L->tol = tol;
  L->Bp = B;
  find_lspol(ffun, *L);
  *vdata = L;
}

void pEVSL_SetupLSPolRec(int max_deg, double tol, double lmin, double lmax, pevsl_Parcsr *B,
                         void **vdata) {
  pEVSL_SetupLSPol(max_deg, tol, lmin, lmax, B, rec, vdata);
}

void pEVSL_SetupLSPolSqrt(int max_deg, double tol, double lmin, double lmax, pevsl_Parcsr *B,
                          void **vdata) {
  pEVSL_SetupLSPol(max_deg, tol, lmin, lmax, B, isqrt, vdata);
}

void pEVSL_LSPolSol(double *db, double *dx, void *data) {
  LSPolData &L = *static_cast<LSPolData *>(data);
  const int n = L.n;
  if (!L.s) PV_HIP(hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking));
  double *w = L.wk.ensure((size_t)5 * n);
  PV_HIP(hipMemcpyAsync(w + 3 * (size_t)n, db, sizeof(double) * n, hipMemcpyHostToDevice, L.s));
  lspol_apply(L, w + 3 * (size_t)n, w + 4 * (size_t)n, w, L.s);
  PV_HIP(hipMemcpyAsync(dx, w + 4 * (size_t)n, sizeof(double) * n, hipMemcpyDeviceToHost, L.s));
  PV_HIP(hipStreamSynchronize(L.s));
}

int pEVSL_LSPolInfo(void *vdata, int *deg, double *mu, double *cc, double *dd) {
  if (!vdata) return -1;
  const LSPolData &L = *static_cast<const LSPolData *>(vdata);
  if (deg) *deg = L.deg;
  if (mu) std::copy(L.mu.begin(), L.mu.begin() + L.deg + 1, mu);
  if (cc) *cc = L.cc;
  if (dd) *dd = L.dd;
  return 0;
}

void pEVSL_LSPolFree(void *vdata) {
  LSPolData *L = static_cast<LSPolData *>(vdata);
  if (!L) return;
  if (L->s) {
    (void)hipStreamSynchronize(L->s);
    (void)hipStreamDestroy(L->s);
  }
  delete L;
}

}  // extern "C"
