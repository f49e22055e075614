// lantrbnd.cpp -- spectrum bounds by thick-restart Lanczos
// (pEVSL_LanTrbounds, SRC/lantrbnd.c:31-408), standard and generalized
// problems.
//
// The restart basis V (lanm + 1 FP64 columns) lives in HBM and uses the
// batched Lanczos kernels with one slice (kernels.hip k_lan_*): products with
// A on the device matrix, the three-term update fused with alpha, full
// CGS-DGKS reorthogonalization with the DGKS decision on the device.  The
// projected matrix T (at most 51 x 51, tridiagonal plus the arrowhead of the
// restart) is diagonalized on the host by cyclic Jacobi; the two extreme
// Ritz vectors are DGEMVs over the basis.  Control flow, bounds (simple or
// Kato-Temple) and the stopping rule are the reference's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.hpp"

namespace pevsl {
namespace {

// Eigenpairs of the symmetric n x n matrix A (column-major, destroyed):
// cyclic Jacobi rotations until the off-diagonal mass is below eps times the
// matrix norm; w ascending, Z column-major eigenvectors.
void jacobi_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &Z) {
  auto at = [&](int r, int c) -> double & { return A[(size_t)c * n + r]; };
  Z.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) Z[(size_t)i * n + i] = 1.0;
  double fro = 0.0;
  for (double v : A) fro += v * v;
  fro = std::sqrt(fro);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int c = 0; c < n; c++)
      for (int r = 0; r < c; r++) off += at(r, c) * at(r, c);
    if (std::sqrt(2.0 * off) <= 1e-15 * fro || off == 0.0) break;
    for (int p = 0; p < n - 1; p++) {
      for (int q = p + 1; q < n; q++) {
        const double apq = at(p, q);
        if (apq == 0.0) continue;
        const double theta = (at(q, q) - at(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        // A <- J^T A J on rows/columns p, q
        for (int k = 0; k < n; k++) {
          const double akp = at(k, p), akq = at(k, q);
          at(k, p) = cs * akp - sn * akq;
          at(k, q) = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = at(p, k), aqk = at(q, k);
          at(p, k) = cs * apk - sn * aqk;
          at(q, k) = sn * apk + cs * aqk;
        }
        at(p, q) = at(q, p) = 0.0;
        for (int k = 0; k < n; k++) {
          double &zp = Z[(size_t)p * n + k], &zq = Z[(size_t)q * n + k];
          const double a = zp, b = zq;
          zp = cs * a - sn * b;
          zq = sn * a + cs * b;
        }
      }
    }
  }
  std::vector<int> idx((size_t)n);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<double> d((size_t)n);
  for (int i = 0; i < n; i++) d[(size_t)i] = at(i, i);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[(size_t)a] < d[(size_t)b]; });
  w.resize((size_t)n);
  std::vector<double> Zs((size_t)n * n);
  for (int j = 0; j < n; j++) {
    w[(size_t)j] = d[(size_t)idx[(size_t)j]];
    std::copy(Z.begin() + (long)idx[(size_t)j] * n, Z.begin() + (long)(idx[(size_t)j] + 1) * n,
              Zs.begin() + (long)j * n);
  }
  Z.swap(Zs);
}

// Generalized problem (SRC/lantrbnd.c ifGenEv branches): bases V and
// Z = B V in HBM, B-inner products, B-solves through the instance's solve,
// CGS_DGKS2 by rocBLAS DGEMVs; T and the bounds as in the standard problem.
int lan_trbounds_gen(pevsl_Data *pevsl, Ctx &c, int lanm, int maxit, double tol,
                     pevsl_Parvec *vinit, int bndtype, double *lammin, double *lammax,
                     FILE *fstats) {
  const int n = pevsl->n, lanm1 = lanm + 1;
  const int min_inner_step = 5;
  const hipStream_t st = c.stream;
  const size_t ld = ((size_t)n + 63) / 64 * 64;
  DBuf<double> Vb, Zb, rb, sb;
  double *V = Vb.ensure(ld * lanm1), *Z = Zb.ensure(ld * lanm1);
  double *rv = rb.ensure(4 * ld), *sv = sb.ensure(2 * (size_t)lanm1 + 8);
  auto vcol = [&](int j) { return V + (size_t)j * ld; };
  auto zcol = [&](int j) { return Z + (size_t)j * ld; };
  auto dot = [&](const double *x, const double *y) { return gen_dot(c, n, x, y); };
  const double one = 1.0, zero = 0.0, mone = -1.0;
  // V(:,0) = vinit / |vinit|_B, Z(:,0) = B V(:,0)
  PV_HIP(hipMemcpyAsync(vcol(0), vinit->data, sizeof(double) * n, hipMemcpyHostToDevice, st));
  apply_B(c, pevsl, vcol(0), zcol(0));
  double t = 1.0 / std::sqrt(dot(vcol(0), zcol(0)));
  scale_vec(n, zcol(0), nullptr, t, st);
  scale_vec(n, vcol(0), nullptr, t, st);
  // a random vector B-orthogonalized against V(:,0:k-1) into (V, Z)(:,k),
  // unit B-norm (lucky breakdown)
  auto restart_random = [&](int k) {
    rand_vectors(c, n, 1, vcol(k), ld);
    cgs2(c, n, k, NGS_MAX, Z, V, ld, vcol(k));
    apply_B(c, pevsl, vcol(k), zcol(k));
    const double ib = 1.0 / std::sqrt(dot(vcol(k), zcol(k)));
    scale_vec(n, vcol(k), nullptr, ib, st);
    scale_vec(n, zcol(k), nullptr, ib, st);
  };
  std::vector<double> T((size_t)lanm1 * lanm1, 0.0), Rval((size_t)lanm1, 0.0), Tk, Ev, w;
  auto Tat = [&](int r, int cc) -> double & { return T[(size_t)cc * lanm1 + r]; };
  double s[3] = {0.0, 0.0, 0.0}, lmin = 0.0, lmax = 0.0, t1 = 0.0, t2 = 0.0;
  int it = 0, trlen = 0, rc = 0;
  while (it < maxit) {
    double wn = 0.0, beta = 0.0;
    int nwn = 0;
    int k = trlen > 0 ? trlen + 1 : 0;
    if (k + min_inner_step > lanm1) {
      fprintf(stderr, "Krylov dim too small for this problem. Try a larger dim\n");
      rc = -1;
      break;
    }
    if (trlen > 0) {
      const int k1 = k - 1;
      double *znew = zcol(k1 + 1), *vnew = vcol(k1 + 1);
      apply_A(c, vcol(k1), znew);
      for (int i = 0; i < trlen; i++) {
        Tat(i, i) = Rval[(size_t)i];
        wn += std::fabs(Rval[(size_t)i]);
      }
      s[k1] = dot(vcol(k1), znew);
      // znew -= Z(:,0:k-1) s(0:k-1)
      PV_HIP(hipMemcpyAsync(sv, s, sizeof(double) * k, hipMemcpyHostToDevice, st));
      PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_none, n, k, &mone, Z, (int)ld, sv, 1,
                            &one, znew, 1));
      for (int i = 0; i < k1; i++) {
        Tat(i, trlen) = s[i];
        Tat(trlen, i) = s[i];
        wn += 2.0 * std::fabs(s[i]);
      }
      Tat(trlen, trlen) = s[k1];
      wn += std::fabs(s[k1]);
      solve_B(c, pevsl, znew, vnew);
      beta = std::sqrt(dot(vnew, znew));
      wn += 2.0 * beta;
      nwn += 3 * k;
      if (beta * nwn < orthTol * wn) {
        restart_random(k1 + 1);
        beta = 0.0;
      } else {
        scale_vec(n, vnew, nullptr, 1.0 / beta, st);
        scale_vec(n, znew, nullptr, 1.0 / beta, st);
      }
      Tat(k, k1) = beta;
      Tat(k1, k) = beta;
    }
    const double *zold = k > 0 ? zcol(k - 1) : nullptr;
    while (k < lanm && it < maxit) {
      k++;
      double *v = vcol(k - 1), *z = zcol(k - 1), *vnew = vcol(k), *znew = zcol(k);
      apply_A(c, v, znew);
      it++;
      if (zold) axpy_vec(n, -beta, zold, znew, st);
      const double alpha = dot(v, znew);
      Tat(k - 1, k - 1) = alpha;
      wn += std::fabs(alpha);
      axpy_vec(n, -alpha, z, znew, st);
      cgs2(c, n, k, NGS_MAX, V, Z, ld, znew);
      solve_B(c, pevsl, znew, vnew);
      beta = std::sqrt(dot(vnew, znew));
      wn += 2.0 * beta;
      nwn += 3;
      zold = z;
      if (beta * nwn < orthTol * wn) {
        if (fstats) fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", it, beta);
        restart_random(k);
        beta = 0.0;
      } else {
        scale_vec(n, vnew, nullptr, 1.0 / beta, st);
        scale_vec(n, znew, nullptr, 1.0 / beta, st);
      }
      Tat(k, k - 1) = beta;
      Tat(k - 1, k) = beta;
    }
    Tk.assign((size_t)k * k, 0.0);
    for (int cc = 0; cc < k; cc++)
      for (int r = 0; r < k; r++) Tk[(size_t)cc * k + r] = Tat(r, cc);
    jacobi_eig(k, Tk, w, Ev);
    for (int i = 0; i < k; i++) Rval[(size_t)i] = w[(size_t)i];
    s[0] = beta * Ev[(size_t)k - 1];
    s[1] = beta * Ev[(size_t)(k - 1) * k + (k - 1)];
    if (bndtype <= 1) {
      t1 = std::fabs(s[0]);
      t2 = std::fabs(s[1]);
    } else {
      t1 = 2.0 * s[0] * s[0] / (Rval[1] - Rval[0]);
      t2 = 2.0 * s[1] * s[1] / (Rval[(size_t)k - 1] - Rval[(size_t)k - 2]);
    }
    lmin = Rval[0] - t1;
    lmax = Rval[(size_t)k - 1] + t2;
    // extreme Ritz vectors and their B-images become (V, Z)(:, 0:1)
    std::vector<double> e2((size_t)2 * lanm1, 0.0);
    std::copy(Ev.begin(), Ev.begin() + k, e2.begin());
    std::copy(Ev.begin() + (long)(k - 1) * k, Ev.begin() + (long)k * k, e2.begin() + lanm1);
    PV_HIP(hipMemcpyAsync(sv, e2.data(), sizeof(double) * 2 * lanm1, hipMemcpyHostToDevice, st));
    for (int j = 0; j < 2; j++) {
      PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_none, n, k, &one, V, (int)ld,
                            sv + (size_t)j * lanm1, 1, &zero, rv + (size_t)j * ld, 1));
      PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_none, n, k, &one, Z, (int)ld,
                            sv + (size_t)j * lanm1, 1, &zero, rv + (size_t)(2 + j) * ld, 1));
    }
    trlen = 2;
    for (int j = 0; j < 2; j++) {
      PV_HIP(hipMemcpyAsync(vcol(j), rv + (size_t)j * ld, sizeof(double) * ld,
                            hipMemcpyDeviceToDevice, st));
      PV_HIP(hipMemcpyAsync(zcol(j), rv + (size_t)(2 + j) * ld, sizeof(double) * ld,
                            hipMemcpyDeviceToDevice, st));
    }
    Rval[1] = Rval[(size_t)k - 1];
    if (fstats)
      fprintf(fstats, "it %4d, k %3d: ritz %.15e %.15e, t1,t2 %e %e, res %.15e %.15e\n", it, k,
              Rval[0], Rval[1], t1, t2, std::fabs(s[0]), std::fabs(s[1]));
    if (t1 + t2 < tol * (std::fabs(lmin) + std::fabs(lmax))) break;
    std::fill(T.begin(), T.end(), 0.0);
    PV_HIP(hipMemcpyAsync(vcol(trlen), vcol(k), sizeof(double) * ld, hipMemcpyDeviceToDevice,
                          st));
    PV_HIP(hipMemcpyAsync(zcol(trlen), zcol(k), sizeof(double) * ld, hipMemcpyDeviceToDevice,
                          st));
  }
  PV_HIP(hipStreamSynchronize(st));
  *lammin = lmin;
  *lammax = lmax;
  return rc;
}

}  // namespace
}  // namespace pevsl

using namespace pevsl;

extern "C" {

int pEVSL_LanTrbounds(pevsl_Data *pevsl, int lanm, int maxit, double tol, pevsl_Parvec *vinit,
                      int bndtype, double *lammin, double *lammax, FILE *fstats) {
  const double tms = wtime();
  if (!pevsl || !vinit || !lammin || !lammax || lanm < 1) return -1;
  Ctx &c = ctx_of(pevsl);
  const int n = pevsl->n, N = pevsl->N;
  lanm = std::min(lanm, N);
  const int lanm1 = lanm + 1;
  if (lanm == N) maxit = std::min(maxit, N);
  if (fstats) fprintf(fstats, " LanTR for bounds: dim %d, maxits %d\n", lanm, maxit);
  if (pevsl->ifGenEv) {
    const int rc = lan_trbounds_gen(pevsl, c, lanm, maxit, tol, vinit, bndtype, lammin,
                                    lammax, fstats);
    pevsl->stats->t_eigbounds += wtime() - tms;
    return rc;
  }
  const int min_inner_step = 5;
  const hipStream_t st = c.stream;
  // basis: one chunk of >= lanm + 1 columns, rows [n, ld) zero
  LanBatch L;
  L.n = n;
  L.ld = ((size_t)n + 63) / 64 * 64;
  L.nslice = 1;
  int shift = 5;
  while ((1 << shift) < lanm1) shift++;
  L.chunk_shift = shift;
  L.maxchunks = 1;
  const size_t ld = L.ld, ncols = (size_t)1 << shift;
  DBuf<double> V, sc, h, part, rv, ev, pt;
  DBuf<void *> vtab;
  DBuf<int> fl;
  V.ensure(ld * ncols);
  PV_HIP(hipMemsetAsync(V.p, 0, sizeof(double) * ld * ncols, st));
  void *vp = V.p;
  L.vtab = vtab.ensure(1);
  PV_HIP(hipMemcpyAsync(L.vtab, &vp, sizeof(void *), hipMemcpyHostToDevice, st));
  double *s4 = sc.ensure(4);
  L.alpha = s4;
  L.beta = s4 + 1;
  L.nrm2a = s4 + 2;
  L.nrm2b = s4 + 3;
  int *f2 = fl.ensure(2);
  L.pass2 = f2;
  L.active = f2 + 1;
  const int flags[2] = {0, 1};
  PV_HIP(hipMemcpyAsync(f2, flags, sizeof(int) * 2, hipMemcpyHostToDevice, st));
  PV_HIP(hipMemsetAsync(s4, 0, sizeof(double) * 4, st));
  L.hcap = lanm1 + 1;
  L.h = h.ensure((size_t)L.hcap);
  L.part_cap = lan_part_doubles(n, 1, L.hcap);
  L.part = part.ensure(L.part_cap);
  pt.ensure((size_t)dot_part_blocks(n) + 8);
  rv.ensure(2 * ld);
  ev.ensure(2 * (size_t)lanm1);
  PV_HIP(hipMemsetAsync(rv.p, 0, sizeof(double) * 2 * ld, st));
  auto col = [&](int j) { return V.p + (size_t)j * ld; };
  auto dotd = [&](const double *x, const double *y) {
    const int nb = dot_partials(n, x, y, pt.p, st);
    reduce_partials(pt.p, nb, 1, pt.p + nb, st);
    double r = 0.0;
    PV_HIP(hipMemcpyAsync(&r, pt.p + nb, sizeof(double), hipMemcpyDeviceToHost, st));
    PV_HIP(hipStreamSynchronize(st));
    return r;
  };
  auto set_dev = [&](double *dst, double v) {
    PV_HIP(hipMemcpyAsync(dst, &v, sizeof(double), hipMemcpyHostToDevice, st));
    PV_HIP(hipStreamSynchronize(st));
  };
  auto get_dev = [&](const double *src) {
    double v = 0.0;
    PV_HIP(hipMemcpyAsync(&v, src, sizeof(double), hipMemcpyDeviceToHost, st));
    PV_HIP(hipStreamSynchronize(st));
    return v;
  };
  // vnew = random vector orthogonalized against V(:,0:k-1), unit norm,
  // beta = 0 (lucky breakdown, SRC/lantrbnd.c:192-211, 278-301)
  auto restart_random = [&](int k) {
    rand_vectors(c, n, 1, col(k), ld);
    set_dev(L.alpha, 0.0);
    lan_subtract_alpha(L, k - 1, st);  // alpha = 0: only |v|^2
    lan_cgs_pass(L, k - 1, 0, st);
    lan_cgs_pass(L, k - 1, 1, st);
    const double b = std::sqrt(get_dev(L.nrm2b));
    scale_vec(n, col(k), nullptr, 1.0 / b, st);
  };
  // V(:,0) = vinit / |vinit|
  PV_HIP(hipMemcpyAsync(col(0), vinit->data, sizeof(double) * n, hipMemcpyHostToDevice, st));
  scale_vec(n, col(0), nullptr, 1.0 / std::sqrt(dotd(col(0), col(0))), st);

  std::vector<double> T((size_t)lanm1 * lanm1, 0.0), Rval((size_t)lanm1, 0.0), Tk, Ev, w;
  auto Tat = [&](int r, int cc) -> double & { return T[(size_t)cc * lanm1 + r]; };
  double s[3] = {0.0, 0.0, 0.0}, lmin = 0.0, lmax = 0.0, t1 = 0.0, t2 = 0.0;
  int it = 0, trlen = 0, rc = 0;
  while (it < maxit) {
    double wn = 0.0, beta = 0.0;
    int nwn = 0;
    int k = trlen > 0 ? trlen + 1 : 0;
    if (k + min_inner_step > lanm1) {
      fprintf(stderr, "Krylov dim too small for this problem. Try a larger dim\n");
      rc = -1;
      break;
    }
    if (trlen > 0) {
      // thick-restart step (SRC/lantrbnd.c:142-223): arrowhead T
      const int k1 = k - 1;
      apply_A(c, col(k1), col(k1 + 1));
      for (int i = 0; i < trlen; i++) {
        Tat(i, i) = Rval[(size_t)i];
        wn += std::fabs(Rval[(size_t)i]);
      }
      s[k1] = dotd(col(k1), col(k1 + 1));
      for (int j = 0; j < k; j++) axpy_vec(n, -s[j], col(j), col(k1 + 1), st);
      for (int i = 0; i < k1; i++) {
        Tat(i, trlen) = s[i];
        Tat(trlen, i) = s[i];
        wn += 2.0 * std::fabs(s[i]);
      }
      Tat(trlen, trlen) = s[k1];
      wn += std::fabs(s[k1]);
      beta = std::sqrt(dotd(col(k1 + 1), col(k1 + 1)));
      wn += 2.0 * beta;
      nwn += 3 * k;
      if (beta * nwn < orthTol * wn) {
        restart_random(k1 + 1);
        beta = 0.0;
      } else {
        scale_vec(n, col(k1 + 1), nullptr, 1.0 / beta, st);
      }
      Tat(k, k1) = beta;
      Tat(k1, k) = beta;
    }
    // Lanczos steps with full reorthogonalization (SRC/lantrbnd.c:236-314)
    while (k < lanm && it < maxit) {
      k++;
      apply_A(c, col(k - 1), col(k));
      it++;
      set_dev(L.beta, beta);  // znew -= beta zold (k > 1), alpha = <v, znew>
      lan_alpha(L, k - 1, st);
      lan_subtract_alpha(L, k - 1, st);
      lan_cgs_pass(L, k - 1, 0, st);
      lan_cgs_pass(L, k - 1, 1, st);
      double sc2[4];
      PV_HIP(hipMemcpyAsync(sc2, s4, sizeof(double) * 4, hipMemcpyDeviceToHost, st));
      PV_HIP(hipStreamSynchronize(st));
      const double alpha = sc2[0];
      Tat(k - 1, k - 1) = alpha;
      wn += std::fabs(alpha);
      beta = std::sqrt(sc2[3]);
      wn += 2.0 * beta;
      nwn += 3;
      if (beta * nwn < orthTol * wn) {
        if (fstats) fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", it, beta);
        restart_random(k);
        beta = 0.0;
      } else {
        set_dev(L.beta, beta);
        lan_scale_new(L, k - 1, st);
      }
      Tat(k, k - 1) = beta;
      Tat(k - 1, k) = beta;
    }
    // eigenpairs of T(0:k, 0:k) (SRC/lantrbnd.c:315-324)
    Tk.assign((size_t)k * k, 0.0);
    for (int cc = 0; cc < k; cc++)
      for (int r = 0; r < k; r++) Tk[(size_t)cc * k + r] = Tat(r, cc);
    jacobi_eig(k, Tk, w, Ev);
    for (int i = 0; i < k; i++) Rval[(size_t)i] = w[(size_t)i];
    s[0] = beta * Ev[(size_t)k - 1];
    s[1] = beta * Ev[(size_t)(k - 1) * k + (k - 1)];
    if (bndtype <= 1) {
      t1 = std::fabs(s[0]);
      t2 = std::fabs(s[1]);
    } else {
      t1 = 2.0 * s[0] * s[0] / (Rval[1] - Rval[0]);
      t2 = 2.0 * s[1] * s[1] / (Rval[(size_t)k - 1] - Rval[(size_t)k - 2]);
    }
    lmin = Rval[0] - t1;
    lmax = Rval[(size_t)k - 1] + t2;
    // the two extreme Ritz vectors become V(:,0), V(:,1)
    std::vector<double> e2((size_t)2 * lanm1, 0.0);
    std::copy(Ev.begin(), Ev.begin() + k, e2.begin());
    std::copy(Ev.begin() + (long)(k - 1) * k, Ev.begin() + (long)k * k, e2.begin() + lanm1);
    PV_HIP(hipMemcpyAsync(ev.p, e2.data(), sizeof(double) * 2 * lanm1, hipMemcpyHostToDevice, st));
    const double one = 1.0, zero = 0.0;
    PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_none, n, k, &one, V.p, (int)ld, ev.p, 1,
                          &zero, rv.p, 1));
    PV_BLAS(rocblas_dgemv(c.blas, rocblas_operation_none, n, k, &one, V.p, (int)ld,
                          ev.p + lanm1, 1, &zero, rv.p + ld, 1));
    trlen = 2;
    PV_HIP(hipMemcpyAsync(col(0), rv.p, sizeof(double) * ld, hipMemcpyDeviceToDevice, st));
    PV_HIP(hipMemcpyAsync(col(1), rv.p + ld, sizeof(double) * ld, hipMemcpyDeviceToDevice, st));
    Rval[1] = Rval[(size_t)k - 1];
    if (fstats)
      fprintf(fstats, "it %4d, k %3d: ritz %.15e %.15e, t1,t2 %e %e, res %.15e %.15e\n", it, k,
              Rval[0], Rval[1], t1, t2, std::fabs(s[0]), std::fabs(s[1]));
    if (t1 + t2 < tol * (std::fabs(lmin) + std::fabs(lmax))) break;
    std::fill(T.begin(), T.end(), 0.0);
    // V(:,trlen) = V(:,k)
    PV_HIP(hipMemcpyAsync(col(trlen), col(k), sizeof(double) * ld, hipMemcpyDeviceToDevice, st));
  }
  PV_HIP(hipStreamSynchronize(st));
  *lammin = lmin;
  *lammax = lmax;
  pevsl->stats->t_eigbounds += wtime() - tms;
  return rc;
}

}  // extern "C"
