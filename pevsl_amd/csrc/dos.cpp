// dos.cpp -- density of states and spectrum slicing on the GPU.
//
// pEVSL_Kpmdos (SRC/kpmdos.c:36-197): all nvec random vectors advance
// together as the columns of one interleaved block; each degree is one fused
// launch (kernels.hip k_kpm_blk: SpMM + three-term recurrence + per-column
// dot with v0; k_kpm_bal, in blocks of at most 32 vectors, when a row has
// more than 64 entries).  For the generalized problem the moments come from
// the B-solve form (gen.cpp).  The per-(vector, degree) dots come back to the host, where
// mu and the eigenvalue count are accumulated in the reference's order
// (vector-major), so the host arithmetic matches the reference exactly.
//
// pEVSL_SpslicerKpm (SRC/spslicer.c:62-156): the closed-form DOS integral at
// the npts grid points is evaluated on the GPU, the greedy equal-share walk on
// the host.
//
// pEVSL_LanDosG (SRC/landos.c:40-314): nvec independent msteps-step Lanczos
// runs with full reorthogonalization advance together on the batched Lanczos
// kernels; the msteps x msteps tridiagonal problems are solved by rocSOLVER;
// the Gaussian blur and Simpson integration run on the host.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>

#include <rocsolver/rocsolver.h>

#include "internal.hpp"

namespace pevsl {
namespace {

int check_intv(const double *intv, FILE *f) {
  const double a = intv[0], b = intv[1], lmin = intv[2], lmax = intv[3];
  if (a >= b) {
    fprintf(f, " error: invalid interval (%e, %e)\n", a, b);
    return -1;
  }
  if (a >= lmax || b <= lmin) {
    fprintf(f, " error: interval (%e, %e) is outside (%e %e) \n", a, b, lmin, lmax);
    return -2;
  }
  return 0;
}

void jackson_like(int m, int damping, std::vector<double> &jac) {
  jac.assign((size_t)m + 1, 0.0);
  const double dm = m;
  double thetJ = 0, thetL = 0, a1 = 0, a2 = 0;
  if (damping == 1) {
    thetJ = M_PI / (dm + 2);
    a1 = 1 / (dm + 2);
    a2 = sin(thetJ);
  } else if (damping == 2) {
    thetL = M_PI / (dm + 1);
  }
  jac[0] = 0.5;
  for (int k = 1; k <= m; k++) {
    if (damping == 0)
      jac[(size_t)k] = 1.0;
    else if (damping == 1)
      jac[(size_t)k] = a1 * sin((k + 1) * thetJ) / a2 + (1 - (k + 1) * a1) * cos(k * thetJ);
    else
      jac[(size_t)k] = sin(k * thetL) / (k * thetL);
  }
}

// Context for entry points that take no instance (SpslicerKpm): a private
// stream on the current device.
struct Aux {
  hipStream_t s = nullptr;
  Aux() { PV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
};
Aux &aux() {
  static Aux *a = new Aux();  // lives until process exit
  return *a;
}

template <typename T>
void kpm_moments(Ctx &c, int n, int Mdeg, int nvec, double ctr, double wid,
                 std::vector<double> &dots /* [nvec][Mdeg] */) {
  const DevCsr &A = *c.A;
  const hipStream_t s = c.stream;
  // uneven rows: nnz-balanced step, blocks of at most 32 vectors
  const bool bal = A.max_row > kUnevenRow;
  const int bmax = bal ? 32 : 64;
  int B = 1;
  while (B < nvec && B < bmax) B *= 2;
  const size_t blk = (size_t)n * B;
  DBuf<double> raw, part, dd;
  DBuf<T> v0, b0, b1, b2;
  raw.ensure((size_t)n * B);
  v0.ensure(blk);
  b0.ensure(blk);
  b1.ensure(blk);
  b2.ensure(blk);
  int nrb = kpm_part_blocks(n, B);
  if (bal) nrb = std::max(nrb, A.nrowblk_q[bal_level((int)sizeof(T), B)]);
  part.ensure((size_t)nrb * B);
  dd.ensure((size_t)Mdeg * B + (size_t)dot_part_blocks(n) + 8);
  DBuf<double *> cp;
  cp.ensure((size_t)B);
  std::vector<double> hd((size_t)Mdeg * B);
  std::vector<double> nrm((size_t)B);
  for (int v0i = 0; v0i < nvec; v0i += B) {
    const int nb = std::min(B, nvec - v0i);
    // random vectors, unit 2-norm (SRC/kpmdos.c:117-121)
    rand_vectors(c, n, nb, raw.p, (size_t)n);
    for (int j = 0; j < nb; j++) {
      double *col = raw.p + (size_t)j * n;
      double *pt = dd.p + (size_t)Mdeg * B;
      const int np = dot_partials(n, col, col, pt, s);
      reduce_partials(pt, np, 1, dd.p + j, s);
    }
    PV_HIP(hipMemcpyAsync(nrm.data(), dd.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    PV_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nb; j++)
      scale_vec(n, raw.p + (size_t)j * n, nullptr, 1.0 / sqrt(nrm[(size_t)j]), s);
    std::vector<double *> ptrs((size_t)B, nullptr);
    for (int j = 0; j < nb; j++) ptrs[(size_t)j] = raw.p + (size_t)j * n;
    PV_HIP(hipMemcpyAsync(cp.p, ptrs.data(), sizeof(double *) * B, hipMemcpyHostToDevice, s));
    pack_cols<T>(n, B, nb, cp.p, v0.p, s);
    // vk = v0, vkm1 = 0
    PV_HIP(hipMemcpyAsync(b0.p, v0.p, sizeof(T) * blk, hipMemcpyDeviceToDevice, s));
    PV_HIP(hipMemsetAsync(b1.p, 0, sizeof(T) * blk, s));
    T *vk = b0.p, *vkm1 = b1.p, *vkp1 = b2.p;
    for (int k = 0; k < Mdeg; k++) {
      const double scal = (k == 0 ? 1.0 : 2.0) / wid;
      const int nbk = bal ? kpm_step_bal<T>(A, B, vk, vkm1, vkp1, v0.p, ctr, scal, part.p, s)
                          : kpm_step_blk<T>(A, B, vk, vkm1, vkp1, v0.p, ctr, scal, part.p, s);
      reduce_partials(part.p, nbk, B, dd.p + (size_t)k * B, s);
      T *t = vkm1;
      vkm1 = vk;
      vk = vkp1;
      vkp1 = t;
      c.kstats.filter_steps += nb;
    }
    PV_HIP(hipMemcpyAsync(hd.data(), dd.p, sizeof(double) * Mdeg * B, hipMemcpyDeviceToHost, s));
    PV_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nb; j++)
      for (int k = 0; k < Mdeg; k++)
        dots[(size_t)(v0i + j) * Mdeg + k] = hd[(size_t)k * B + j];
  }
}

// KPM moments through an operator given as a host callback (pEVSL_SetAMatvec):
// the reference's per-vector loop (SRC/kpmdos.c:102-157), one product per
// degree through apply_A and the vector part of the recurrence on the device.
void kpm_moments_op(Ctx &c, int n, int Mdeg, int nvec, double ctr, double wid,
                    std::vector<double> &dots) {
  const hipStream_t s = c.stream;
  DBuf<double> buf, dd;
  double *v = buf.ensure((size_t)4 * n), *vk = v + n, *vkm1 = v + 2 * (size_t)n,
         *vkp1 = v + 3 * (size_t)n;
  const int npb = dot_part_blocks(n);
  double *pt = dd.ensure((size_t)Mdeg + npb + 8);
  std::vector<double> hd((size_t)Mdeg + 1);
  for (int m = 0; m < nvec; m++) {
    rand_vectors(c, n, 1, v, (size_t)n);
    const int nb0 = dot_partials(n, v, v, pt + Mdeg + 1, s);
    reduce_partials(pt + Mdeg + 1, nb0, 1, pt + Mdeg, s);
    PV_HIP(hipMemcpyAsync(hd.data() + Mdeg, pt + Mdeg, sizeof(double), hipMemcpyDeviceToHost,
                          s));
    PV_HIP(hipStreamSynchronize(s));
    scale_vec(n, v, nullptr, 1.0 / sqrt(hd[(size_t)Mdeg]), s);
    copy_f64((size_t)n, v, vk, s);
    fill_zero_f64((size_t)n, vkm1, s);
    for (int k = 0; k < Mdeg; k++) {
      apply_A(c, vk, vkp1);
      const double scal = (k == 0 ? 1.0 : 2.0) / wid;
      cheb_update(n, vk, vkm1, vkp1, nullptr, ctr, scal, 0.0, s);
      double *t = vkm1;
      vkm1 = vk;
      vk = vkp1;
      vkp1 = t;
      const int nb = dot_partials(n, vk, v, pt + Mdeg + 1, s);
      reduce_partials(pt + Mdeg + 1, nb, 1, pt + k, s);
      c.kstats.filter_steps++;
    }
    PV_HIP(hipMemcpyAsync(hd.data(), pt, sizeof(double) * Mdeg, hipMemcpyDeviceToHost, s));
    PV_HIP(hipStreamSynchronize(s));
    std::copy(hd.begin(), hd.begin() + Mdeg, dots.begin() + (long)m * Mdeg);
  }
}

}  // namespace
}  // namespace pevsl

using namespace pevsl;

extern "C" {

int pEVSL_Kpmdos(pevsl_Data *pevsl, int Mdeg, int damping, int nvec, double *intv,
                 int ngroups, int groupid, MPI_Comm gl_comm, double *mu, double *ecnt) {
  (void)groupid;
  (void)gl_comm;
  const double tm = wtime();
  Ctx &c = ctx_of(pevsl);
  if (ngroups > 1) return PEVSL_NOT_IMPLEMENT;  // one process: one group
  if (check_intv(intv, stdout) < 0) return -1;
  if (!c.A && !c.host_mv && !pevsl->ifGenEv) return -3;  // no operator installed
  const int n = pevsl->n;
  const double aa = std::max(intv[0], intv[2]), bb = std::min(intv[1], intv[3]);
  if (intv[0] < intv[2] || intv[1] > intv[3])
    fprintf(stdout, " warning: interval (%e, %e) is adjusted to (%e, %e)\n", intv[0],
            intv[1], aa, bb);
  const double ctr = (intv[3] + intv[2]) / 2.0, wid = (intv[3] - intv[2]) / 2.0;
  double t = std::max(-1.0 + DBL_EPSILON, (aa - ctr) / wid);
  const double beta1 = acos(t);
  t = std::min(1.0 - DBL_EPSILON, (bb - ctr) / wid);
  const double beta2 = acos(t);
  std::vector<double> jac;
  jackson_like(Mdeg, damping, jac);
  jac[0] = 1.0;
  std::vector<double> dots((size_t)nvec * Mdeg);
  EvTimer ev;
  ev.start(c.stream);
  if (pevsl->ifGenEv) {
    kpm_moments_gen(c, pevsl, Mdeg, nvec, ctr, wid, dots);
  } else if (!c.A) {
    // host-callback operator: FP64 vectors whatever the filter precision
    if ((int)c.h_x.size() != n) {
      c.h_x.assign((size_t)n, 0.0);
      c.h_y.assign((size_t)n, 0.0);
    }
    kpm_moments_op(c, n, Mdeg, nvec, ctr, wid, dots);
  } else if (c.filter_prec == PEVSL_PREC_FP32)
    kpm_moments<float>(c, n, Mdeg, nvec, ctr, wid, dots);
  else
    kpm_moments<double>(c, n, Mdeg, nvec, ctr, wid, dots);
  ev.stop(c.stream);
  c.kstats.t_kpm += ev.seconds();
  if (c.A && !pevsl->ifGenEv) pevsl->stats->n_mvA += (size_t)nvec * Mdeg;
  // accumulation in the reference order (SRC/kpmdos.c:104-157)
  memset(mu, 0, sizeof(double) * (size_t)(Mdeg + 1));
  double tcnt = 0.0;
  for (int m = 0; m < nvec; m++) {
    mu[0] += jac[0];
    tcnt -= jac[0] * (beta2 - beta1);
    for (int k = 0; k < Mdeg; k++) {
      const int k1 = k + 1;
      t = dots[(size_t)m * Mdeg + k] * 2.0 * jac[(size_t)k1];
      mu[k1] += t;
      tcnt -= t * (sin(k1 * beta2) - sin(k1 * beta1)) / k1;
    }
  }
  t = 1.0 / (((double)nvec) * M_PI);
  for (int k = 0; k <= Mdeg; k++) mu[k] *= t;
  tcnt *= t * ((double)pevsl->N);
  *ecnt = tcnt;
  pevsl->stats->t_dos = wtime() - tm;
  return 0;
}

int pEVSL_SpslicerKpm(double *sli, double *mu, int Mdeg, double *intv, int n_int, int npts) {
  if (check_intv(intv, stdout) < 0) return -1;
  const double aa = std::max(intv[0], intv[2]), bb = std::min(intv[1], intv[3]);
  if (intv[0] < intv[2] || intv[1] > intv[3])
    fprintf(stdout, " warning:  interval (%e, %e) is adjusted to (%e, %e)\n", intv[0],
            intv[1], aa, bb);
  memset(sli, 0, sizeof(double) * (size_t)(n_int + 1));
  if (n_int == 1) {
    sli[0] = intv[0];
    sli[1] = intv[1];
    return 0;
  }
  const double ctr = (intv[3] + intv[2]) / 2, wid = (intv[3] - intv[2]) / 2;
  double aL = (aa - ctr) / wid, bL = (bb - ctr) / wid;
  aL = std::max(aL, -1.0);
  bL = std::min(bL, 1.0);
  npts = std::max(npts, 2 * n_int + 1);
  std::vector<double> xi((size_t)npts), yi((size_t)npts), th((size_t)npts);
  linspace(aL, bL, npts, xi.data());
  for (int j = 0; j < npts; j++) th[(size_t)j] = acos(xi[(size_t)j]);
  // integrals at all grid points on the device
  const hipStream_t s = aux().s;
  DBuf<double> d;
  d.ensure((size_t)2 * npts + Mdeg + 1);
  PV_HIP(hipMemcpyAsync(d.p, th.data(), sizeof(double) * npts, hipMemcpyHostToDevice, s));
  PV_HIP(hipMemcpyAsync(d.p + 2 * (size_t)npts, mu, sizeof(double) * (Mdeg + 1),
                        hipMemcpyHostToDevice, s));
  kpm_intchx(Mdeg, d.p + 2 * (size_t)npts, npts, d.p, th[0], d.p + npts, s);
  PV_HIP(hipMemcpyAsync(yi.data(), d.p + npts, sizeof(double) * npts, hipMemcpyDeviceToHost,
                        s));
  PV_HIP(hipStreamSynchronize(s));
  // greedy equal-share walk (SRC/spslicer.c:101-122)
  double target = yi[(size_t)npts - 1] / (double)n_int;
  int ls = 0, ii = 0, err = 0;
  sli[0] = intv[0];
  while (++ls < n_int) {
    while (ii < npts && yi[(size_t)ii] < target) ii++;
    if (ii == npts) break;
    // best of the two grid points around the target; at ii = 0 (target <= 0:
    // the DOS integral over the interval is not positive) the first point,
    // where SRC/spslicer.c:116-118 would read yi[-1]
    if (ii > 0 && (target - yi[(size_t)ii - 1]) < (yi[(size_t)ii] - target)) ii--;
    sli[ls] = ctr + wid * xi[(size_t)ii];
    target = yi[(size_t)ii] + (yi[(size_t)npts - 1] - yi[(size_t)ii]) / (n_int - ls);
  }
  sli[n_int] = intv[1];
  if (ls != n_int) err = 1;
  for (int i = 1; i <= n_int; i++)
    if (sli[i] < sli[i - 1]) {
      err += 2;
      break;
    }
  return err;
}

// Composite Simpson (odd points by the trapezoid), in place
// (SRC/simpson.c:20-45).
static void simpson_inplace(const double *xi, double *yi, int npts) {
  double tc = 0.0, ysav = yi[0];
  int i = 1;
  yi[0] = 0;
  while (i < npts - 1) {
    const double tm = tc + (xi[i] - xi[i - 1]) * (ysav + yi[i]) / 2.0;
    tc += (xi[i + 1] - xi[i - 1]) * (ysav + 4 * yi[i] + yi[i + 1]) / 6.0;
    yi[i] = tm;
    ysav = yi[i + 1];
    yi[i + 1] = tc;
    i += 2;
  }
  if (i == npts - 1) yi[i] = tc + (xi[i] - xi[i - 1]) * (yi[i] + ysav) / 2.0;
}

int pEVSL_SymmTridEigvals(int n, const double *diag, const double *sdiag, double thr,
                          double *eigval, int *count, double *trace) {
  if (n <= 0) return -1;
  const hipStream_t s = aux().s;
  DBuf<double> d;
  double *dd = d.ensure((size_t)3 * n + 16);
  double *de = dd + n, *lam = de + n, *out = lam + n;
  PV_HIP(hipMemcpyAsync(dd, diag, sizeof(double) * n, hipMemcpyHostToDevice, s));
  if (n > 1)
    PV_HIP(hipMemcpyAsync(de, sdiag, sizeof(double) * (n - 1), hipMemcpyHostToDevice, s));
  tridiag_trace_above(n, dd, de, thr, lam, out, s);
  double o[6];
  PV_HIP(hipMemcpyAsync(o, out, sizeof(double) * 6, hipMemcpyDeviceToHost, s));
  PV_HIP(hipStreamSynchronize(s));
  const int first = (int)o[5];
  *count = n - first;
  *trace = o[0];
  PV_HIP(hipMemcpyAsync(eigval, lam + first, sizeof(double) * (n - first),
                        hipMemcpyDeviceToHost, s));
  PV_HIP(hipStreamSynchronize(s));
  return 0;
}

void pEVSL_SpslicerLan(double *xi, double *yi, int n_int, int npts, double *sli) {
  simpson_inplace(xi, yi, npts);
  double t = yi[0];
  const double want = (yi[npts - 1] - yi[0]) / (double)n_int;
  int ls = 0;
  sli[0] = xi[0];
  for (int k = 1; k < npts; k++) {
    if (yi[k] - t >= want) {
      ls++;
      if (ls <= n_int) sli[ls] = xi[k];
      t = yi[k];
    }
  }
  sli[n_int] = xi[npts - 1];
}

int pEVSL_LanDosG(pevsl_Data *pevsl, const int nvec, int msteps, int npts, double *xdos,
                  double *ydos, double *neig, double *intv, int ngroups, int groupid,
                  MPI_Comm gl_comm) {
  (void)groupid;
  (void)gl_comm;
  if (ngroups > 1) return PEVSL_NOT_IMPLEMENT;
  Ctx &c = ctx_of(pevsl);
  const double tm = wtime();
  const int n = pevsl->n;
  const int maxit = std::min(pevsl->N, msteps);
  const hipStream_t s = c.stream;
  std::vector<double> dT((size_t)nvec * maxit), eT((size_t)nvec * maxit);
  if (pevsl->ifGenEv) {
    // generalized problem: one random vector after the other (SRC/landos.c
    // ifGenEv branches)
    landos_gen(c, pevsl, nvec, maxit, dT, eT);
  } else {
  // bases: one per random vector, maxit+1 columns each (single chunk)
  LanBatch L;
  L.n = n;
  L.ld = ((size_t)n + 63) / 64 * 64;
  L.nslice = nvec;
  int shift = 0;
  while ((1 << shift) < maxit + 1 || shift < 5) shift++;
  L.chunk_shift = shift;
  L.maxchunks = 1;
  DBuf<double> V, sc, h, part, raw;
  DBuf<double *> vtab;
  DBuf<int> fl;
  V.ensure(L.ld * ((size_t)1 << shift) * nvec);
  PV_HIP(hipMemsetAsync(V.p, 0, sizeof(double) * L.ld * ((size_t)1 << shift) * nvec, s));
  std::vector<double *> vt((size_t)nvec);
  for (int v = 0; v < nvec; v++) vt[(size_t)v] = V.p + L.ld * ((size_t)1 << shift) * v;
  L.vtab = reinterpret_cast<void **>(vtab.ensure((size_t)nvec));
  PV_HIP(hipMemcpyAsync(L.vtab, vt.data(), sizeof(double *) * nvec, hipMemcpyHostToDevice, s));
  double *scp = sc.ensure((size_t)4 * nvec);
  L.alpha = scp;
  L.beta = scp + nvec;
  L.nrm2a = scp + 2 * nvec;
  L.nrm2b = scp + 3 * nvec;
  int *flp = fl.ensure((size_t)2 * nvec);
  L.pass2 = flp;
  L.active = flp + nvec;
  PV_HIP(hipMemsetAsync(scp, 0, sizeof(double) * 4 * nvec, s));
  L.hcap = maxit + 1;
  L.h = h.ensure((size_t)nvec * L.hcap);
  L.part_cap = lan_part_doubles(n, nvec, L.hcap);
  L.part = part.ensure(L.part_cap);
  std::vector<int> ones((size_t)nvec, 1);
  PV_HIP(hipMemcpyAsync(L.active, ones.data(), sizeof(int) * nvec, hipMemcpyHostToDevice, s));
  // start vectors v / |v| (SRC/landos.c:133-159)
  for (int v = 0; v < nvec; v++) rand_vectors(c, n, 1, vt[(size_t)v], L.ld);
  {
    DBuf<double> pt;
    pt.ensure((size_t)dot_part_blocks(n) + 8);
    std::vector<double> nr(1);
    for (int v = 0; v < nvec; v++) {
      const int nb = dot_partials(n, vt[(size_t)v], vt[(size_t)v], pt.p, s);
      reduce_partials(pt.p, nb, 1, pt.p + nb, s);
      PV_HIP(hipMemcpyAsync(nr.data(), pt.p + nb, sizeof(double), hipMemcpyDeviceToHost, s));
      PV_HIP(hipStreamSynchronize(s));
      scale_vec(n, vt[(size_t)v], nullptr, 1.0 / sqrt(nr[0]), s);
    }
  }
  std::vector<double> wn((size_t)nvec, 0.0);
  std::vector<int> nwn((size_t)nvec, 0);
  std::vector<double> ha((size_t)2 * nvec), betas((size_t)nvec);
  const int lanes = c.lanes ? c.lanes : auto_lanes((double)c.A->nnz / std::max(1, n));
  const int msk = (1 << shift) - 1;
  (void)msk;
  for (int k = 0; k < maxit; k++) {
    for (int v = 0; v < nvec; v++) {
      double *base = vt[(size_t)v];
      apply_A(c, base + (size_t)k * L.ld, base + (size_t)(k + 1) * L.ld);
    }
    (void)lanes;
    lan_alpha(L, k, s);
    lan_subtract_alpha(L, k, s);
    lan_cgs_pass(L, k, 0, s);
    lan_cgs_pass(L, k, 1, s);
    PV_HIP(hipMemcpyAsync(ha.data(), L.alpha, sizeof(double) * nvec, hipMemcpyDeviceToHost, s));
    PV_HIP(hipMemcpyAsync(ha.data() + nvec, L.nrm2b, sizeof(double) * nvec,
                          hipMemcpyDeviceToHost, s));
    PV_HIP(hipStreamSynchronize(s));
    std::vector<int> brk;
    for (int v = 0; v < nvec; v++) {
      const double alpha = ha[(size_t)v], beta = sqrt(ha[(size_t)nvec + v]);
      dT[(size_t)v * maxit + k] = alpha;
      wn[(size_t)v] += fabs(alpha) + 2.0 * beta;
      nwn[(size_t)v] += 3;
      betas[(size_t)v] = beta;
      if (beta * nwn[(size_t)v] < orthTol * wn[(size_t)v]) brk.push_back(v);
    }
    PV_HIP(hipMemcpyAsync(L.beta, betas.data(), sizeof(double) * nvec, hipMemcpyHostToDevice, s));
    lan_scale_new(L, k, s);  // breakdown columns are overwritten below
    for (int v : brk) {
      // fresh random vector orthogonalized against the basis (SRC/landos.c:209-232)
      double *col = vt[(size_t)v] + (size_t)(k + 1) * L.ld;
      rand_vectors(c, n, 1, col, L.ld);
      std::vector<int> one((size_t)nvec, 0);
      one[(size_t)v] = 1;
      PV_HIP(hipMemcpyAsync(L.active, one.data(), sizeof(int) * nvec, hipMemcpyHostToDevice, s));
      PV_HIP(hipMemsetAsync(L.alpha + v, 0, sizeof(double), s));
      lan_subtract_alpha(L, k, s);
      lan_cgs_pass(L, k, 0, s);
      lan_cgs_pass(L, k, 1, s);
      double nb2 = 0;
      PV_HIP(hipMemcpyAsync(&nb2, L.nrm2b + v, sizeof(double), hipMemcpyDeviceToHost, s));
      PV_HIP(hipStreamSynchronize(s));
      const double b = sqrt(nb2);
      PV_HIP(hipMemcpyAsync(L.beta + v, &b, sizeof(double), hipMemcpyHostToDevice, s));
      lan_scale_new(L, k, s);
      const double zero = 0.0;
      PV_HIP(hipMemcpyAsync(L.beta + v, &zero, sizeof(double), hipMemcpyHostToDevice, s));
      PV_HIP(hipMemcpyAsync(L.active, ones.data(), sizeof(int) * nvec, hipMemcpyHostToDevice, s));
      PV_HIP(hipStreamSynchronize(s));
      betas[(size_t)v] = 0.0;
    }
    for (int v = 0; v < nvec; v++) eT[(size_t)v * maxit + k] = betas[(size_t)v];
  }
  }
  // Ritz values and first components of the eigenvectors of each T
  DBuf<double> D, E, C;
  DBuf<int> info;
  D.ensure((size_t)maxit);
  E.ensure((size_t)maxit);
  C.ensure((size_t)maxit * maxit);
  info.ensure(1);
  const double lm = intv[2], lM = intv[3];
  const double aa = std::max(intv[0], intv[2]), bb = std::min(intv[1], intv[3]);
  const double kappa = 1.25;
  const int M = std::min(msteps, 30);
  const double H = (lM - lm) / (M - 1);
  const double sigma = H / sqrt(8 * log(kappa)) * pevsl->sigma_mult;
  const double sigma2 = 2 * sigma * sigma;
  const double width = sigma * sqrt(-2.0 * log(1e-08));
  linspace(aa, bb, npts, xdos);
  std::vector<double> y((size_t)npts, 0.0), ev((size_t)maxit), Ch((size_t)maxit * maxit);
  for (int v = 0; v < nvec; v++) {
    PV_HIP(hipMemcpyAsync(D.p, dT.data() + (size_t)v * maxit, sizeof(double) * maxit,
                          hipMemcpyHostToDevice, s));
    PV_HIP(hipMemcpyAsync(E.p, eT.data() + (size_t)v * maxit, sizeof(double) * maxit,
                          hipMemcpyHostToDevice, s));
    PV_BLAS(rocsolver_dstedc(c.blas, rocblas_evect_tridiagonal, maxit, D.p, E.p, C.p, maxit,
                             info.p));
    PV_HIP(hipMemcpyAsync(ev.data(), D.p, sizeof(double) * maxit, hipMemcpyDeviceToHost, s));
    PV_HIP(hipMemcpyAsync(Ch.data(), C.p, sizeof(double) * maxit * maxit,
                          hipMemcpyDeviceToHost, s));
    PV_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < maxit; i++) {
      const double g2 = Ch[(size_t)i * maxit] * Ch[(size_t)i * maxit];
      const double th = ev[(size_t)i];
      for (int j = 0; j < npts; j++) {
        if (fabs(xdos[j] - th) < width) {
          const double dx = xdos[j] - th;
          y[(size_t)j] += g2 * exp(-(dx * dx) / sigma2);
        }
      }
    }
  }
  const double scaling = 1.0 / (nvec * sqrt(sigma2 * M_PI));
  for (int j = 0; j < npts; j++) {
    y[(size_t)j] *= scaling;
    ydos[j] = y[(size_t)j];
  }
  simpson_inplace(xdos, y.data(), npts);
  *neig = y[(size_t)npts - 1] * pevsl->N;
  pevsl->stats->t_dos = wtime() - tm;
  return 0;
}

}  // extern "C"
