// tune.cpp -- autotuner of the fused filter step and reorthogonalization
// timing (pevsl_amd.h: pEVSL_Autotune, pEVSL_BenchReorth).  No counterpart
// in the reference: it picks the lanes per row (or the CSR-stream kernel) of
// the one-column filter, and the block width and block kernel (lane per
// (row, column), row per thread, Clenshaw, nnz-balanced) of the batched SpMM,
// for the matrix at hand, from HIP-event timings of back-to-back launches on
// the instance's stream.  pEVSL_BenchBlock, pEVSL_BenchCopy and
// pEVSL_BenchReorth time single kernels for profiling.
#include <algorithm>
#include <cmath>

#include "internal.hpp"

using namespace pevsl;

namespace {

// idx: index bytes per nonzero the kernel reads (2 for the row-block kernels
// on a matrix with 16-bit column offsets)
double step_bytes(const DevCsr &A, int elem, int ncol, int idx = 4) {
  return (double)A.nnz * (elem + idx) + 4.0 * (A.n + 1) + 5.0 * (double)A.n * elem * ncol;
}

// lanes == 0: the CSR-stream kernel
template <typename T>
double time_vec(Ctx &c, int lanes, int reps) {
  const DevCsr &A = *c.A;
  const int n = A.n;
  DBuf<T> b;
  b.ensure((size_t)4 * n);
  T *v0 = b.p, *v1 = b.p + n, *v2 = b.p + 2 * (size_t)n, *y = b.p + 3 * (size_t)n;
  PV_HIP(hipMemsetAsync(b.p, 0, sizeof(T) * 4 * (size_t)n, c.stream));
  // warm-up
  auto step = [&](const T *vk, const T *vkm1, T *vkp1) {
    if (lanes == 0)
      cheb_step_stream<T>(A, vk, vkm1, vkp1, y, 6.0, 1.0 / 6, 0.1, 0.1, false, c.stream);
    else
      cheb_step_vec<T>(A, vk, vkm1, vkp1, y, 6.0, 1.0 / 6, 0.1, 0.1, false, lanes, c.stream);
  };
  step(v0, v1, v2);
  EvTimer ev;
  ev.start(c.stream);
  T *vk = v0, *vkm1 = v1, *vkp1 = v2;
  for (int r = 0; r < reps; r++) {
    step(vk, vkm1, vkp1);
    T *t = vkm1;
    vkm1 = vk;
    vk = vkp1;
    vkp1 = t;
  }
  ev.stop(c.stream);
  return ev.seconds() / reps;
}

// kind: 0 lane per (row, column), 1 row per thread, 2 row per thread in
// Clenshaw form, 3 / 4 the nnz-balanced forms of 1 / 2, 5 the segmented
// Clenshaw form (B a multiple of 16 bytes of columns, at most kMaxSegs pieces)
template <typename T>
double time_blk(Ctx &c, int B, int reps, DBuf<double> &mus, int kind) {
  const DevCsr &A = *c.A;
  const size_t blk = (size_t)A.n * B;
  DBuf<T> b;
  b.ensure(4 * blk);
  PV_HIP(hipMemsetAsync(b.p, 0, sizeof(T) * 4 * blk, c.stream));
  T *v0 = b.p, *v1 = b.p + blk, *v2 = b.p + 2 * blk, *y = b.p + 3 * blk;
  SegTops tops{};
  for (int p = 0; p < kMaxSegs; p++) tops.top[p] = 100;  // every piece forms the product
  const int nseg = B * (int)sizeof(T) / 16;
  // the three-term row kernels alternate skip and pair steps, as a batched
  // solve runs them (on average four vector passes per degree)
  int parity = 0;
  auto step = [&](const T *vk, const T *vkm1, T *vkp1) {
    const int ymode = (parity++ & 1) ? kYPair : kYSkip;
    if (kind == 5)
      clen_step_seg<T>(A, B, vk, vkm1, y, vkp1, 6.0, 1.0 / 3, mus.p, 2, nseg, tops, c.stream);
    else if (kind == 4)
      clen_step_bal<T>(A, B, vk, vkm1, y, vkp1, 6.0, 1.0 / 3, mus.p, 2, c.stream);
    else if (kind == 3)
      cheb_step_bal<T>(A, B, vk, vkm1, vkp1, y, 6.0, 1.0 / 6, mus.p, 2, B, c.stream, ymode);
    else if (kind == 2)
      clen_step_blkrow<T>(A, B, vk, vkm1, y, vkp1, 6.0, 1.0 / 3, mus.p, 2, c.stream);
    else if (kind == 1)
      cheb_step_blkrow<T>(A, B, vk, vkm1, vkp1, y, 6.0, 1.0 / 6, mus.p, 2, B, c.stream, ymode);
    else
      cheb_step_blk<T>(A, B, vk, vkm1, vkp1, y, 6.0, 1.0 / 6, mus.p, 2, B, c.stream);
  };
  step(v0, v1, v2);
  EvTimer ev;
  ev.start(c.stream);
  T *vk = v0, *vkm1 = v1, *vkp1 = v2;
  for (int r = 0; r < reps; r++) {
    step(vk, vkm1, vkp1);
    T *t = vkm1;
    vkm1 = vk;
    vk = vkp1;
    vkp1 = t;
  }
  ev.stop(c.stream);
  return ev.seconds() / reps;
}

}  // namespace

extern "C" {

int pEVSL_Autotune(pevsl_Data *p, int reps, pevsl_TuneResult *res) {
  Ctx &c = ctx_of(p);
  if (!c.A) return -1;
  if (reps < 1) reps = 1;
  const bool f32 = c.filter_prec == PEVSL_PREC_FP32;
  const int elem = f32 ? 4 : 8;
  *res = pevsl_TuneResult{};
  double best = 1e30;
  int bl = 1;
  for (int i = 0; i < 6; i++) {
    const int L = 1 << i;
    const double t = f32 ? time_vec<float>(c, L, reps) : time_vec<double>(c, L, reps);
    res->t_lanes[i] = t;
    if (t < best) {
      best = t;
      bl = L;
    }
  }
  res->bytes_lanes = step_bytes(*c.A, elem, 1);
  res->t_stream = f32 ? time_vec<float>(c, 0, reps) : time_vec<double>(c, 0, reps);
  DBuf<double> mus;
  mus.ensure(3 * 64);
  std::vector<double> mh(3 * 64, 0.01);
  PV_HIP(hipMemcpy(mus.p, mh.data(), sizeof(double) * mh.size(), hipMemcpyHostToDevice));
  double bestc = 1e30;
  int bw = 1, rowk = 1;
  for (int i = 0; i < 6; i++) {
    const int B = 2 << i;
    // skip widths whose blocks would not fit comfortably (4 blocks)
    const double need = 4.0 * c.A->n * (double)B * elem;
    if (need > 32e9) break;
    const double t = f32 ? time_blk<float>(c, B, reps, mus, 0)
                         : time_blk<double>(c, B, reps, mus, 0);
    res->t_block[i] = t;
    res->bytes_block[i] = step_bytes(*c.A, elem, B);
    if (t / B < bestc) {
      bestc = t / B;
      bw = B;
      rowk = 0;
    }
    // segmented Clenshaw kernel (FP32): one pass over the matrix for all
    // B columns, so its cost per column falls with B
    if (f32 && B % 4 == 0) {
      const double ts = time_blk<float>(c, B, reps, mus, 5);
      res->t_blockseg[i] = ts;
      if (ts / B < bestc) {
        bestc = ts / B;
        bw = B;
        rowk = 5;
      }
    }
    if (B <= 32) {
      const double tr = f32 ? time_blk<float>(c, B, reps, mus, 1)
                            : time_blk<double>(c, B, reps, mus, 1);
      res->t_blockrow[i] = tr;
      if (tr / B < bestc) {
        bestc = tr / B;
        bw = B;
        rowk = 1;
      }
      const double tb = f32 ? time_blk<float>(c, B, reps, mus, 3)
                            : time_blk<double>(c, B, reps, mus, 3);
      res->t_blockbal[i] = tb;
      if (tb / B < bestc) {
        bestc = tb / B;
        bw = B;
        rowk = 3;
      }
      // Clenshaw forms: FP32 only (the FP64 path keeps the reference's
      // recurrence)
      if (f32) {
        const double tc = time_blk<float>(c, B, reps, mus, 2);
        res->t_blockclen[i] = tc;
        res->bytes_clen[i] =
            step_bytes(*c.A, elem, B, index_bytes(*c.A)) - (double)c.A->n * B * elem;
        if (tc / B < bestc) {
          bestc = tc / B;
          bw = B;
          rowk = 2;
        }
        const double tcb = time_blk<float>(c, B, reps, mus, 4);
        res->t_blockbalclen[i] = tcb;
        if (tcb / B < bestc) {
          bestc = tcb / B;
          bw = B;
          rowk = 4;
        }
      }
    }
  }
  res->lanes = res->t_stream < best ? PEVSL_SPMV_STREAM : bl;
  res->block_width = bw;
  c.use_stream = res->t_stream < best ? 1 : 0;
  c.lanes = bl;
  c.block_width = bw;
  c.blk_rowkernel = rowk;
  res->block_rowkernel = rowk;
  return 0;
}

int pEVSL_BenchBlock(pevsl_Data *p, int B, int reps, double *sec, double *bytes) {
  Ctx &c = ctx_of(p);
  if (!c.A || reps < 1 || !sec || !bytes || B < 1 || B > 32 || (B & (B - 1))) return -1;
  const bool f32 = c.filter_prec == PEVSL_PREC_FP32;
  const int elem = f32 ? 4 : 8;
  // the kernel a batched solve of width B uses (see BatchSolver::filter_block)
  int kind = c.blk_rowkernel;
  if (kind < 0) kind = (f32 ? 2 : 1) + (c.A->max_row > kUnevenRow ? 2 : 0);
  if (!f32 && (kind == 2 || kind == 4 || kind == 5)) return -1;  // Clenshaw forms: FP32 only
  if (kind == 5 && B < 4) return -1;  // a whole 16-byte piece of columns
  DBuf<double> mus;
  mus.ensure(3 * 64);
  std::vector<double> mh(3 * 64, 0.01);
  PV_HIP(hipMemcpy(mus.p, mh.data(), sizeof(double) * mh.size(), hipMemcpyHostToDevice));
  *sec = f32 ? time_blk<float>(c, B, reps, mus, kind) : time_blk<double>(c, B, reps, mus, kind);
  const int idx = (kind == 1 || kind == 2) ? index_bytes(*c.A) : 4;
  // four vector passes per degree for all but the lane-per-(row, column)
  // kernel (Clenshaw; three-term row kernels with y every other degree)
  *bytes = kind != 0 ? step_bytes(*c.A, elem, B, idx) - (double)c.A->n * B * elem
                     : step_bytes(*c.A, elem, B, idx);
  return 0;
}

int pEVSL_BenchCopy(size_t n, int reps, double *sec, double *bytes) {
  if (n < 2 || reps < 1 || !sec || !bytes) return -1;
  n &= ~(size_t)1;
  hipStream_t s;
  PV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  {
    DBuf<double> x, y;
    x.ensure(n);
    y.ensure(n);
    rand_fill(n, x.p, 11, 0, 0, s);
    copy_f64(n, x.p, y.p, s);  // warm-up
    EvTimer ev;
    ev.start(s);
    for (int r = 0; r < reps; r++) copy_f64(n, x.p, y.p, s);
    ev.stop(s);
    *sec = ev.seconds() / reps;
    *bytes = 16.0 * (double)n;
    PV_HIP(hipStreamSynchronize(s));
  }
  PV_HIP(hipStreamDestroy(s));
  return 0;
}

int pEVSL_BenchReorth(pevsl_Data *p, int k, int reps, double *sec, double *bytes) {
  Ctx &c = ctx_of(p);
  const int n = p->n;
  if (n <= 0 || k < 1) return -1;
  LanBatch L;
  L.n = n;
  L.ld = ((size_t)n + 63) / 64 * 64;
  L.nslice = 1;
  int shift = 0;
  while ((1 << shift) < k + 1 || shift < 5) shift++;
  L.chunk_shift = shift;
  L.maxchunks = 1;
  // basis storage as the instance's Lanczos runs use it
  const bool b32 = c.basis_prec == PEVSL_PREC_FP32;
  L.vbytes = b32 ? 4 : 8;
  const size_t cnt = L.ld * ((size_t)1 << shift);
  DBuf<double> V, sc, h, part;
  DBuf<float> V32;
  DBuf<void *> vt;
  DBuf<int> fl;
  V.ensure(cnt);
  rand_fill(cnt, V.p, 7, 0, 0, c.stream);
  void *vp = V.p;
  if (b32) {
    V32.ensure(cnt);
    copy_f64_to_f32(cnt, V.p, V32.p, c.stream);
    PV_HIP(hipStreamSynchronize(c.stream));
    V.release();
    vp = V32.p;
  }
  L.vtab = vt.ensure(1);
  PV_HIP(hipMemcpy(L.vtab, &vp, sizeof(void *), hipMemcpyHostToDevice));
  double *s = sc.ensure(4);
  L.alpha = s;
  L.beta = s + 1;
  L.nrm2a = s + 2;
  L.nrm2b = s + 3;
  int *f = fl.ensure(2);
  L.pass2 = f;
  L.active = f + 1;
  const int one[2] = {0, 1};
  PV_HIP(hipMemcpy(f, one, sizeof(int) * 2, hipMemcpyHostToDevice));
  PV_HIP(hipMemset(s, 0, sizeof(double) * 4));
  L.hcap = k + 1;
  L.h = h.ensure((size_t)L.hcap);
  L.part_cap = lan_part_doubles(n, 1, L.hcap);
  L.part = part.ensure(L.part_cap);
  const int kk = k - 1;  // basis columns 0..k-1, vector in column k
  lan_cgs_pass(L, kk, 0, c.stream);
  EvTimer ev;
  ev.start(c.stream);
  for (int r = 0; r < reps; r++) lan_cgs_pass(L, kk, 0, c.stream);
  ev.stop(c.stream);
  *sec = ev.seconds() / reps;
  *bytes = 2.0 * L.vbytes * (double)n * k + 4.0 * L.vbytes * n;
  return 0;
}

// CGS_DGKS on host vectors (INC/pevsl_int.h:37, SRC/miscla.c:211-240): the
// reorthogonalization kernels of the Lanczos solve (GEMV-T and GEMV-N over a
// device basis, FP64 sums) on Q(:, 0:k-1) and v, with the reference's DGKS
// loop: after each pass the new norm is compared with 1/sqrt(2) of the old
// one on the device, and another pass runs (at most i_max in all) while it
// is not larger.
int pEVSL_CGS_DGKS(pevsl_Data *p, int k, int i_max, pevsl_Parvecs *Q, pevsl_Parvec *v,
                   double *nrmv, double *w) {
  (void)w;
  Ctx &c = ctx_of(p);
  const int n = p->n;
  if (k < 0 || i_max < 1 || !v || v->n_local != n || (k > 0 && (!Q || Q->n_local != n ||
                                                                   Q->n_vecs < k)))
    return -4;
  const double t0 = wtime();
  LanBatch L;
  L.n = n;
  L.ld = ((size_t)n + 63) / 64 * 64;
  L.nslice = 1;
  int shift = 5;
  while ((1 << shift) < k + 1) shift++;
  L.chunk_shift = shift;
  L.maxchunks = 1;
  L.vbytes = 8;  // the reference's Q and v are FP64
  const size_t cnt = L.ld * ((size_t)1 << shift);
  DBuf<double> V, sc, h, part, pt;
  DBuf<void *> vt;
  DBuf<int> fl;
  V.ensure(cnt);
  fill_zero_f64(cnt, V.p, c.stream);  // rows [n, ld) stay zero
  if (k > 0)
    PV_HIP(hipMemcpy2DAsync(V.p, L.ld * sizeof(double), Q->data, (size_t)Q->ld * sizeof(double),
                            (size_t)n * sizeof(double), (size_t)k, hipMemcpyHostToDevice,
                            c.stream));
  double *z = V.p + (size_t)k * L.ld;
  PV_HIP(hipMemcpyAsync(z, v->data, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
  void *vp = V.p;
  L.vtab = vt.ensure(1);
  PV_HIP(hipMemcpyAsync(L.vtab, &vp, sizeof(void *), hipMemcpyHostToDevice, c.stream));
  double *s = sc.ensure(4);
  L.alpha = s;
  L.beta = s + 1;
  L.nrm2a = s + 2;
  L.nrm2b = s + 3;
  int *f = fl.ensure(2);
  L.pass2 = f;
  L.active = f + 1;
  const int flags[2] = {0, 1};
  PV_HIP(hipMemcpyAsync(f, flags, sizeof(int) * 2, hipMemcpyHostToDevice, c.stream));
  L.hcap = k + 1;
  L.h = h.ensure((size_t)L.hcap);
  L.part_cap = lan_part_doubles(n, 1, L.hcap);
  L.part = part.ensure(L.part_cap);
  // |v|^2 before the first pass
  pt.ensure((size_t)dot_part_blocks(n) + 8);
  const int nb = dot_partials(n, z, z, pt.p, c.stream);
  reduce_partials(pt.p, nb, 1, L.nrm2a, c.stream);
  double nrm2 = 0.0;
  if (k == 0) {
    PV_HIP(hipMemcpyAsync(&nrm2, L.nrm2a, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  }
  for (int it = 0; k > 0 && it < i_max; it++) {
    lan_cgs_pass(L, k - 1, 0, c.stream);  // basis columns 0..k-1, vector in column k
    int again = 0;
    PV_HIP(hipMemcpyAsync(&nrm2, L.nrm2b, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    PV_HIP(hipMemcpyAsync(&again, L.pass2, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PV_HIP(hipStreamSynchronize(c.stream));
    if (!again) break;  // new norm > old / sqrt(2)
    PV_HIP(hipMemcpyAsync(L.nrm2a, L.nrm2b, sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  }
  PV_HIP(hipMemcpyAsync(v->data, z, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
  PV_HIP(hipStreamSynchronize(c.stream));
  if (nrmv) *nrmv = sqrt(nrm2);
  p->stats->t_reorth += wtime() - t0;
  return 0;
}

}  // extern "C"
