// api.cpp -- instance, matrix, vector and utility entry points of the C API.
//
// Instance / matrix setup follows SRC/pevsl.c and SRC/parcsr.c with one
// process: the parallel CSR keeps the whole matrix in `diag`, `offd` has no
// columns and no halo plan is built (SRC/parcsr.c:70-283 collapses).  The
// matrix is copied to HBM once, in ParcsrSetup, and every product runs on
// the GPU.  Vector and multi-vector utilities operate on the host arrays the
// API hands out, as the reference's do (SRC/parvec.c, SRC/parvecs.c).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <ctime>

#include "internal.hpp"

namespace pevsl {

void fatal(const char *what, const char *file, int line) {
  fprintf(stderr, "pevsl_amd fatal error: %s (%s:%d)\n", what, file, line);
  fflush(stderr);
  abort();
}

double wtime() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void check_device() {
  int cnt = 0;
  hipError_t e = hipGetDeviceCount(&cnt);
  if (e != hipSuccess || cnt < 1)
    fatal("no HIP device available: this library runs on the GPU only",
          __FILE__, __LINE__);
}

Ctx &ctx_of(pevsl_Data *p) {
  if (!p || !p->dev) fatal("pevsl instance not started", __FILE__, __LINE__);
  return *static_cast<Ctx *>(p->dev);
}

const DevCsr *dev_matrix_of(pevsl_Parcsr *A) {
  if (!A || !A->dev) return nullptr;
  return &static_cast<DevMatrix *>(A->dev)->csr;
}

KernTimers::~KernTimers() {
  for (Pend &q : pend) {
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

hipEvent_t KernTimers::take() {
  if (pool.empty()) {
    hipEvent_t e;
    PV_HIP(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = pool.back();
  pool.pop_back();
  return e;
}

long KernTimers::begin(hipStream_t s, const char *name) {
  // bound the events in flight: resolve the recorded runs first
  if (pend.size() >= 256) flush();
  size_t i = 0;
  while (i < acc.size() && acc[i].name != name) i++;
  if (i == acc.size()) {
    acc.emplace_back();
    acc.back().name = name;
  }
  Pend q{i, take(), take(), -1, 0.0};
  PV_HIP(hipEventRecord(q.a, s));
  pend.push_back(q);
  return (long)pend.size() - 1;
}

void KernTimers::end(hipStream_t s, long tok, long long launches, double bytes) {
  Pend &q = pend[(size_t)tok];
  q.launches = launches;
  q.bytes = bytes;
  PV_HIP(hipEventRecord(q.b, s));
}

void KernTimers::flush() {
  std::vector<Pend> open;  // begun, end not yet recorded
  for (Pend &q : pend) {
    if (q.launches < 0) {
      open.push_back(q);
      continue;
    }
    PV_HIP(hipEventSynchronize(q.b));
    float ms = 0.f;
    PV_HIP(hipEventElapsedTime(&ms, q.a, q.b));
    Acc &a = acc[q.acc];
    a.launches += q.launches;
    a.sec += 1e-3 * ms;
    a.bytes += q.bytes;
    pool.push_back(q.a);
    pool.push_back(q.b);
  }
  pend.swap(open);
}

void KernTimers::reset() {
  flush();
  acc.clear();
}

double KernTimers::seconds_matching(const char *prefix) {
  flush();
  double t = 0.0;
  const size_t L = strlen(prefix);
  for (const Acc &a : acc)
    if (a.name.compare(0, L, prefix) == 0) t += a.sec;
  return t;
}

// about three nonzeros per lane (measured best for 7-point stencils: 2)
int auto_lanes(double mean_row) {
  int L = 1;
  while (L < 64 && 2 * L <= mean_row / 3.0) L *= 2;
  return L;
}

void apply_A(Ctx &c, const double *dx, double *dy) {
  if (c.stats) c.stats->n_mvA++;
  if (c.A) {
    const int lanes =
        c.lanes ? c.lanes
                : auto_lanes(c.A->n ? (double)c.A->nnz / c.A->n : 1.0);
    spmv_f64(*c.A, dx, dy, lanes, c.stream);
    return;
  }
  if (!c.host_mv) fatal("matrix A not set", __FILE__, __LINE__);
  // host callback: the operator lives on the host (MVFunc, INC/pevsl_struct.h:208)
  const size_t n = c.h_x.size();
  PV_HIP(hipMemcpyAsync(c.h_x.data(), dx, n * sizeof(double),
                        hipMemcpyDeviceToHost, c.stream));
  PV_HIP(hipStreamSynchronize(c.stream));
  const double t0 = wtime();
  c.host_mv(c.h_x.data(), c.h_y.data(), c.host_mv_data);
  if (c.stats) c.stats->t_mvA += wtime() - t0;
  PV_HIP(hipMemcpyAsync(dy, c.h_y.data(), n * sizeof(double),
                        hipMemcpyHostToDevice, c.stream));
}

// The reference draws its random vectors from libc rand() after srand(rank)
// (SRC/pevsl.c:36, SRC/parvec.c:65-72).  The library keeps that stream in a
// generator of its own -- glibc's rand() algorithm (additive feedback,
// r_i = r_{i-31} + r_{i-3} mod 2^32 over a state seeded by the multiplier-16807
// congruential generator, first 310 outputs discarded, output r_i >> 1) -- so
// that the values equal the reference's and no other user of libc rand() in
// the process (runtime or math libraries initialising themselves) can shift
// them.
namespace {
struct RandStream {
  uint32_t r[34];  // r[k mod 34] = r_k for the last 34 k
  int i = 4;       // next k mod 34
  void seed(unsigned s) {
    int32_t w = s ? (int32_t)s : 1;
    uint32_t full[344];
    full[0] = (uint32_t)w;
    for (int k = 1; k < 31; k++) {
      // 16807 * w mod (2^31 - 1) by Schrage's method
      const int32_t hi = w / 127773, lo = w % 127773;
      w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      full[k] = (uint32_t)w;
    }
    for (int k = 31; k < 34; k++) full[k] = full[k - 31];
    for (int k = 34; k < 344; k++) full[k] = full[k - 31] + full[k - 3];
    for (int k = 310; k < 344; k++) r[k % 34] = full[k];
    i = 344 % 34;
  }
  int next() {
    // r_k = r_{k-31} + r_{k-3}: slots (k + 3) and (k + 31) mod 34
    const uint32_t v = r[(i + 3) % 34] + r[(i + 31) % 34];
    r[i] = v;
    i = (i + 1) % 34;
    return (int)(v >> 1);
  }
};
RandStream g_rand;
int lib_rand() { return g_rand.next(); }
}  // namespace

void rand_vectors(Ctx &c, int n, int nvec, double *d_out, size_t ld) {
  if (c.rand_mode == PEVSL_RAND_LIBC) {
    std::vector<double> h((size_t)n);
    const double t = ((double)RAND_MAX) / 2.0;
    for (int v = 0; v < nvec; v++) {
      for (int i = 0; i < n; i++) h[i] = (lib_rand() - t) / t;
      PV_HIP(hipMemcpyAsync(d_out + (size_t)v * ld, h.data(),
                            sizeof(double) * n, hipMemcpyHostToDevice,
                            c.stream));
      PV_HIP(hipStreamSynchronize(c.stream));
    }
    return;
  }
  for (int v = 0; v < nvec; v++) {
    rand_fill((size_t)n, d_out + (size_t)v * ld, c.rand_seed, c.rand_offset,
              c.rand_mode == PEVSL_RAND_RADEMACHER, c.stream);
    c.rand_offset += (unsigned long long)n;
  }
}

}  // namespace pevsl

using namespace pevsl;

// host matvec installed in Amv->func by SetAParcsr, for callers that invoke
// the hook directly with host arrays (pEVSL_ParcsrMatvec0, SRC/parcsrmv.c:66)
static void parcsr_mv_host(double *x, double *y, void *data) {
  pevsl_Parcsr *A = (pevsl_Parcsr *)data;
  pevsl_Parvec X, Y;
  pEVSL_ParvecCreateShell(A->ncol_global, A->ncol_local, A->first_col, A->comm,
                          &X, x);
  pEVSL_ParvecCreateShell(A->nrow_global, A->nrow_local, A->first_row, A->comm,
                          &Y, y);
  pEVSL_ParcsrMatvec(A, &X, &Y);
}

extern "C" {

// ---------------------------------------------------------------------------
// Instance (SRC/pevsl.c)
// ---------------------------------------------------------------------------

int pEVSL_Start(MPI_Comm comm, pevsl_Data **data) {
  check_device();
  pevsl_Data *p;
  PEVSL_MALLOC(p, 1, pevsl_Data);
  memset(p, 0, sizeof(*p));
  p->comm = comm;
  p->N = PEVSL_NOT_DEFINED;
  p->n = PEVSL_NOT_DEFINED;
  p->nfirst = PEVSL_NOT_DEFINED;
  p->sigma_mult = 1.0;
  // one process: rank 0 seeds the rand() stream (SRC/pevsl.c:31-36)
  g_rand.seed(0);
  PEVSL_MALLOC(p->stats, 1, pevsl_Stat);
  Ctx *c = new Ctx();
  PV_HIP(hipGetDevice(&c->device));
  PV_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  PV_BLAS(rocblas_create_handle(&c->blas));
  PV_BLAS(rocblas_set_stream(c->blas, c->stream));
  c->stats = p->stats;
  p->dev = c;
  pEVSL_StatsReset(p);
  *data = p;
  return 0;
}

int pEVSL_Finish(pevsl_Data *p) {
  if (!p) return 0;
  if (p->dev) {
    Ctx *c = static_cast<Ctx *>(p->dev);
    (void)hipStreamSynchronize(c->stream);
    if (c->blas) rocblas_destroy_handle(c->blas);
    hipStream_t st = c->stream;
    delete c;  // device buffers are released before the stream
    (void)hipStreamDestroy(st);
  }
  PEVSL_FREE(p->Amv);
  PEVSL_FREE(p->Bmv);
  PEVSL_FREE(p->Bsol);
  PEVSL_FREE(p->LTsol);
  PEVSL_FREE(p->stats);
  free(p);
  return 0;
}

int pEVSL_SetProbSizes(pevsl_Data *p, int N, int n, int nfirst) {
  p->N = N;
  p->n = n;
  p->nfirst = nfirst < 0 ? PEVSL_NOT_DEFINED : nfirst;
  return 0;
}

int pEVSL_SetAParcsr(pevsl_Data *p, pevsl_Parcsr *A) {
  Ctx &c = ctx_of(p);
  // the eigenvalue problem needs a square operator
  if (A->nrow_global != A->ncol_global || A->nrow_local != A->ncol_local) return -1;
  p->N = A->ncol_global;
  p->n = A->ncol_local;
  p->nfirst = A->first_col;
  if (!p->Amv) PEVSL_CALLOC(p->Amv, 1, pevsl_Matvec);
  p->Amv->func = parcsr_mv_host;
  p->Amv->data = (void *)A;
  c.A_src = dev_matrix_of(A);
  if (!c.A_src) fatal("SetAParcsr: matrix has no device copy (ParcsrSetup not called)",
                      __FILE__, __LINE__);
  c.A_view = *c.A_src;
  if (!c.col16) c.A_view.jo16 = nullptr;
  c.A = &c.A_view;
  c.host_mv = nullptr;
  return 0;
}

int pEVSL_SetAMatvec(pevsl_Data *p, MVFunc func, void *data) {
  Ctx &c = ctx_of(p);
  if (!p->Amv) PEVSL_CALLOC(p->Amv, 1, pevsl_Matvec);
  p->Amv->func = func;
  p->Amv->data = data;
  c.A = nullptr;
  c.A_src = nullptr;
  c.host_mv = func;
  c.host_mv_data = data;
  if (p->n > 0) {
    c.h_x.assign((size_t)p->n, 0.0);
    c.h_y.assign((size_t)p->n, 0.0);
  }
  return 0;
}

int pEVSL_SetStdEig(pevsl_Data *p) {
  p->ifGenEv = 0;
  return 0;
}
int pEVSL_SetSigmaMult(pevsl_Data *p, double mult) {
  p->sigma_mult = mult;
  return 0;
}

// ---------------------------------------------------------------------------
// Extensions (pevsl_amd.h)
// ---------------------------------------------------------------------------

int pEVSL_SetFilterPrecision(pevsl_Data *p, int prec) {
  if (prec != PEVSL_PREC_FP64 && prec != PEVSL_PREC_FP32) return -1;
  ctx_of(p).filter_prec = prec;
  return 0;
}

int pEVSL_SetBasisPrecision(pevsl_Data *p, int prec) {
  if (prec != PEVSL_PREC_FP64 && prec != PEVSL_PREC_FP32) return -1;
  ctx_of(p).basis_prec = prec;
  return 0;
}

int pEVSL_SetRitzRefine(pevsl_Data *p, int mode) {
  if (mode < -1 || mode > 1) return -1;
  ctx_of(p).ritz_refine = mode;
  return 0;
}

int pEVSL_SetMixedTol(pevsl_Data *p, double tol) {
  if (!(tol > 0)) return -1;
  ctx_of(p).mixed_tol = tol;
  return 0;
}

int pEVSL_SetRandMode(pevsl_Data *p, int mode, unsigned long long seed) {
  if (mode < PEVSL_RAND_LIBC || mode > PEVSL_RAND_UNIFORM) return -1;
  Ctx &c = ctx_of(p);
  c.rand_mode = mode;
  c.rand_seed = seed;
  c.rand_offset = 0;
  return 0;
}

int pEVSL_SetKeepVectors(pevsl_Data *p, int keep) {
  ctx_of(p).keep_vectors = keep ? 1 : 0;
  return 0;
}

int pEVSL_SetSpmvTuning(pevsl_Data *p, int lanes, int bw) {
  auto pow2 = [](int v) { return v >= 0 && v <= 64 && (v & (v - 1)) == 0; };
  if (!(pow2(lanes) || lanes == PEVSL_SPMV_STREAM) || !pow2(bw)) return -1;
  Ctx &c = ctx_of(p);
  if (lanes == PEVSL_SPMV_STREAM) {
    c.use_stream = 1;
    c.lanes = 0;
  } else {
    c.use_stream = lanes ? 0 : -1;
    c.lanes = lanes;
  }
  c.block_width = bw;
  return 0;
}

int pEVSL_SetGraphFilter(pevsl_Data *p, int on) {
  if (on != 0 && on != 1) return -1;
  ctx_of(p).graph_filter = on;
  return 0;
}

int pEVSL_SetColumnOffsets16(pevsl_Data *p, int on) {
  if (on != 0 && on != 1) return -1;
  Ctx &c = ctx_of(p);
  c.col16 = on;
  if (c.A_src) {
    c.A_view = *c.A_src;
    if (!on) c.A_view.jo16 = nullptr;
  }
  return 0;
}

int pEVSL_ColumnOffsets16(pevsl_Data *p) {
  const Ctx &c = ctx_of(p);
  return c.A && c.A->jo16 ? 1 : 0;
}

int pEVSL_SetBatchWave(pevsl_Data *p, int wave) {
  if (wave < 0) return -1;
  ctx_of(p).batch_wave = wave;
  return 0;
}

int pEVSL_SetBlockKernel(pevsl_Data *p, int rowkernel) {
  if (rowkernel < -1 || rowkernel > 5) return -1;
  ctx_of(p).blk_rowkernel = rowkernel;
  return 0;
}

int pEVSL_GetLanInfo(pevsl_Data *p, int s, pevsl_LanInfo *info) {
  Ctx &c = ctx_of(p);
  if (s < 0 || s >= (int)c.info.size()) return -1;
  *info = c.info[(size_t)s];
  return 0;
}

int pEVSL_GetKernelStats(pevsl_Data *p, pevsl_KernelStats *st) {
  *st = ctx_of(p).kstats;
  return 0;
}

int pEVSL_GetKernelTimes(pevsl_Data *p, int maxn, pevsl_KernelTime *out, int *count) {
  Ctx &c = ctx_of(p);
  c.kt.flush();
  *count = (int)c.kt.acc.size();
  if (maxn < 0 || (maxn > 0 && !out)) return -1;
  for (int i = 0; i < std::min(maxn, *count); i++) {
    const KernTimers::Acc &a = c.kt.acc[(size_t)i];
    memset(out[i].name, 0, sizeof(out[i].name));
    strncpy(out[i].name, a.name.c_str(), sizeof(out[i].name) - 1);
    out[i].launches = a.launches;
    out[i].seconds = a.sec;
    out[i].bytes = a.bytes;
  }
  return 0;
}

int pEVSL_DeviceMatvec(pevsl_Data *p, const double *dx, double *dy) {
  Ctx &c = ctx_of(p);
  apply_A(c, dx, dy);
  PV_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int pEVSL_DeviceInfo(char *name, int len, size_t *hbm) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
  if (name && len > 0) {
    snprintf(name, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (hbm) *hbm = prop.totalGlobalMem;
  return 0;
}

// ---------------------------------------------------------------------------
// Parallel CSR (SRC/parcsr.c) -- one rank
// ---------------------------------------------------------------------------

int pEVSL_ParcsrCreate(int nrow, int ncol, int *row_starts, int *col_starts,
                       pevsl_Parcsr *A, MPI_Comm comm) {
  memset(A, 0, sizeof(*A));
  A->comm = comm;
  A->nrow_global = nrow;
  A->ncol_global = ncol;
  if (row_starts) {
    if (row_starts[0] != 0 || row_starts[1] != nrow) return -1;
    PEVSL_MALLOC(A->row_starts, 2, int);
    memcpy(A->row_starts, row_starts, 2 * sizeof(int));
  }
  if (col_starts) {
    if (col_starts[0] != 0 || col_starts[1] != ncol) return -1;
    PEVSL_MALLOC(A->col_starts, 2, int);
    memcpy(A->col_starts, col_starts, 2 * sizeof(int));
  }
  A->nrow_local = nrow;
  A->ncol_local = ncol;
  A->first_row = 0;
  A->first_col = 0;
  return 0;
}

int pEVSL_ParcsrSetup(pevsl_Csr *Ai, pevsl_Parcsr *A) {
  const int n = A->nrow_local;
  if (!Ai || Ai->nrows != n) return -1;
  // The kernels index x and their LDS staging with ia/ja: a malformed CSR
  // (ia not starting at 0 or decreasing, a column outside [0, ncol)) is
  // rejected here, before anything is allocated or reaches the device.
  if (n < 0 || !Ai->ia || Ai->ia[0] != 0) return -2;
  for (int i = 0; i < n; i++)
    if (Ai->ia[i + 1] < Ai->ia[i]) return -2;
  const int nnz = Ai->ia[n];
  if (nnz > 0 && (!Ai->ja || !Ai->a)) return -2;
  for (int j = 0; j < nnz; j++)
    if (Ai->ja[j] < 0 || Ai->ja[j] >= A->ncol_local) return -2;
  check_device();
  // diag = the whole local matrix (all columns are local with one rank)
  PEVSL_MALLOC(A->diag, 1, pevsl_Csr);
  pEVSL_CsrResize(n, A->ncol_local, nnz, A->diag);
  memcpy(A->diag->ia, Ai->ia, sizeof(int) * (size_t)(n + 1));
  memcpy(A->diag->ja, Ai->ja, sizeof(int) * (size_t)nnz);
  memcpy(A->diag->a, Ai->a, sizeof(double) * (size_t)nnz);
  PEVSL_MALLOC(A->offd, 1, pevsl_Csr);
  pEVSL_CsrResize(n, 0, 0, A->offd);
  for (int i = 0; i <= n; i++) A->offd->ia[i] = 0;
  A->col_map_offd = NULL;
  PEVSL_CALLOC(A->comm_handle, 1, commHandle);
  // device copy
  DevMatrix *M = new DevMatrix();
  M->csr.n = n;
  M->csr.nnz = nnz;
  M->csr.ia = M->ia.ensure((size_t)n + 1);
  M->csr.ja = M->ja.ensure((size_t)nnz);
  M->csr.a64 = M->a64.ensure((size_t)nnz);
  M->csr.a32 = M->a32.ensure((size_t)nnz);
  int mr = 0;
  for (int i = 0; i < n; i++) mr = std::max(mr, Ai->ia[i + 1] - Ai->ia[i]);
  M->csr.max_row = mr;
  std::vector<int> rb;
  build_row_blocks(n, Ai->ia, rb);
  M->csr.nrowblk = (int)rb.size() - 1;
  M->csr.rowblk = M->rowblk.ensure(rb.size());
  PV_HIP(hipMemcpy(M->csr.rowblk, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice));
  // 16-bit column offsets from each row block's first row, kept when every
  // nonzero's offset fits (banded orderings such as the natural order of a
  // grid Laplacian; not random sparsity)
  std::vector<short> jo((size_t)nnz);
  bool fits = nnz > 0;
  for (size_t b = 0; fits && b + 1 < rb.size(); b++) {
    const int r0 = rb[b];
    for (int j = Ai->ia[r0]; j < Ai->ia[rb[b + 1]]; j++) {
      const int d = Ai->ja[j] - r0;
      if (d < -32768 || d > 32767) {
        fits = false;
        break;
      }
      jo[(size_t)j] = (short)d;
    }
  }
  if (fits) {
    M->csr.jo16 = M->jo16.ensure((size_t)nnz);
    PV_HIP(hipMemcpy(M->csr.jo16, jo.data(), sizeof(short) * (size_t)nnz,
                     hipMemcpyHostToDevice));
  }
  M->csr.rowblk_q[0] = M->csr.rowblk;
  M->csr.nrowblk_q[0] = M->csr.nrowblk;
  for (int l = 1; l < 5; l++) {
    build_row_blocks(n, Ai->ia, rb, kStreamRows >> l, kStreamNnz >> l);
    M->csr.nrowblk_q[l] = (int)rb.size() - 1;
    M->csr.rowblk_q[l] = M->rowblk_q[l].ensure(rb.size());
    PV_HIP(hipMemcpy(M->csr.rowblk_q[l], rb.data(), sizeof(int) * rb.size(),
                     hipMemcpyHostToDevice));
  }
  std::vector<float> a32((size_t)nnz);
  for (int j = 0; j < nnz; j++) a32[j] = (float)Ai->a[j];
  PV_HIP(hipMemcpy(M->csr.ia, Ai->ia, sizeof(int) * (size_t)(n + 1),
                   hipMemcpyHostToDevice));
  PV_HIP(hipMemcpy(M->csr.ja, Ai->ja, sizeof(int) * (size_t)nnz,
                   hipMemcpyHostToDevice));
  PV_HIP(hipMemcpy(M->csr.a64, Ai->a, sizeof(double) * (size_t)nnz,
                   hipMemcpyHostToDevice));
  PV_HIP(hipMemcpy(M->csr.a32, a32.data(), sizeof(float) * (size_t)nnz,
                   hipMemcpyHostToDevice));
  A->dev = M;
  return 0;
}

void pEVSL_ParcsrFree(pevsl_Parcsr *A) {
  if (A->comm_handle) PEVSL_FREE(A->comm_handle);
  PEVSL_FREE(A->row_starts);
  PEVSL_FREE(A->col_starts);
  if (A->diag) {
    pEVSL_FreeCsr(A->diag);
    PEVSL_FREE(A->diag);
  }
  if (A->offd) {
    pEVSL_FreeCsr(A->offd);
    PEVSL_FREE(A->offd);
  }
  PEVSL_FREE(A->col_map_offd);
  if (A->dev) {
    delete static_cast<DevMatrix *>(A->dev);
    A->dev = NULL;
  }
}

int pEVSL_ParcsrGetLocalMat(pevsl_Parcsr *A, int idx, pevsl_Coo *coo,
                            pevsl_Csr *csr, char stype) {
  // whole matrix in diag; SRC/parcsr.c:319-401 with one rank
  stype = (char)toupper(stype);
  if (stype != 'L' && stype != 'U') stype = 'A';
  pevsl_Csr *Ad = A->diag;
  const int n = A->nrow_local;
  pevsl_Csr loc;
  pEVSL_CsrResize(n, A->ncol_global, Ad->ia[n], &loc);
  int k = 0;
  loc.ia[0] = 0;
  for (int i = 0; i < n; i++) {
    for (int j = Ad->ia[i]; j < Ad->ia[i + 1]; j++) {
      const int gc = Ad->ja[j];
      if ((stype == 'L' && i >= gc) || (stype == 'U' && i <= gc) ||
          stype == 'A') {
        loc.ja[k] = gc;
        loc.a[k] = Ad->a[j];
        k++;
      }
    }
    loc.ia[i + 1] = k;
  }
  pEVSL_SortRow(&loc);
  if (coo) pEVSL_CsrToCoo(&loc, idx, coo);
  if (csr) {
    if (idx) {
      for (int i = 0; i < n; i++)
        for (int j = loc.ia[i]; j < loc.ia[i + 1]; j++) loc.ja[j]++;
      for (int i = 0; i <= n; i++) loc.ia[i]++;
    }
    *csr = loc;
  } else {
    pEVSL_FreeCsr(&loc);
  }
  return 0;
}

int pEVSL_ParcsrNnz(pevsl_Parcsr *A) { return A->diag->ia[A->nrow_local]; }
int pEVSL_ParcsrLocalNnz(pevsl_Parcsr *A) {
  return A->diag->ia[A->nrow_local];
}

void pEVSL_ParcsrMatvec(pevsl_Parcsr *A, pevsl_Parvec *x, pevsl_Parvec *y) {
  const DevCsr *M = dev_matrix_of(A);
  if (!M) fatal("ParcsrMatvec: no device matrix", __FILE__, __LINE__);
  // y (nrow_local) = A x (ncol_local); A may be rectangular
  const int n = A->nrow_local, m = A->ncol_local;
  if (x->n_local != m || y->n_local != n) fatal("ParcsrMatvec: vector sizes", __FILE__, __LINE__);
  double *d = nullptr;
  PV_HIP(hipMalloc(&d, sizeof(double) * ((size_t)n + (size_t)m + 1)));
  PV_HIP(hipMemcpy(d, x->data, sizeof(double) * (size_t)m, hipMemcpyHostToDevice));
  spmv_f64(*M, d, d + m, auto_lanes(n ? (double)M->nnz / n : 1.0), 0);
  PV_HIP(hipMemcpy(y->data, d + m, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  PV_HIP(hipFree(d));
}

// ---------------------------------------------------------------------------
// Parallel vectors (SRC/parvec.c) -- host arrays
// ---------------------------------------------------------------------------

void pEVSL_ParvecCreateShell(int nglobal, int nlocal, int nfirst,
                             MPI_Comm comm, pevsl_Parvec *x, double *data) {
  x->comm = comm;
  x->n_global = nglobal;
  x->n_local = nlocal;
  x->n_first = nfirst < 0 ? PEVSL_NOT_DEFINED : nfirst;
  x->data = data;
}

void pEVSL_ParvecCreate(int nglobal, int nlocal, int nfirst, MPI_Comm comm,
                        pevsl_Parvec *x) {
  pEVSL_ParvecCreateShell(nglobal, nlocal, nfirst, comm, x, NULL);
  PEVSL_MALLOC(x->data, nlocal, double);
}

void pEVSL_ParvecDupl(pevsl_Parvec *x, pevsl_Parvec *y) {
  pEVSL_ParvecCreate(x->n_global, x->n_local, x->n_first, x->comm, y);
}

void pEVSL_ParvecFree(pevsl_Parvec *x) { PEVSL_FREE(x->data); }

void pEVSL_ParvecRand(pevsl_Parvec *x) {
  const double t = ((double)RAND_MAX) / 2.0;
  for (int i = 0; i < x->n_local; i++) x->data[i] = (lib_rand() - t) / t;
}

void pEVSL_RandSeed(unsigned seed) { g_rand.seed(seed); }

void pEVSL_ParvecDot(pevsl_Parvec *x, pevsl_Parvec *y, double *t) {
  double s = 0.0;
  for (int i = 0; i < x->n_local; i++) s += x->data[i] * y->data[i];
  *t = s;
}

void pEVSL_ParvecNrm2(pevsl_Parvec *x, double *t) {
  double t2;
  pEVSL_ParvecDot(x, x, &t2);
  *t = sqrt(t2);
}

void pEVSL_ParvecCopy(pevsl_Parvec *x, pevsl_Parvec *y) {
  memcpy(y->data, x->data, sizeof(double) * (size_t)x->n_local);
}

void pEVSL_ParvecSum(pevsl_Parvec *x, double *t) {
  double s = 0.0;
  for (int i = 0; i < x->n_local; i++) s += x->data[i];
  *t = s;
}

void pEVSL_ParvecScal(pevsl_Parvec *x, double t) {
  for (int i = 0; i < x->n_local; i++) x->data[i] *= t;
}

void pEVSL_ParvecSetScalar(pevsl_Parvec *x, double t) {
  for (int i = 0; i < x->n_local; i++) x->data[i] = t;
}

void pEVSL_ParvecSetZero(pevsl_Parvec *x) {
  memset(x->data, 0, sizeof(double) * (size_t)x->n_local);
}

void pEVSL_ParvecAxpy(double a, pevsl_Parvec *x, pevsl_Parvec *y) {
  for (int i = 0; i < x->n_local; i++) y->data[i] += a * x->data[i];
}

int pEVSL_ParvecSameSize(pevsl_Parvec *x, pevsl_Parvec *y) {
  return x->n_global == y->n_global && x->n_local == y->n_local &&
         x->n_first == y->n_first;
}

int pEVSL_ParvecWrite(pevsl_Parvec *x, const char *fn) {
  FILE *fp = fn ? fopen(fn, "w") : stdout;
  if (!fp) return -1;
  fprintf(fp, "%% %d\n", x->n_global);
  for (int j = 0; j < x->n_local; j++) fprintf(fp, "%.15e\n", x->data[j]);
  if (fn) fclose(fp);
  return 0;
}

// ---------------------------------------------------------------------------
// Multi-vectors (SRC/parvecs.c) -- host arrays, column-major
// ---------------------------------------------------------------------------

void pEVSL_ParvecsCreateShell(int nglobal, int nvecs, int ld, int nlocal,
                              int nfirst, MPI_Comm comm, pevsl_Parvecs *x,
                              double *data) {
  x->comm = comm;
  x->n_global = nglobal;
  x->n_local = nlocal;
  x->n_first = nfirst < 0 ? PEVSL_NOT_DEFINED : nfirst;
  x->n_vecs = nvecs;
  x->ld = ld;
  x->data = data;
}

void pEVSL_ParvecsCreate(int nglobal, int nvecs, int ld, int nlocal,
                         int nfirst, MPI_Comm comm, pevsl_Parvecs *x) {
  pEVSL_ParvecsCreateShell(nglobal, nvecs, ld, nlocal, nfirst, comm, x, NULL);
  PEVSL_MALLOC(x->data, (size_t)ld * (size_t)nvecs, double);
}

void pEVSL_ParvecsDuplParvec(pevsl_Parvec *x, int nvecs, int ld,
                             pevsl_Parvecs *y) {
  pEVSL_ParvecsCreate(x->n_global, nvecs, ld, x->n_local, x->n_first, x->comm,
                      y);
}

void pEVSL_ParvecsFree(pevsl_Parvecs *x) { PEVSL_FREE(x->data); }

void pEVSL_ParvecsGetParvecShell(pevsl_Parvecs *X, int i, pevsl_Parvec *x) {
  double *data = i < 0 ? NULL : X->data + (size_t)X->ld * (size_t)i;
  pEVSL_ParvecCreateShell(X->n_global, X->n_local, X->n_first, X->comm, x,
                          data);
}

void pEVSL_ParvecsGemv(double alp, pevsl_Parvecs *A, int nvecs, double *x,
                       double bet, pevsl_Parvec *y) {
  const int n = A->n_local;
  for (int i = 0; i < n; i++) y->data[i] *= bet;
  for (int j = 0; j < nvecs; j++) {
    const double t = alp * x[j];
    const double *a = A->data + (size_t)A->ld * j;
    for (int i = 0; i < n; i++) y->data[i] += t * a[i];
  }
}

void pEVSL_ParvecsGemtvWithWspace(double alp, pevsl_Parvecs *A, int nvecs,
                                  pevsl_Parvec *x, double bet, double *y,
                                  double *w) {
  const int n = A->n_local;
  for (int j = 0; j < nvecs; j++) {
    const double *a = A->data + (size_t)A->ld * j;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * x->data[i];
    w[j] = s;
  }
  for (int j = 0; j < nvecs; j++) y[j] = bet * y[j] + alp * w[j];
}

void pEVSL_ParvecsGemtv(double alp, pevsl_Parvecs *A, int nvecs,
                        pevsl_Parvec *x, double bet, double *y) {
  std::vector<double> w((size_t)nvecs);
  pEVSL_ParvecsGemtvWithWspace(alp, A, nvecs, x, bet, y, w.data());
}

// ---------------------------------------------------------------------------
// Sparse utilities (SRC/spmat.c) -- host
// ---------------------------------------------------------------------------

void pEVSL_CsrResize(int nrow, int ncol, int nnz, pevsl_Csr *csr) {
  csr->nrows = nrow;
  csr->ncols = ncol;
  PEVSL_MALLOC(csr->ia, nrow + 1, int);
  PEVSL_MALLOC(csr->ja, nnz, int);
  PEVSL_MALLOC(csr->a, nnz, double);
}

void pEVSL_FreeCsr(pevsl_Csr *csr) {
  PEVSL_FREE(csr->ia);
  PEVSL_FREE(csr->ja);
  PEVSL_FREE(csr->a);
}

void pEVSL_CooResize(int nrow, int ncol, int nnz, pevsl_Coo *coo) {
  coo->nrows = nrow;
  coo->ncols = ncol;
  coo->nnz = nnz;
  PEVSL_MALLOC(coo->ir, nnz, int);
  PEVSL_MALLOC(coo->jc, nnz, int);
  PEVSL_MALLOC(coo->vv, nnz, double);
}

void pEVSL_FreeCoo(pevsl_Coo *coo) {
  PEVSL_FREE(coo->ir);
  PEVSL_FREE(coo->jc);
  PEVSL_FREE(coo->vv);
}

// counting sort by row, stable within a row (SRC/spmat.c:56-86)
int pEVSL_CooToCsr(int cooidx, pevsl_Coo *coo, pevsl_Csr *csr) {
  cooidx = cooidx != 0;
  const int nnz = coo->nnz, nr = coo->nrows;
  // an index outside the declared size would be counted out of bounds
  for (int i = 0; i < nnz; i++) {
    const int r = coo->ir[i] - cooidx, c = coo->jc[i] - cooidx;
    if (r < 0 || r >= nr || c < 0 || c >= coo->ncols) return -1;
  }
  pEVSL_CsrResize(nr, coo->ncols, nnz, csr);
  std::vector<int> next((size_t)nr + 1, 0);
  for (int i = 0; i < nnz; i++) next[(size_t)(coo->ir[i] - cooidx) + 1]++;
  for (int i = 0; i < nr; i++) next[(size_t)i + 1] += next[(size_t)i];
  memcpy(csr->ia, next.data(), sizeof(int) * (size_t)(nr + 1));
  for (int i = 0; i < nnz; i++) {
    const int r = coo->ir[i] - cooidx;
    const int k = next[(size_t)r]++;
    csr->ja[k] = coo->jc[i] - cooidx;
    csr->a[k] = coo->vv[i];
  }
  return 0;
}

int pEVSL_CsrToCoo(pevsl_Csr *csr, int cooidx, pevsl_Coo *coo) {
  cooidx = cooidx != 0;
  const int nnz = csr->ia[csr->nrows];
  pEVSL_CooResize(csr->nrows, csr->ncols, nnz, coo);
  int k = 0;
  for (int i = 0; i < csr->nrows; i++)
    for (int j = csr->ia[i]; j < csr->ia[i + 1]; j++, k++) {
      coo->ir[k] = i + cooidx;
      coo->jc[k] = csr->ja[j] + cooidx;
      coo->vv[k] = csr->a[j];
    }
  return 0;
}

int pEVSL_MatvecGen(double alp, pevsl_Csr *A, double *x, double bet,
                    double *y) {
  for (int i = 0; i < A->nrows; i++) {
    double r = 0.0;
    for (int j = A->ia[i]; j < A->ia[i + 1]; j++) r += A->a[j] * x[A->ja[j]];
    y[i] = bet * y[i] + alp * r;
  }
  return 0;
}

int pEVSL_Matvec(pevsl_Csr *A, double *x, double *y) {
  for (int i = 0; i < A->nrows; i++) {
    double r = 0.0;
    for (int j = A->ia[i]; j < A->ia[i + 1]; j++) r += A->a[j] * x[A->ja[j]];
    y[i] = r;
  }
  return 0;
}

void pEVSL_SortRow(pevsl_Csr *A) {
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < A->nrows; i++) {
    row.clear();
    for (int j = A->ia[i]; j < A->ia[i + 1]; j++)
      row.emplace_back(A->ja[j], A->a[j]);
    std::sort(row.begin(), row.end(),
              [](const std::pair<int, double> &a,
                 const std::pair<int, double> &b) { return a.first < b.first; });
    int k = A->ia[i];
    for (auto &e : row) {
      A->ja[k] = e.first;
      A->a[k] = e.second;
      k++;
    }
  }
}

// ---------------------------------------------------------------------------
// Utilities (SRC/utils.c, SRC/vector.c)
// ---------------------------------------------------------------------------

void pEVSL_SortInt(int *x, int n) { std::sort(x, x + n); }

void sort_double(int n, double *v, int *ind) {
  if (!ind) {
    std::sort(v, v + n);
    return;
  }
  std::vector<std::pair<double, int>> t((size_t)n);
  for (int i = 0; i < n; i++) t[(size_t)i] = {v[i], i};
  std::stable_sort(t.begin(), t.end(),
                   [](const std::pair<double, int> &a,
                      const std::pair<double, int> &b) {
                     return a.first < b.first;
                   });
  for (int i = 0; i < n; i++) {
    v[i] = t[(size_t)i].first;
    ind[i] = t[(size_t)i].second;
  }
}

void pEVSL_SortDouble(int n, double *v, int *ind) { sort_double(n, v, ind); }

void linspace(double a, double b, int num, double *arr) {
  const double h = num == 1 ? 0 : (b - a) / (num - 1);
  arr[0] = a;
  arr[num - 1] = b;
  for (int i = 1; i < num - 1; i++) arr[i] = a + i * h;
}

void pEVSL_LinSpace(double a, double b, int num, double *arr) {
  linspace(a, b, num, arr);
}

// 1-D block partition (SRC/utils.c:128-149)
void pEVSL_Part1d(int len, int pnum, int *idx, int *j1, int *j2, int job) {
  int size = (len + pnum - 1) / pnum;
  const int cc = pnum - (size * pnum - len);
  if (job == 1) {
    if (*idx < cc) {
      *j1 = *idx * size;
    } else {
      --size;
      *j1 = len - size * (pnum - *idx);
    }
    *j2 = *j1 + size;
  } else {
    if (*j1 < size * cc) {
      *idx = *j1 / size;
    } else {
      *idx = cc + (*j1 - size * cc) / (size - 1);
      --size;
    }
    if (j2) *j2 = size;
  }
}

int pEVSL_BinarySearch(int *x, int n, int key) {
  int *p = std::lower_bound(x, x + n, key);
  return (p != x + n && *p == key) ? (int)(p - x) : -1;
}

void vecset(int n, double t, double *v) {
  for (int i = 0; i < n; i++) v[i] = t;
}
void pEVSL_Vecset(int n, double t, double *v) { vecset(n, t, v); }

double pEVSL_Wtime(void) { return wtime(); }

// ---------------------------------------------------------------------------
// Stats (SRC/stats.c)
// ---------------------------------------------------------------------------

void pEVSL_StatsReset(pevsl_Data *p) {
  memset(p->stats, 0, sizeof(pevsl_Stat));
  if (p->dev) {
    memset(&ctx_of(p).kstats, 0, sizeof(pevsl_KernelStats));
    ctx_of(p).kt.reset();
  }
}

void pEVSL_StatsPrint(pevsl_Data *p, FILE *f) {
  if (!f) return;
  const pevsl_Stat *s = p->stats;
  fprintf(f, "= = = = = = = = = = = = = = = = = = = Time & Memory Stats = = = "
             "= = = = = = = = = = = = = = = = =\n");
  fprintf(f, " Timing (sec):\n");
  if (s->t_dos) fprintf(f, "   Compute DOS               :  %f\n", s->t_dos);
  if (s->t_iter) fprintf(f, "   Iteration time (tot)      :  %f\n", s->t_iter);
  fprintf(f, "   - - - - - - - - - - - - - - - - -\n");
  if (s->n_polAv)
    fprintf(f, "   Pol(A)*v                  :  %f (%8lu, avg %f)\n",
            s->t_polAv, (unsigned long)s->n_polAv, s->t_polAv / s->n_polAv);
  if (s->n_mvA)
    fprintf(f, "   Matvec matrix A           :  %f (%8lu, avg %f)\n", s->t_mvA,
            (unsigned long)s->n_mvA, s->t_mvA / s->n_mvA);
  if (s->t_reorth) fprintf(f, "   Reorthogonalization       :  %f\n", s->t_reorth);
  if (s->t_eig) fprintf(f, "   LAPACK eig                :  %f\n", s->t_eig);
  if (s->t_ritz) fprintf(f, "   Compute Ritz vectors      :  %f\n", s->t_ritz);
  if (s->t_sth) fprintf(f, "   Other                     :  %f\n", s->t_sth);
  fprintf(f, "= = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = = "
             "= = = = = = = = = =\n");
}

}  // extern "C"
