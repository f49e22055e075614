// lanczos.cpp -- filtered Lanczos without restart (pEVSL_ChebLanNr) and its
// batched form (pEVSL_ChebLanNrBatch) on the GPU.
//
// Per slice the control flow is SRC/cheblanNr.c:54-448: filter the initial
// vector, then for k = 0.. : z_{k+1} = p(A) z_k, three-term update with
// alpha/beta, full reorthogonalization by classical Gram-Schmidt with the
// DGKS test (SRC/miscla.c:211-240), lucky-breakdown restart from a fresh
// filtered random vector, and every `cycle` steps from `Ntest` on the
// restricted trace of the Ritz values of p(A) above `bar` as the stopping
// test; at the end the Ritz pairs of the selected Ritz values are formed and
// accepted if their Rayleigh quotient lies in the slice and the residual
// norm is below tol.
//
// GPU mapping:
//   * all slices of a batch step in lockstep; their current Lanczos vectors
//     are packed, in groups of similar degree, into interleaved blocks that
//     a block SpMM filters with per-column coefficients and degrees
//     (kernels.hip: the row kernels k_cheb_blkrow / k_clen_blkrow, the
//     nnz-balanced k_blk_bal for uneven rows); a single slice uses the
//     one-column CSR-stream kernel (k_cheb_stream) or, when tuned, k_cheb_vec;
//   * on the mixed path the filter runs in FP32 and the basis may be stored
//     in FP32; the Ritz stage then refines the selected vectors by an FP64
//     Rayleigh-Ritz projection (MFMA DGEMMs + rocSOLVER SYGVD);
//   * the Lanczos bases live in HBM in chunks of `chunk_cols` columns that
//     are allocated as the basis grows, so memory follows the actual step
//     count rather than maxit;
//   * alpha/beta, the DGKS decision and the norms stay on the device; the
//     host reads two scalars per slice per step for the breakdown test and
//     the tridiagonal matrix;
//   * the restricted trace uses GPU Sturm multisection on T_k; the final
//     eigenvectors of T come from rocSOLVER (divide and conquer), the Ritz
//     vectors from rocBLAS DGEMM over the basis chunks, the residuals from the
//     SpMV and column kernels.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <type_traits>

#include <rocsolver/rocsolver.h>

#include "internal.hpp"

namespace pevsl {

namespace {

constexpr int kNtest = 30;   // first convergence test (SRC/cheblanNr.c:82)
constexpr int kCycle = 20;   // test period (SRC/cheblanNr.c:84)

struct Slice {
  int idx = 0;
  double aa = 0, bb = 0, bar = 0, gam = 0;
  int maxit = 0;
  const pevsl_Polparams *pol = nullptr;
  std::vector<void *> chunks;  // basis chunks (FP64 or FP32 elements)
  std::vector<double> dT, eT;
  double wn = 0.0;
  int nwn = 0;
  double tr0 = 0.0, tr1 = 0.0;
  int kdim = 0, nconv = 0, nbreak = 0, its = 0;
  bool done = false;
  long long nmatvec = 0;
  double t_start = 0.0;
  // results
  int rc = 0;
  std::vector<double> lam, res;
  std::vector<double> vecs;  // accepted Ritz vectors (host, n x nev)
};

// Clenshaw step: gathered b1 (counted once), b2 (if any), v, out; idx =
// index bytes per nonzero the kernel reads (2 with 16-bit column offsets)
size_t clen_launch_bytes(const DevCsr &A, int elem, int ncol, bool has_b2, int idx) {
  const double vec = (double)A.n * elem;
  return (size_t)((double)A.nnz * (elem + idx) + 4.0 * (A.n + 1) +
                  (has_b2 ? 4.0 : 3.0) * vec * ncol);
}

// segmented Clenshaw step: the matrix once if any piece forms the product,
// then per active piece of `seg` columns two (top term: v, out), three (b1,
// v, out) or four vector passes
double clen_seg_launch_bytes(const DevCsr &A, int elem, int seg, int d, int nseg,
                             const SegTops &tops) {
  const double piece = (double)A.n * elem * seg;
  double b = tops.top[0] > d ? (double)A.nnz * (elem + 4) + 4.0 * (A.n + 1) : 0.0;
  for (int s = 0; s < nseg; s++)
    b += (d == tops.top[s] ? 2.0 : d == tops.top[s] - 1 ? 3.0 : 4.0) * piece;
  return b;
}

// three-term step: passes = vector passes per column (5: vk gathered, vkm1,
// vkp1, y read and written; 3 on the first degree, or with y skipped)
size_t filter_launch_bytes(const DevCsr &A, int elem, int ncol, int passes, int idx) {
  const double vec = (double)A.n * elem;
  return (size_t)((double)A.nnz * (elem + idx) + 4.0 * (A.n + 1) + passes * vec * ncol);
}

// name of a filter kernel for the per-kernel timers
std::string filter_kernel_name(bool f32, int kind, int B, bool stream, int lanes) {
  const char *t = f32 ? "float" : "double";
  char nm[48];
  if (B == 0) {
    if (stream)
      snprintf(nm, sizeof nm, "k_cheb_stream<%s>", t);
    else
      snprintf(nm, sizeof nm, "k_cheb_vec<%s,%d>", t, lanes);
  } else if (kind == 5) {
    snprintf(nm, sizeof nm, "k_clen_seg<%s,%d>", t, B);
  } else if (kind >= 3) {
    snprintf(nm, sizeof nm, "k_blk_bal<%s,%d,%s>", t, B, kind == 4 ? "EpiClen" : "EpiCheb");
  } else if (kind == 2) {
    snprintf(nm, sizeof nm, "k_clen_blkrow<%s,%d>", t, B);
  } else if (kind == 1) {
    snprintf(nm, sizeof nm, "k_cheb_blkrow<%s,%d>", t, B);
  } else {
    snprintf(nm, sizeof nm, "k_cheb_blk<%s,%d>", t, B);
  }
  return nm;
}

}  // namespace

// FP32 filter y = p((A - cc)/dd) v on FP32 columns; w holds 3n floats.
// `timed`: record the run in the instance's per-kernel HIP-event timers.
static void cheb_filter_f32(Ctx &c, const pevsl_Polparams *pol, const float *x32, float *y32,
                            float *w, bool timed = true) {
  const DevCsr &A = *c.A;
  const int n = A.n, m = pol->deg;
  const double *mu = pol->mu;
  const double cc = pol->cc, dd = pol->dd;
  const hipStream_t s = c.stream;
  const int lanes = c.lanes ? c.lanes : auto_lanes((double)A.nnz / std::max(1, n));
  const bool stream = c.use_stream != 0;
  float *b[3] = {w, w + (size_t)n, w + 2 * (size_t)n};
  const float *vk = x32, *vkm1 = nullptr;
  int nb = 0;
  const long tok = m > 0 && timed
                       ? c.kt.begin(s, filter_kernel_name(true, 0, 0, stream, lanes).c_str())
                       : -1;
  double bytes = 0.0;
  for (int k = 1; k <= m; k++) {
    float *vkp1 = b[nb];
    nb = (nb + 1) % 3;
    if (stream)
      cheb_step_stream<float>(A, vk, vkm1, vkp1, y32, cc, k == 1 ? 1.0 / dd : 2.0 / dd, mu[0],
                              mu[k], k == 1, s);
    else
      cheb_step_vec<float>(A, vk, vkm1, vkp1, y32, cc, k == 1 ? 1.0 / dd : 2.0 / dd, mu[0],
                           mu[k], k == 1, lanes, s);
    vkm1 = vk;
    vk = vkp1;
    bytes += (double)filter_launch_bytes(A, 4, 1, k == 1 ? 3 : 5, stream ? index_bytes(A) : 4);
  }
  if (tok >= 0) c.kt.end(s, tok, m, bytes);
  c.kstats.filter_bytes += bytes;
  c.kstats.filter_steps += m;
  c.kstats.filter_launches += m;
}

// y = p((A - cc)/dd) v for one FP64 column in HBM (SRC/chebpol.c:509-574).
void cheb_filter_col(Ctx &c, const pevsl_Polparams *pol, const double *dv,
                     double *dy, double *work) {
  const int n = c.A ? c.A->n : (int)c.h_x.size();
  const int m = pol->deg;
  const double *mu = pol->mu;
  const double cc = pol->cc, dd = pol->dd;
  const hipStream_t s = c.stream;
  if (work && !c.A) fatal("cheb_filter_col: caller work vectors need a device matrix",
                          __FILE__, __LINE__);
  if (m == 0 || !c.A) {
    // operator given as a host callback (or trivial degree): unfused steps;
    // with a device matrix this branch only runs at degree 0 (one vector)
    double *w = work ? work : c.scratch64.ensure((size_t)4 * n);
    double *vk = w, *vkp1 = w + n, *vkm1 = w + 2 * (size_t)n;
    PV_HIP(hipMemcpyAsync(vk, dv, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    PV_HIP(hipMemcpyAsync(dy, vk, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    scale_vec(n, dy, nullptr, mu[0], s);
    for (int k = 1; k <= m; k++) {
      const double t = k == 1 ? 1.0 / dd : 2.0 / dd;
      apply_A(c, vk, vkp1);
      axpy_vec(n, -cc, vk, vkp1, s);
      scale_vec(n, vkp1, nullptr, t, s);
      if (k > 1) axpy_vec(n, -1.0, vkm1, vkp1, s);
      axpy_vec(n, mu[k], vkp1, dy, s);
      std::swap(vkm1, vk);
      std::swap(vk, vkp1);
    }
    c.kstats.filter_steps += m;
    return;
  }
  const DevCsr &A = *c.A;
  const int lanes = c.lanes ? c.lanes : auto_lanes((double)A.nnz / std::max(1, n));
  // default (-1): the CSR-stream kernel (5.7 vs 4.4 TB/s for the best
  // vector-CSR lane count on the 7-point stencil)
  const bool stream = c.use_stream != 0;
  if (c.filter_prec == PEVSL_PREC_FP64) {
    double *w = work ? work : c.scratch64.ensure((size_t)3 * n);
    double *b[3] = {w, w + n, w + 2 * (size_t)n};
    const double *vk = dv, *vkm1 = nullptr;
    int nb = 0;
    const long tok =
        work ? -1 : c.kt.begin(s, filter_kernel_name(false, 0, 0, stream, lanes).c_str());
    const double bytes0 = c.kstats.filter_bytes;
    for (int k = 1; k <= m; k++) {
      double *vkp1 = b[nb];
      nb = (nb + 1) % 3;
      if (stream)
        cheb_step_stream<double>(A, vk, vkm1, vkp1, dy, cc, k == 1 ? 1.0 / dd : 2.0 / dd,
                                 mu[0], mu[k], k == 1, s);
      else
        cheb_step_vec<double>(A, vk, vkm1, vkp1, dy, cc,
                              k == 1 ? 1.0 / dd : 2.0 / dd, mu[0], mu[k], k == 1,
                              lanes, s);
      vkm1 = vk;
      vk = vkp1;
      c.kstats.filter_bytes +=
          (double)filter_launch_bytes(A, 8, 1, k == 1 ? 3 : 5, stream ? index_bytes(A) : 4);
    }
    if (tok >= 0) c.kt.end(s, tok, m, c.kstats.filter_bytes - bytes0);
  } else {
    // x32, y32 and three work columns: 5n floats (within 3n doubles)
    float *w = reinterpret_cast<float *>(work ? work
                                              : c.scratch64.ensure((size_t)5 * n / 2 + 8));
    float *x32 = w, *y32 = w + n;
    copy_f64_to_f32((size_t)n, dv, x32, s);
    cheb_filter_f32(c, pol, x32, y32, w + 2 * (size_t)n, work == nullptr);
    copy_f32_to_f64((size_t)n, y32, dy, s);
    return;
  }
  c.kstats.filter_steps += m;
  c.kstats.filter_launches += m;
}

void cheb_filter_col32(Ctx &c, const pevsl_Polparams *pol, const float *dv, float *dy) {
  const int n = c.A->n;
  float *w = reinterpret_cast<float *>(c.scratch64.ensure((size_t)3 * n / 2 + 8));
  cheb_filter_f32(c, pol, dv, dy, w);
}

namespace {

// C (m1 x m2, ldc = m1) = A^T B for tall column-major A (n x m1) and B
// (n x m2).  A single DGEMM with the long n as its inner dimension leaves most
// of the chip idle at m1, m2 ~ 1e3; the rows are split into kGramParts
// chunks computed as one strided-batched DGEMM (MFMA), whose partial
// products are then summed in a fixed order.
constexpr int kGramParts = 64;
// upper (m1 == m2): the caller reads only the upper triangle of C (the
// Rayleigh-Ritz pencil, solved with rocblas_fill_upper).  The MFMA kernel then
// forms only the tiles on and above the diagonal and returns C symmetric.
void gram_tn(rocblas_handle h, hipStream_t st, int n, int m1, const double *A, int lda, int m2,
             const double *B, int ldb, double *C, DBuf<double> &work, int kernel,
             bool upper = false) {
  const double one = 1.0, zero = 0.0;
  // default (-1): the library's MFMA kernel for the upper triangle, and for
  // the full product where the row split below would fall back to one
  // unsplit DGEMM (n < 64 x 1024: 6.4 ms against 0.11 ms at n = 40,000,
  // m = 130); rocBLAS for the full product above (at n = 1e6, m = 800:
  // 23.1 ms against 29.0 ms)
  upper = upper && m1 == m2;
  if (kernel < 0) kernel = (upper || n < kGramParts * 1024) ? 1 : 0;
  if (kernel == 1) {
    // hand-written FP64 MFMA tiles over at most kGramParts row parts of >= 1024 rows
    const int chunk = std::max(1024, (n + kGramParts - 1) / kGramParts);
    const int parts = (n + chunk - 1) / chunk;
    const size_t blk = (size_t)m1 * m2;
    double *wk = work.ensure(blk * parts);
    gram_tn_mfma(n, m1, A, (size_t)lda, m2, B, (size_t)ldb, chunk, parts, upper, wk, st);
    if (upper)
      sum_parts_upper(wk, parts, m1, C, st);
    else
      sum_parts(wk, parts, blk, C, st);
    return;
  }
  const int chunk = n / kGramParts;
  if (chunk < 1024) {
    PV_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, m1, m2, n,
                          &one, A, lda, B, ldb, &zero, C, m1));
    return;
  }
  const int rem = n - kGramParts * chunk;
  const int parts = kGramParts + (rem > 0 ? 1 : 0);
  const size_t blk = (size_t)m1 * m2;
  double *wk = work.ensure(blk * parts);
  PV_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none,
                                        m1, m2, chunk, &one, A, lda, chunk, B, ldb, chunk,
                                        &zero, wk, m1, (rocblas_stride)blk, kGramParts));
  if (rem > 0)
    PV_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, m1, m2, rem,
                          &one, A + (size_t)kGramParts * chunk, lda,
                          B + (size_t)kGramParts * chunk, ldb, &zero, wk + blk * kGramParts,
                          m1));
  sum_parts(wk, parts, blk, C, st);
}

// One filter group replayed from a captured hipGraph (pEVSL_SetGraphFilter):
// the group keeps its own coefficients and column pointers on the device, so
// that the graphs of the groups of one step replay one after the other, and
// the graph holds kernel launches only.
struct GroupGraph {
  std::vector<int> key;  // slices, width, kernel
  DBuf<double> mus;
  DBuf<void *> colptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  char *blocks = nullptr;  // filter-block allocation the graph was captured on
  double bytes = 0.0;      // algorithmic bytes of one replay
  long long steps = 0, launches = 0;
  std::string kname;
  bool used = false;
  GroupGraph() = default;
  GroupGraph(const GroupGraph &) = delete;
  GroupGraph &operator=(const GroupGraph &) = delete;
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~GroupGraph() { drop(); }
};

// Raised when a basis chunk cannot be allocated: the batch does not fit in
// device memory (pEVSL_ChebLanNrBatch returns PEVSL_ERR_DEVICE_OOM).
struct DeviceOom {
  size_t bytes;
};

class BatchSolver {
 public:
  BatchSolver(Ctx &c, pevsl_Data *p, int ns) : c_(c), p_(p), ns_(ns), S_(ns) {}
  // releases the bases and staging buffers a run left (an interrupted run)
  ~BatchSolver() {
    (void)hipStreamSynchronize(c_.stream);
    for (Slice &sl : S_)
      for (void *q : sl.chunks) (void)hipFree(q);
    if (pinned_) (void)hipHostFree(pinned_);
    if (pinned_i_) (void)hipHostFree(pinned_i_);
  }
  BatchSolver(const BatchSolver &) = delete;
  BatchSolver &operator=(const BatchSolver &) = delete;

  int run(double *intvs, int *maxits, double tol, pevsl_Parvec *vinit,
          pevsl_Polparams *pols, FILE *fstats);
  int bench(double *intvs, pevsl_Polparams *pols, int k, int warmup, int reps,
            int graph_reps, pevsl_StepBench *out);
  void outputs(int *nevOut, double **lamo, pevsl_Parvecs **Wo, double **reso,
               pevsl_Parvec *vinit);

 private:
  void *col(int s, int j) {
    const int sh = L_.chunk_shift;
    return static_cast<char *>(S_[(size_t)s].chunks[(size_t)(j >> sh)]) +
           (size_t)(j & ((1 << sh) - 1)) * L_.ld * vb_;
  }
  double *cold(int s, int j) { return static_cast<double *>(col(s, j)); }
  float *colf(int s, int j) { return static_cast<float *>(col(s, j)); }
  // FP64 view of basis chunk ci of slice s (kc columns): the chunk itself, or
  // its widened copy in conv_ when the basis is stored in FP32
  const double *chunk64(int s, int ci, int kc);
  // residual threshold of slice s: tol, or on the FP32-filter path
  // max(tol, mixed_tol * rho) with rho = max(|lmin|, |lmax|) of the frame
  double accept_tol(int s, double tol) const {
    if (c_.filter_prec != PEVSL_PREC_FP32) return tol;
    const pevsl_Polparams *pol = S_[(size_t)s].pol;
    const double rho = std::max(std::fabs(pol->cc - pol->dd), std::fabs(pol->cc + pol->dd));
    return std::max(tol, c_.mixed_tol * rho);
  }
  // z = p(A) x between basis columns / an FP64 input, in the basis' type
  void filter_into(int s, const double *x64, int kin, int kout);
  void ensure_column(int s, int j);
  void sync_vtab();
  void set_active(const std::vector<int> &act);
  void filter_step(const std::vector<int> &act, int k);
  void filter_block(const std::vector<int> &grp, int k);
  int group_width(const std::vector<int> &grp, int &kind) const;
  char *blocks_for(int B);
  void stage_mus(const std::vector<int> &grp, int B, DBuf<double> &dst);
  void stage_colptr(const std::vector<int> &grp, int B, int k, DBuf<void *> &dst);
  void enqueue_block(const std::vector<int> &grp, int B, int kind, const double *musp,
                     void *const *colp);
  void filter_group_graph(const std::vector<int> &grp, int k);
  void lanczos_update(const std::vector<int> &act, int k);
  void cgs(const std::vector<int> &act, int k);
  void breakdown(int s, int k);
  void convergence_tests(const std::vector<int> &ts, int k);
  void ritz(int s, double tol);
  void ritz_rr(int s, double tol);
  void free_basis(int s);
  int prepare(double *intvs, int *maxits, pevsl_Polparams *pols, double tall);

  // the filter groups of one step: at most the tuned width, one interval map
  std::vector<std::vector<int>> filter_groups(const std::vector<int> &act) const;
  // block SpMM kernel of the batched filter: 0 lane per (row, column), 1
  // row per thread, 2 its Clenshaw form, 3 / 4 the nnz-balanced forms of 1 /
  // 2, 5 the segmented Clenshaw form (FP32).  Default: 2 on the FP32 path,
  // the reference's three-term recurrence (1) on FP64, the balanced forms
  // when some row is long.
  int block_kind() const {
    const bool f32 = c_.filter_prec == PEVSL_PREC_FP32;
    int kind = c_.blk_rowkernel;
    if (kind < 0) kind = (f32 ? 2 : 1) + (c_.A->max_row > kUnevenRow ? 2 : 0);
    if (kind == 5 && !f32) kind = 1;  // Clenshaw forms: FP32 only
    return kind;
  }

  Ctx &c_;
  pevsl_Data *p_;
  int ns_;
  bool capture_ = false;  // filter_block enqueues kernels only (graph capture)
  int n_ = 0;
  bool b32_ = false;  // FP32 basis storage
  size_t vb_ = 8;     // basis element bytes
  std::vector<Slice> S_;
  LanBatch L_;
  DBuf<void *> vtab_;
  std::vector<void *> vtab_h_;
  DBuf<double> conv_;   // FP64 copy of one FP32 basis chunk (Ritz stage)
  DBuf<float> f32tmp_;  // FP32 staging of an FP64 input vector
  DBuf<double> scal_;  // alpha, beta, nrm2a, nrm2b
  DBuf<int> flags_;    // pass2, active
  DBuf<double> h_, part_;
  DBuf<double> blk_;   // filter blocks
  DBuf<double> mus_;
  DBuf<void *> colptr_;
  DBuf<double> tri_;   // tridiagonal work
  std::vector<double> tri_h_;
  double *pinned_ = nullptr;
  int *pinned_i_ = nullptr;
  std::vector<int> mus_key_;
  std::vector<double> mus_h_, betas_h_;
  std::vector<void *> cp_h_;
  std::vector<std::unique_ptr<GroupGraph>> graphs_;  // captured filter groups
};

const double *BatchSolver::chunk64(int s, int ci, int kc) {
  void *p = S_[(size_t)s].chunks[(size_t)ci];
  if (!b32_) return static_cast<const double *>(p);
  double *d = conv_.ensure(L_.ld * ((size_t)1 << L_.chunk_shift));
  copy_f32_to_f64(L_.ld * (size_t)kc, static_cast<const float *>(p), d, c_.stream);
  return d;
}

void BatchSolver::filter_into(int s, const double *x64, int kin, int kout) {
  const pevsl_Polparams *pol = S_[(size_t)s].pol;
  if (!b32_) {
    cheb_filter_col(c_, pol, x64 ? x64 : cold(s, kin), cold(s, kout));
    return;
  }
  const float *x = kin >= 0 ? colf(s, kin) : nullptr;
  if (x64) {
    f32tmp_.ensure((size_t)n_);
    copy_f64_to_f32((size_t)n_, x64, f32tmp_.p, c_.stream);
    x = f32tmp_.p;
  }
  cheb_filter_col32(c_, pol, x, colf(s, kout));
}

void BatchSolver::ensure_column(int s, int j) {
  Slice &sl = S_[(size_t)s];
  const size_t need = (size_t)(j >> L_.chunk_shift) + 1;
  bool grew = false;
  while (sl.chunks.size() < need) {
    void *p = nullptr;
    const size_t bytes = vb_ * L_.ld * ((size_t)1 << L_.chunk_shift);
    const hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();  // clear the error state for the caller
      throw DeviceOom{bytes};
    }
    PV_HIP(e);
    // rows [n, ld) of every column stay zero: the reorthogonalization
    // kernels stream whole row pairs up to ld
    PV_HIP(hipMemsetAsync(p, 0, bytes, c_.stream));
    sl.chunks.push_back(p);
    vtab_h_[(size_t)s * L_.maxchunks + sl.chunks.size() - 1] = p;
    grew = true;
  }
  if (grew) sync_vtab();
}

void BatchSolver::sync_vtab() {
  PV_HIP(hipMemcpyAsync(vtab_.p, vtab_h_.data(), sizeof(void *) * vtab_h_.size(),
                        hipMemcpyHostToDevice, c_.stream));
  PV_HIP(hipStreamSynchronize(c_.stream));
}

void BatchSolver::set_active(const std::vector<int> &act) {
  // the staging buffer may still feed an earlier copy
  PV_HIP(hipStreamSynchronize(c_.stream));
  for (int s = 0; s < ns_; s++) pinned_i_[s] = 0;
  for (int s : act) pinned_i_[s] = 1;
  PV_HIP(hipMemcpyAsync(L_.active, pinned_i_, sizeof(int) * ns_,
                        hipMemcpyHostToDevice, c_.stream));
}

void BatchSolver::free_basis(int s) {
  PV_HIP(hipStreamSynchronize(c_.stream));
  for (size_t i = 0; i < S_[(size_t)s].chunks.size(); i++) {
    PV_HIP(hipFree(S_[(size_t)s].chunks[i]));
    vtab_h_[(size_t)s * L_.maxchunks + i] = nullptr;
  }
  S_[(size_t)s].chunks.clear();
}

// Block filter for a group of at most B slices (sorted by degree, highest
// first): z_{k+1} = p_s(A) z_k for every slice s of the group.
// Width and kernel of the block that filters a group of g slices: the
// narrowest width that holds the group (groups never exceed the tuned width:
// a width-8 block holding 4 columns would stream twice the vector bytes it
// needs); the segmented kernel takes any multiple of its 16-byte piece.
int BatchSolver::group_width(const std::vector<int> &grp, int &kind) const {
  const int g = (int)grp.size();
  const int seg = c_.filter_prec == PEVSL_PREC_FP32 ? 4 : 2;
  kind = block_kind();
  // Clenshaw sums lose accuracy with the degree faster than the three-term
  // recurrence: on the n = 2e6 power-law matrix at degree 508 the FP32
  // Clenshaw filter left accepted residuals up to 8.6e-5 (threshold 1.1e-4)
  // and missed one of 32 eigenvalues; the three-term filter, 1.1e-5 and all
  // 32.  Above kClenMaxDeg a group runs the three-term form of its kernel.
  if (kind == 2 || kind == 4 || kind == 5) {
    int degmax = 0;
    for (int s : grp) degmax = std::max(degmax, S_[(size_t)s].pol->deg);
    if (degmax > kClenMaxDeg) kind = (kind == 4 || (kind == 5 && c_.A->max_row > kUnevenRow)) ? 3 : 1;
  }
  int B = 1;
  if (kind == 5)
    B = (g + seg - 1) / seg * seg;
  else
    while (B < g) B *= 2;
  if (kind != 5 && (B > 32 || c_.A->nrowblk == 0)) kind = 0;
  return B;
}

// Interleaved filter blocks (x, y and two work blocks) for width B; returns
// the base of the allocation (it moves when a wider block needs more room).
char *BatchSolver::blocks_for(int B) {
  const size_t elem = c_.filter_prec == PEVSL_PREC_FP32 ? 4 : 8;
  const size_t blk_elems = (size_t)n_ * B;
  return reinterpret_cast<char *>(blk_.ensure((4 * blk_elems * elem) / 8 + 8));
}

// Per-column coefficients [(degmax+1) x B] of a group into dst (host staging
// rewritten only after the stream has drained the previous copy).
void BatchSolver::stage_mus(const std::vector<int> &grp, int B, DBuf<double> &dst) {
  int degmax = 0;
  for (int s : grp) degmax = std::max(degmax, S_[(size_t)s].pol->deg);
  PV_HIP(hipStreamSynchronize(c_.stream));
  std::vector<double> &mh = mus_h_;
  mh.assign((size_t)(degmax + 1) * B, 0.0);
  for (size_t ci = 0; ci < grp.size(); ci++) {
    const pevsl_Polparams *pol = S_[(size_t)grp[ci]].pol;
    for (int j = 0; j <= pol->deg; j++) mh[(size_t)j * B + ci] = pol->mu[j];
  }
  dst.ensure(mh.size());
  PV_HIP(hipMemcpyAsync(dst.p, mh.data(), sizeof(double) * mh.size(), hipMemcpyHostToDevice,
                        c_.stream));
}

// Column pointers of a group at step k: inputs z_k, outputs z_{k+1}.
void BatchSolver::stage_colptr(const std::vector<int> &grp, int B, int k, DBuf<void *> &dst) {
  PV_HIP(hipStreamSynchronize(c_.stream));
  std::vector<void *> &cp = cp_h_;
  cp.assign((size_t)2 * B, nullptr);
  for (size_t ci = 0; ci < grp.size(); ci++) {
    cp[ci] = col(grp[ci], k);
    cp[(size_t)B + ci] = col(grp[ci], k + 1);
  }
  dst.ensure(cp.size());
  PV_HIP(hipMemcpyAsync(dst.p, cp.data(), sizeof(void *) * cp.size(), hipMemcpyHostToDevice,
                        c_.stream));
}

void BatchSolver::filter_block(const std::vector<int> &grp, int k) {
  int kind;
  const int B = group_width(grp, kind);
  std::vector<int> key(grp);
  key.push_back(B);
  if (key != mus_key_) {
    if (capture_) fatal("filter_block: coefficients not staged before capture", __FILE__,
                        __LINE__);
    stage_mus(grp, B, mus_);
    mus_key_ = key;
  }
  // under graph capture the pointers of this group and step are already on
  // the device
  if (!capture_) stage_colptr(grp, B, k, colptr_);
  enqueue_block(grp, B, kind, mus_.p, colptr_.p);
}

// The kernels of one filter application to a group (pack, the degree loop,
// unpack), with coefficients musp and column pointers colp on the device:
// nothing but kernel launches, so that it can be captured in a hipGraph.
void BatchSolver::enqueue_block(const std::vector<int> &grp, int B, int kind,
                                const double *musp, void *const *colp) {
  const DevCsr &A = *c_.A;
  const int g = (int)grp.size();
  const bool f32 = c_.filter_prec == PEVSL_PREC_FP32;
  const size_t elem = f32 ? 4 : 8;
  const int seg = 16 / (int)elem;
  const size_t blk_elems = (size_t)n_ * B;
  char *base = blocks_for(B);
  void *xb[4];
  for (int i = 0; i < 4; i++) xb[i] = base + (size_t)i * blk_elems * elem;
  int degmax = 0;
  for (int s : grp) degmax = std::max(degmax, S_[(size_t)s].pol->deg);
  const double cc = S_[(size_t)grp[0]].pol->cc, dd = S_[(size_t)grp[0]].pol->dd;
  std::vector<int> degs;
  for (int s : grp) degs.push_back(S_[(size_t)s].pol->deg);
  const bool rowk = kind != 0;
  const bool segk = kind == 5;
  const bool clen = kind == 2 || kind == 4;
  const bool bal = kind == 3 || kind == 4;
  auto run = [&](auto *tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    T *x = (T *)xb[0], *y = (T *)xb[1];
    T *b[3] = {(T *)xb[2], (T *)xb[3], nullptr};
    if constexpr (std::is_same<T, float>::value) {
      if (b32_)
        pack_cols<float, float>(n_, B, g, reinterpret_cast<const float *const *>(colp), x,
                                c_.stream);
      else
        pack_cols<float, double>(n_, B, g, reinterpret_cast<const double *const *>(colp),
                                 x, c_.stream);
    } else {
      pack_cols<T, double>(n_, B, g, reinterpret_cast<const double *const *>(colp), x,
                           c_.stream);
    }
    T *yout = y;
    const std::string kname = filter_kernel_name(f32, kind, B, false, 0);
    if (segk) {
      // pieces of `seg` columns, each active from its own top degree down
      // (columns are sorted by degree, so a piece's first column has its top)
      const int np = B / seg;
      SegTops tops{};
      for (int p = 0; p < np; p++) tops.top[p] = p * seg < g ? degs[(size_t)p * seg] : -1;
      T *ring[3] = {y, b[0], b[1]};
      const T *b1 = nullptr, *b2 = nullptr;
      int nr = 0;
      long tok = -1;
      double kbytes = 0.0;
      for (int d = degmax; d >= 0; d--) {
        T *out = ring[nr];
        if ((const T *)out == b1 || (const T *)out == b2) {
          nr = (nr + 1) % 3;
          out = ring[nr];
        }
        nr = (nr + 1) % 3;
        int nseg = 0;  // pieces whose series has reached degree d
        while (nseg < np && tops.top[nseg] >= d) nseg++;
        if (tok < 0 && !capture_) tok = c_.kt.begin(c_.stream, kname.c_str());
        clen_step_seg<T>(A, B, b1, b2, x, out, cc, d == 0 ? 1.0 / dd : 2.0 / dd, musp, d, nseg,
                         tops, c_.stream);
        const double lb = clen_seg_launch_bytes(A, (int)elem, seg, d, nseg, tops);
        kbytes += lb;
        c_.kstats.filter_bytes += lb;
        int nprod = 0;  // columns that form a product at this degree
        while (nprod < g && degs[(size_t)nprod] > d) nprod++;
        c_.kstats.filter_steps += nprod;
        c_.kstats.filter_launches++;
        b2 = b1;
        b1 = out;
        yout = out;
      }
      if (tok >= 0) c_.kt.end(c_.stream, tok, degmax + 1, kbytes);
    } else if (clen && degmax >= 1) {
      // Clenshaw sums from the top degree down: b_d = mu_d x + (2/dd)(A - cc)
      // b_{d+1} - b_{d+2}; the d = 0 step (1/dd) is the filtered vector.
      // Three rotating buffers; x (the input block) is read by every step.
      // The top term (no product) is a column scaling; the degmax product
      // steps after it are timed as one run of the block kernel.
      T *ring[3] = {y, b[0], b[1]};
      const T *b1 = nullptr, *b2 = nullptr;
      int nr = 0;
      long tok = -1;
      double kbytes = 0.0;
      for (int d = degmax; d >= 0; d--) {
        T *out = ring[nr];
        if ((const T *)out == b1 || (const T *)out == b2) {
          nr = (nr + 1) % 3;
          out = ring[nr];
        }
        nr = (nr + 1) % 3;
        if (b1 && tok < 0 && !capture_) tok = c_.kt.begin(c_.stream, kname.c_str());
        if (bal)
          clen_step_bal<T>(A, B, b1, b2, x, out, cc, d == 0 ? 1.0 / dd : 2.0 / dd, musp, d,
                           c_.stream);
        else
          clen_step_blkrow<T>(A, B, b1, b2, x, out, cc, d == 0 ? 1.0 / dd : 2.0 / dd, musp,
                              d, c_.stream);
        int nact = 0;  // columns whose series has reached degree d + 1
        while (nact < g && degs[(size_t)nact] >= d + 1) nact++;
        if (b1) {
          const double lb = (double)clen_launch_bytes(A, (int)elem, nact, b2 != nullptr,
                                                      bal ? 4 : index_bytes(A));
          kbytes += lb;
          c_.kstats.filter_bytes += lb;
          c_.kstats.filter_steps += nact;
          c_.kstats.filter_launches++;
        }
        b2 = b1;
        b1 = out;
        yout = out;
      }
      if (tok >= 0) c_.kt.end(c_.stream, tok, degmax, kbytes);
    } else {
    // third work buffer: reuse x once it is no longer read (after step 2)
    const T *vk = x, *vkm1 = nullptr;
    T *bufs[3] = {b[0], b[1], x};
    int nb = 0;
    const long tok = degmax > 0 && !capture_ ? c_.kt.begin(c_.stream, kname.c_str()) : -1;
    double kbytes = 0.0;
    int nprev = g;  // columns whose series reaches degree d - 1
    for (int d = 1; d <= degmax; d++) {
      int nact = 0;
      while (nact < g && degs[(size_t)nact] >= d) nact++;
      T *vkp1 = bufs[nb];
      // never overwrite the buffers read in this step
      if ((const T *)vkp1 == vk || (const T *)vkp1 == vkm1) {
        nb = (nb + 1) % 3;
        vkp1 = bufs[nb];
      }
      nb = (nb + 1) % 3;
      // the row kernels update y every other degree (skip, then pair: the
      // same additions in the same order, one pass over y per two degrees);
      // the last degree updates it alone when it would be skipped
      int ymode = kYPlain, ncol = nact, passes = d == 1 ? 3 : 5;
      if (rowk && d > 1) {
        if (d % 2 == 1) {
          ymode = kYPair;
          ncol = nprev;
        } else if (d < degmax) {
          ymode = kYSkip;
          passes = 3;
        }
      }
      if (bal)
        cheb_step_bal<T>(A, B, vk, vkm1, vkp1, y, cc, d == 1 ? 1.0 / dd : 2.0 / dd, musp, d,
                         ncol, c_.stream, ymode);
      else if (rowk)
        cheb_step_blkrow<T>(A, B, vk, vkm1, vkp1, y, cc, d == 1 ? 1.0 / dd : 2.0 / dd,
                            musp, d, ncol, c_.stream, ymode);
      else
        cheb_step_blk<T>(A, B, vk, vkm1, vkp1, y, cc, d == 1 ? 1.0 / dd : 2.0 / dd,
                         musp, d, nact, c_.stream);
      nprev = nact;
      // pair steps count y over the columns of degree d - 1 (a product of
      // its own over the columns of degree d); skip steps touch no y
      const int ib = rowk && !bal ? index_bytes(A) : 4;  // only the row kernel reads jo16
      const double lb =
          ymode == kYPair
              ? (double)filter_launch_bytes(A, (int)elem, nact, 3, ib) +
                    2.0 * (double)n_ * (double)elem * ncol
              : (double)filter_launch_bytes(A, (int)elem, nact, passes, ib);
      kbytes += lb;
      c_.kstats.filter_bytes += lb;
      c_.kstats.filter_steps += nact;
      c_.kstats.filter_launches++;
      vkm1 = vk;
      vk = vkp1;
    }
    if (tok >= 0) c_.kt.end(c_.stream, tok, degmax, kbytes);
    }
    if constexpr (std::is_same<T, float>::value) {
      if (b32_)
        unpack_cols<float, float>(n_, B, g, yout,
                                  reinterpret_cast<float *const *>(colp + B), c_.stream);
      else
        unpack_cols<float, double>(n_, B, g, yout,
                                   reinterpret_cast<double *const *>(colp + B), c_.stream);
    } else {
      unpack_cols<T, double>(n_, B, g, yout, reinterpret_cast<double *const *>(colp + B),
                             c_.stream);
    }
  };
  if (f32)
    run((float *)nullptr);
  else
    run((double *)nullptr);
}

std::vector<std::vector<int>> BatchSolver::filter_groups(const std::vector<int> &act) const {
  // groups of at most gmax slices that share the filter's interval map
  // (cc, dd: FindPol sets them from each slice's own [lmin, lmax],
  // SRC/chebpol.c:386-389, and a block step applies one map to all its
  // columns), each sorted by degree (descending)
  std::vector<int> order(act);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const pevsl_Polparams *pa = S_[(size_t)a].pol, *pb = S_[(size_t)b].pol;
    if (pa->cc != pb->cc) return pa->cc < pb->cc;
    if (pa->dd != pb->dd) return pa->dd < pb->dd;
    return pa->deg > pb->deg;
  });
  // group width: the tuned one, else the widest block of the segmented
  // kernel (one pass over the matrix for all columns), else 8 (FP32) / 4
  // (FP64), which pEVSL_Autotune picks for the row kernels on the 7-point
  // stencil (wider row gathers cost more per column)
  const int kind = block_kind();
  const int gmax = c_.block_width ? c_.block_width
                   : kind == 5    ? kMaxSegs * 16 / (c_.filter_prec == PEVSL_PREC_FP32 ? 4 : 8)
                   : (c_.filter_prec == PEVSL_PREC_FP32 ? 8 : 4);
  std::vector<std::vector<int>> out;
  for (size_t i = 0; i < order.size();) {
    const pevsl_Polparams *p0 = S_[(size_t)order[i]].pol;
    size_t j = i + 1;
    while (j < order.size() && j - i < (size_t)gmax && S_[(size_t)order[j]].pol->cc == p0->cc &&
           S_[(size_t)order[j]].pol->dd == p0->dd)
      j++;
    out.emplace_back(order.begin() + (long)i, order.begin() + (long)j);
    i = j;
  }
  return out;
}

void BatchSolver::filter_step(const std::vector<int> &act, int k) {
  for (int s : act) ensure_column(s, k + 1);
  const bool single = act.size() == 1 || !c_.A || c_.block_width == 1;
  if (single) {
    for (int s : act) {
      filter_into(s, nullptr, k, k + 1);
      S_[(size_t)s].nmatvec += S_[(size_t)s].pol->deg;
    }
    return;
  }
  const std::vector<std::vector<int>> groups = filter_groups(act);
  const bool graphs = c_.graph_filter != 0;
  if (graphs) {
    // the blocks sized for the widest group first, so that no allocation
    // moves them between a capture and its replays within the step
    int bmax = 1;
    for (const std::vector<int> &grp : groups) {
      int kind;
      if (grp.size() > 1) bmax = std::max(bmax, group_width(grp, kind));
    }
    blocks_for(bmax);
    for (auto &gg : graphs_) gg->used = false;
  }
  for (const std::vector<int> &grp : groups) {
    if (grp.size() == 1) {
      filter_into(grp[0], nullptr, k, k + 1);
    } else if (graphs) {
      filter_group_graph(grp, k);
    } else {
      filter_block(grp, k);
    }
    for (int s : grp) S_[(size_t)s].nmatvec += S_[(size_t)s].pol->deg;
  }
  if (graphs) {
    // groups only shrink as slices finish: a graph unused this step is not
    // replayed again
    graphs_.erase(std::remove_if(graphs_.begin(), graphs_.end(),
                                 [](const std::unique_ptr<GroupGraph> &g) { return !g->used; }),
                  graphs_.end());
  }
}

// Filter a group by replaying its captured hipGraph; the first use of a
// group (or of a new filter-block allocation) captures the kernels of one
// application.  Per step only the column pointers are restaged.
void BatchSolver::filter_group_graph(const std::vector<int> &grp, int k) {
  int kind;
  const int B = group_width(grp, kind);
  std::vector<int> key(grp);
  key.push_back(B);
  key.push_back(kind);
  GroupGraph *gg = nullptr;
  for (auto &g : graphs_)
    if (g->key == key) gg = g.get();
  if (!gg) {
    graphs_.push_back(std::unique_ptr<GroupGraph>(new GroupGraph()));
    gg = graphs_.back().get();
    gg->key = key;
    gg->kname = filter_kernel_name(c_.filter_prec == PEVSL_PREC_FP32, kind, B, false, 0);
    stage_mus(grp, B, gg->mus);
  }
  gg->used = true;
  stage_colptr(grp, B, k, gg->colptr);
  char *base = blocks_for(B);
  if (gg->exec && gg->blocks != base) gg->drop();
  if (!gg->exec) {
    const pevsl_KernelStats ks = c_.kstats;
    capture_ = true;
    PV_HIP(hipStreamBeginCapture(c_.stream, hipStreamCaptureModeThreadLocal));
    enqueue_block(grp, B, kind, gg->mus.p, gg->colptr.p);
    PV_HIP(hipStreamEndCapture(c_.stream, &gg->graph));
    capture_ = false;
    PV_HIP(hipGraphInstantiate(&gg->exec, gg->graph, nullptr, nullptr, 0));
    gg->blocks = base;
    // the capture counted one application that has not run
    gg->bytes = c_.kstats.filter_bytes - ks.filter_bytes;
    gg->steps = c_.kstats.filter_steps - ks.filter_steps;
    gg->launches = c_.kstats.filter_launches - ks.filter_launches;
    c_.kstats = ks;
  }
  const long tok = c_.kt.begin(c_.stream, gg->kname.c_str());
  PV_HIP(hipGraphLaunch(gg->exec, c_.stream));
  c_.kt.end(c_.stream, tok, gg->launches, gg->bytes);
  c_.kstats.filter_bytes += gg->bytes;
  c_.kstats.filter_steps += gg->steps;
  c_.kstats.filter_launches += gg->launches;
}

// three-term update: alpha and |z_{k+1}|^2 before reorthogonalization
void BatchSolver::lanczos_update(const std::vector<int> &act, int k) {
  set_active(act);
  lan_alpha(L_, k, c_.stream);
  lan_subtract_alpha(L_, k, c_.stream);
  const double vecb = (double)vb_ * n_;
  c_.kstats.reorth_bytes += act.size() * (4 * vecb + 3 * vecb);
}

// classical Gram-Schmidt with the DGKS test: pass 0 always, pass 1 only
// for slices whose norm dropped below 1/sqrt(2) (decided on the device)
void BatchSolver::cgs(const std::vector<int> &act, int k) {
  const double vecb = (double)vb_ * n_;
  const double pass = 2.0 * (k + 1) * vecb + 3 * vecb;
  const long tok = c_.kt.begin(c_.stream, b32_ ? "cgs_pass<float>" : "cgs_pass<double>");
  lan_cgs_pass(L_, k, 0, c_.stream);
  c_.kt.end(c_.stream, tok, 1, act.size() * pass);
  lan_cgs_pass(L_, k, 1, c_.stream);
  c_.kstats.reorth_bytes += act.size() * pass;
  c_.kstats.cgs_bytes += act.size() * pass;
  c_.kstats.cgs_passes += (long long)act.size();
}

// Lucky breakdown (SRC/cheblanNr.c:248-281): replace z_{k+1} by a filtered
// random vector orthogonalized against the basis; beta = 0.
void BatchSolver::breakdown(int s, int k) {
  Slice &sl = S_[(size_t)s];
  sl.nbreak++;
  double *vr = c_.cb_x.ensure((size_t)n_);
  rand_vectors(c_, n_, 1, vr, (size_t)n_);
  filter_into(s, vr, -1, k + 1);
  sl.nmatvec += sl.pol->deg;
  std::vector<int> one{s};
  set_active(one);
  // nrm2a of the fresh vector, then the DGKS passes
  PV_HIP(hipMemsetAsync(L_.alpha + s, 0, sizeof(double), c_.stream));
  lan_subtract_alpha(L_, k, c_.stream);  // alpha = 0: only |v|^2
  lan_cgs_pass(L_, k, 0, c_.stream);
  lan_cgs_pass(L_, k, 1, c_.stream);
  PV_HIP(hipMemcpyAsync(pinned_, L_.nrm2b + s, sizeof(double), hipMemcpyDeviceToHost,
                        c_.stream));
  PV_HIP(hipStreamSynchronize(c_.stream));
  const double beta = sqrt(pinned_[0]);
  PV_HIP(hipMemcpyAsync(L_.beta + s, &beta, sizeof(double), hipMemcpyHostToDevice,
                        c_.stream));
  lan_scale_new(L_, k, c_.stream);
  const double zero = 0.0;
  PV_HIP(hipMemcpyAsync(L_.beta + s, &zero, sizeof(double), hipMemcpyHostToDevice,
                        c_.stream));
  PV_HIP(hipStreamSynchronize(c_.stream));
}

// Restricted-trace stopping test (SRC/cheblanNr.c:311-333) for every slice
// in ts at step k: the sum of the Ritz values of p(A) above bar by batched
// Sturm multisection on the device, one launch sequence and one
// synchronization for all slices.
void BatchSolver::convergence_tests(const std::vector<int> &ts, int k) {
  const int np = (int)ts.size();
  if (np == 0) return;
  const int kdim = k + 1;
  const size_t ld = (size_t)kdim;
  double *d = tri_.ensure((size_t)np * (3 * ld + 8) + 16);
  double *e = d + np * ld, *lam = e + np * ld, *ws = lam + np * ld, *out = ws + 4 * (size_t)np,
         *thr = out + 2 * (size_t)np;
  tri_h_.assign((size_t)np * (2 * ld + 1), 0.0);
  for (int p = 0; p < np; p++) {
    const Slice &sl = S_[(size_t)ts[(size_t)p]];
    std::copy(sl.dT.begin(), sl.dT.begin() + kdim, tri_h_.begin() + (long)(p * ld));
    std::copy(sl.eT.begin(), sl.eT.begin() + (kdim - 1),
              tri_h_.begin() + (long)(np * ld + p * ld));
    tri_h_[(size_t)2 * np * ld + p] = sl.bar - 5 * DBL_EPSILON;
  }
  const double t0 = wtime();
  PV_HIP(hipMemcpyAsync(d, tri_h_.data(), sizeof(double) * 2 * np * ld, hipMemcpyHostToDevice,
                        c_.stream));
  PV_HIP(hipMemcpyAsync(thr, tri_h_.data() + 2 * np * ld, sizeof(double) * np,
                        hipMemcpyHostToDevice, c_.stream));
  tridiag_trace_above_batch(np, kdim, ld, d, e, thr, lam, ws, out, c_.stream);
  PV_HIP(hipMemcpyAsync(pinned_, out, sizeof(double) * 2 * np, hipMemcpyDeviceToHost,
                        c_.stream));
  PV_HIP(hipStreamSynchronize(c_.stream));
  p_->stats->t_eig += wtime() - t0;
  c_.kstats.t_eig += wtime() - t0;
  for (int p = 0; p < np; p++) {
    Slice &sl = S_[(size_t)ts[(size_t)p]];
    sl.kdim = kdim;
    sl.tr1 = pinned_[2 * p];
    sl.nconv = (int)pinned_[2 * p + 1];
    if (fabs(sl.tr1 - sl.tr0) < 2e-12 || fabs(sl.tr1) + fabs(sl.tr0) < 1e-10) sl.done = true;
    sl.tr0 = sl.tr1;
    if (k == sl.maxit - 1) sl.done = true;
  }
}

// Ritz pairs of slice s and their acceptance (SRC/cheblanNr.c:339-413).
void BatchSolver::ritz(int s, double tol) {
  Slice &sl = S_[(size_t)s];
  const hipStream_t st = c_.stream;
  const int kdim = sl.kdim;
  const double t0 = wtime();
  // eigenpairs of T_kdim (ascending), divide and conquer on the device
  DBuf<double> D, E, C;
  DBuf<int> info;
  D.ensure((size_t)kdim);
  E.ensure((size_t)std::max(kdim, 1));
  C.ensure((size_t)kdim * kdim);
  info.ensure(1);
  PV_HIP(hipMemcpyAsync(D.p, sl.dT.data(), sizeof(double) * kdim, hipMemcpyHostToDevice, st));
  if (kdim > 1)
    PV_HIP(hipMemcpyAsync(E.p, sl.eT.data(), sizeof(double) * (kdim - 1),
                          hipMemcpyHostToDevice, st));
  PV_BLAS(rocsolver_dstedc(c_.blas, rocblas_evect_tridiagonal, kdim, D.p, E.p, C.p,
                           kdim, info.p));
  std::vector<double> ev((size_t)kdim);
  int hinfo = 0;
  PV_HIP(hipMemcpyAsync(ev.data(), D.p, sizeof(double) * kdim, hipMemcpyDeviceToHost, st));
  PV_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  PV_HIP(hipStreamSynchronize(st));
  if (hinfo) fprintf(stderr, "pevsl_amd: STEDC info %d\n", hinfo);
  p_->stats->t_eig += wtime() - t0;
  c_.kstats.t_eig += wtime() - t0;
  int i0 = kdim;
  while (i0 > 0 && ev[(size_t)i0 - 1] >= sl.bar) i0--;
  const int nsel = kdim - i0;
  const int CH = 256;  // Ritz vectors per batch
  DBuf<double> U, W, part, red;
  const size_t ld = L_.ld;
  U.ensure(ld * (size_t)std::min(CH, std::max(nsel, 1)));
  W.ensure(ld * (size_t)std::min(CH, std::max(nsel, 1)));
  part.ensure((size_t)CH * col_part_blocks(n_) + 8);
  red.ensure((size_t)4 * CH);
  std::vector<double> hv((size_t)CH), fac((size_t)CH), aneg((size_t)CH), tq((size_t)CH),
      rr((size_t)CH);
  const int csz = 1 << L_.chunk_shift;
  const double one = 1.0, zero = 0.0;
  const double acc_tol = accept_tol(s, tol);
  // accepted vectors are copied straight into sl.vecs by asynchronous D2H
  // copies: its storage must not move while they are in flight
  if (c_.keep_vectors) sl.vecs.reserve(sl.vecs.size() + (size_t)nsel * n_);
  for (int cs = 0; cs < nsel && sl.rc == 0; cs += CH) {
    const int nb = std::min(CH, nsel - cs);
    // U = V(:, 0:kdim) * C(:, i0+cs : i0+cs+nb), accumulated over basis chunks
    for (int j0 = 0, ci = 0; j0 < kdim; j0 += csz, ci++) {
      const int kc = std::min(csz, kdim - j0);
      PV_BLAS(rocblas_dgemm(c_.blas, rocblas_operation_none, rocblas_operation_none, n_, nb,
                            kc, &one, chunk64(s, ci, kc), (int)ld,
                            C.p + (size_t)(i0 + cs) * kdim + j0, kdim,
                            ci == 0 ? &zero : &one, U.p, (int)ld));
    }
    col_dots(n_, nb, U.p, ld, U.p, ld, part.p, red.p, st);
    PV_HIP(hipMemcpyAsync(hv.data(), red.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    PV_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nb; i++) {
      const double t = sqrt(hv[(size_t)i]);
      if (t == 0.0) {
        sl.rc = 2;  // zero Ritz vector (SRC/cheblanNr.c:377-379)
        break;
      }
      fac[(size_t)i] = 1.0 / t;
    }
    if (sl.rc) break;
    PV_HIP(hipMemcpyAsync(red.p, fac.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st));
    col_scale(n_, nb, U.p, ld, red.p, st);
    // W = A U, Rayleigh quotients
    if (c_.A) {
      spmv_cols_f64(*c_.A, U.p, W.p, nb, ld,
                    c_.lanes ? c_.lanes : auto_lanes((double)c_.A->nnz / std::max(1, n_)),
                    st);
    } else {
      for (int i = 0; i < nb; i++) apply_A(c_, U.p + (size_t)i * ld, W.p + (size_t)i * ld);
    }
    sl.nmatvec += nb;
    col_dots(n_, nb, W.p, ld, U.p, ld, part.p, red.p, st);
    PV_HIP(hipMemcpyAsync(tq.data(), red.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    PV_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nb; i++) aneg[(size_t)i] = -tq[(size_t)i];
    PV_HIP(hipMemcpyAsync(red.p, aneg.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st));
    col_axpy(n_, nb, red.p, U.p, ld, W.p, ld, st);
    col_dots(n_, nb, W.p, ld, W.p, ld, part.p, red.p + CH, st);
    PV_HIP(hipMemcpyAsync(rr.data(), red.p + CH, sizeof(double) * nb, hipMemcpyDeviceToHost,
                          st));
    PV_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nb; i++) {
      const double t = tq[(size_t)i];
      if (t < sl.aa - DBL_EPSILON || t > sl.bb + DBL_EPSILON) continue;
      const double r = sqrt(rr[(size_t)i]);
      if (r < acc_tol) {
        sl.lam.push_back(t);
        sl.res.push_back(r);
        if (c_.keep_vectors) {
          const size_t off = sl.vecs.size();
          sl.vecs.resize(off + (size_t)n_);
          PV_HIP(hipMemcpyAsync(sl.vecs.data() + off, U.p + (size_t)i * ld,
                                sizeof(double) * n_, hipMemcpyDeviceToHost, st));
        }
      }
    }
    PV_HIP(hipStreamSynchronize(st));
  }
  p_->stats->t_ritz += wtime() - t0;
  c_.kstats.t_ritz += wtime() - t0;
}


// Ritz pairs with an FP64 Rayleigh-Ritz refinement: the selected Ritz
// vectors U = V Y are re-diagonalized against A itself,
//   (U^T A U) c = theta (U^T U) c,
// which separates eigenvectors that the filter cannot resolve to the basis'
// precision (an FP32 filter is flat to ~1e-7 near its peak).  The Gram
// products are DGEMMs (MFMA); the small generalized problem is rocSOLVER
// SYGVD.  Acceptance is the reference's: Rayleigh quotient of the refined
// vector inside the slice and residual below the threshold.
void BatchSolver::ritz_rr(int s, double tol) {
  Slice &sl = S_[(size_t)s];
  const hipStream_t st = c_.stream;
  const int kdim = sl.kdim;
  const double t0 = wtime();
  DBuf<double> D, E, C;
  DBuf<int> info;
  D.ensure((size_t)kdim);
  E.ensure((size_t)std::max(kdim, 1));
  C.ensure((size_t)kdim * kdim);
  info.ensure(1);
  PV_HIP(hipMemcpyAsync(D.p, sl.dT.data(), sizeof(double) * kdim, hipMemcpyHostToDevice, st));
  if (kdim > 1)
    PV_HIP(hipMemcpyAsync(E.p, sl.eT.data(), sizeof(double) * (kdim - 1),
                          hipMemcpyHostToDevice, st));
  PV_BLAS(rocsolver_dstedc(c_.blas, rocblas_evect_tridiagonal, kdim, D.p, E.p, C.p, kdim,
                           info.p));
  std::vector<double> ev((size_t)kdim);
  PV_HIP(hipMemcpyAsync(ev.data(), D.p, sizeof(double) * kdim, hipMemcpyDeviceToHost, st));
  PV_HIP(hipStreamSynchronize(st));
  int i0 = kdim;
  while (i0 > 0 && ev[(size_t)i0 - 1] >= sl.bar) i0--;
  const int nsel = kdim - i0;
  p_->stats->t_eig += wtime() - t0;
  c_.kstats.t_eig += wtime() - t0;
  if (nsel == 0) return;
  const size_t ld = L_.ld;
  const int csz = 1 << L_.chunk_shift;
  const double one = 1.0, zero = 0.0;
  DBuf<double> U, W;
  U.ensure(ld * nsel);
  W.ensure(ld * nsel);
  for (int j0 = 0, ci = 0; j0 < kdim; j0 += csz, ci++) {
    const int kc = std::min(csz, kdim - j0);
    PV_BLAS(rocblas_dgemm(c_.blas, rocblas_operation_none, rocblas_operation_none, n_, nsel, kc,
                          &one, chunk64(s, ci, kc), (int)ld, C.p + (size_t)i0 * kdim + j0,
                          kdim, ci == 0 ? &zero : &one, U.p, (int)ld));
  }
  C.release();
  free_basis(s);  // the basis is no longer needed
  conv_.release();
  const int lanes = c_.lanes ? c_.lanes : auto_lanes((double)c_.A->nnz / std::max(1, n_));
  spmv_cols_f64(*c_.A, U.p, W.p, nsel, ld, lanes, st);
  sl.nmatvec += nsel;
  // Gram matrices (upper triangles are read by SYGVD)
  DBuf<double> G, H, th, ework;
  G.ensure((size_t)nsel * nsel);
  H.ensure((size_t)nsel * nsel);
  th.ensure((size_t)nsel);
  ework.ensure((size_t)nsel);
  {
    DBuf<double> gw;
    gram_tn(c_.blas, st, n_, nsel, U.p, (int)ld, nsel, U.p, (int)ld, G.p, gw, c_.gram_kernel,
            true);
    gram_tn(c_.blas, st, n_, nsel, U.p, (int)ld, nsel, W.p, (int)ld, H.p, gw, c_.gram_kernel,
            true);
  }
  PV_BLAS(rocsolver_dsygvd(c_.blas, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper,
                           nsel, H.p, nsel, G.p, nsel, th.p, ework.p, info.p));
  int hinfo = 0;
  PV_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  PV_HIP(hipStreamSynchronize(st));
  if (hinfo) fprintf(stderr, "pevsl_amd: SYGVD info %d\n", hinfo);
  // refined vectors U2 = U Cr (into W), then A U2 (into U)
  PV_BLAS(rocblas_dgemm(c_.blas, rocblas_operation_none, rocblas_operation_none, n_, nsel, nsel,
                        &one, U.p, (int)ld, H.p, nsel, &zero, W.p, (int)ld));
  G.release();
  H.release();
  spmv_cols_f64(*c_.A, W.p, U.p, nsel, ld, lanes, st);
  sl.nmatvec += nsel;
  // normalize, Rayleigh quotients, residuals
  DBuf<double> part, red;
  part.ensure((size_t)nsel * col_part_blocks(n_) + 8);
  red.ensure((size_t)nsel);
  std::vector<double> nr((size_t)nsel), f((size_t)nsel), tq((size_t)nsel), rr((size_t)nsel);
  col_dots(n_, nsel, W.p, ld, W.p, ld, part.p, red.p, st);
  PV_HIP(hipMemcpyAsync(nr.data(), red.p, sizeof(double) * nsel, hipMemcpyDeviceToHost, st));
  PV_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < nsel; i++) {
    const double t = sqrt(nr[(size_t)i]);
    if (t == 0.0) {
      sl.rc = 2;
      return;
    }
    f[(size_t)i] = 1.0 / t;
  }
  PV_HIP(hipMemcpyAsync(red.p, f.data(), sizeof(double) * nsel, hipMemcpyHostToDevice, st));
  col_scale(n_, nsel, W.p, ld, red.p, st);
  col_scale(n_, nsel, U.p, ld, red.p, st);  // A (u / |u|) = (A u) / |u|
  col_dots(n_, nsel, U.p, ld, W.p, ld, part.p, red.p, st);
  PV_HIP(hipMemcpyAsync(tq.data(), red.p, sizeof(double) * nsel, hipMemcpyDeviceToHost, st));
  PV_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < nsel; i++) f[(size_t)i] = -tq[(size_t)i];
  PV_HIP(hipMemcpyAsync(red.p, f.data(), sizeof(double) * nsel, hipMemcpyHostToDevice, st));
  col_axpy(n_, nsel, red.p, W.p, ld, U.p, ld, st);
  col_dots(n_, nsel, U.p, ld, U.p, ld, part.p, red.p, st);
  PV_HIP(hipMemcpyAsync(rr.data(), red.p, sizeof(double) * nsel, hipMemcpyDeviceToHost, st));
  PV_HIP(hipStreamSynchronize(st));
  const double acc_tol = accept_tol(s, tol);
  if (c_.keep_vectors) sl.vecs.reserve(sl.vecs.size() + (size_t)nsel * n_);
  for (int i = 0; i < nsel; i++) {
    const double t = tq[(size_t)i];
    if (t < sl.aa - DBL_EPSILON || t > sl.bb + DBL_EPSILON) continue;
    const double r = sqrt(rr[(size_t)i]);
    if (r < acc_tol) {
      sl.lam.push_back(t);
      sl.res.push_back(r);
      if (c_.keep_vectors) {
        const size_t off = sl.vecs.size();
        sl.vecs.resize(off + (size_t)n_);
        PV_HIP(hipMemcpyAsync(sl.vecs.data() + off, W.p + (size_t)i * ld, sizeof(double) * n_,
                              hipMemcpyDeviceToHost, st));
      }
    }
  }
  PV_HIP(hipStreamSynchronize(st));
  p_->stats->t_ritz += wtime() - t0;
  c_.kstats.t_ritz += wtime() - t0;
}

// Slice set-up shared by run() and bench(): filters, bounds, the batch
// workspace and the pinned staging; -1 on an invalid filter centre.
int BatchSolver::prepare(double *intvs, int *maxits, pevsl_Polparams *pols, double tall) {
  const hipStream_t st = c_.stream;
  int kmax = 0;
  for (int s = 0; s < ns_; s++) {
    Slice &sl = S_[(size_t)s];
    const double *iv = intvs + 4 * (size_t)s;
    sl.idx = s;
    sl.aa = iv[0];
    sl.bb = iv[1];
    sl.pol = &pols[s];
    sl.bar = pols[s].bar;
    sl.gam = pols[s].gam;
    sl.maxit = std::min(p_->N, maxits[s]);
    sl.t_start = tall;
    // invalid interval: no eigenvalues, return 0 (SRC/cheblanNr.c:96-100)
    if (iv[0] >= iv[1] || iv[0] >= iv[3] || iv[1] <= iv[2] || sl.maxit <= 0) {
      sl.done = true;
      continue;
    }
    if (sl.gam > 1.0 || sl.gam < -1.0) {
      fprintf(stdout, "gamB error %.15e\n", sl.gam);
      return -1;
    }
    sl.dT.assign((size_t)sl.maxit, 0.0);
    sl.eT.assign((size_t)sl.maxit, 0.0);
    kmax = std::max(kmax, sl.maxit);
  }
  // FP32 basis storage: only with the FP32 filter on the device matrix
  b32_ = c_.basis_prec == PEVSL_PREC_FP32 && c_.filter_prec == PEVSL_PREC_FP32 && c_.A;
  vb_ = b32_ ? 4 : 8;
  L_.vbytes = (int)vb_;
  // batch workspace
  int shift = 0;
  while ((1 << shift) < std::max(64, c_.chunk_cols)) shift++;
  L_.n = n_;
  L_.ld = ((size_t)n_ + 63) / 64 * 64;
  L_.nslice = ns_;
  L_.chunk_shift = shift;
  L_.maxchunks = (kmax + 1 + (1 << shift) - 1) >> shift;
  vtab_h_.assign((size_t)ns_ * L_.maxchunks, nullptr);
  L_.vtab = vtab_.ensure(vtab_h_.size());
  double *sc = scal_.ensure((size_t)4 * ns_);
  L_.alpha = sc;
  L_.beta = sc + ns_;
  L_.nrm2a = sc + 2 * ns_;
  L_.nrm2b = sc + 3 * ns_;
  int *fl = flags_.ensure((size_t)2 * ns_);
  L_.pass2 = fl;
  L_.active = fl + ns_;
  PV_HIP(hipMemsetAsync(sc, 0, sizeof(double) * 4 * ns_, st));
  PV_HIP(hipMemsetAsync(fl, 0, sizeof(int) * 2 * ns_, st));
  L_.hcap = kmax + 1;
  L_.h = h_.ensure((size_t)ns_ * L_.hcap);
  L_.part_cap = lan_part_doubles(n_, ns_, L_.hcap);
  L_.part = part_.ensure(L_.part_cap);
  PV_HIP(hipHostMalloc(&pinned_, sizeof(double) * (2 * ns_ + 8)));
  PV_HIP(hipHostMalloc(&pinned_i_, sizeof(int) * (ns_ + 8)));
  return 0;
}

int BatchSolver::run(double *intvs, int *maxits, double tol, pevsl_Parvec *vinit,
                     pevsl_Polparams *pols, FILE *fstats) {
  n_ = p_->n;
  const hipStream_t st = c_.stream;
  const double tall = wtime();
  const double t_kern0 = c_.kt.seconds_matching("k_");
  if (prepare(intvs, maxits, pols, tall) < 0) return -1;

  std::vector<int> act;
  for (int s = 0; s < ns_; s++)
    if (!S_[(size_t)s].done) act.push_back(s);
  // V(:,0) = p(A) vinit / |.|  (FILTER_VINIT, SRC/cheblanNr.c:169-194)
  double *dvi = c_.cb_y.ensure((size_t)n_);
  PV_HIP(hipMemcpyAsync(dvi, vinit->data, sizeof(double) * n_, hipMemcpyHostToDevice, st));
  EvTimer ev_f, ev_r, ev_c;
  double t_filter = 0, t_reorth = 0, t_cgs = 0;
  for (int s : act) {
    ensure_column(s, 0);
    filter_into(s, dvi, -1, 0);
    S_[(size_t)s].nmatvec += S_[(size_t)s].pol->deg;
  }
  if (!act.empty()) {
    // |v|: reuse the Lanczos kernels with the basis column 0 as "znew"
    // (z = zold = nothing): nrm2a via lan_subtract_alpha with alpha = 0 is
    // not applicable at k = -1, so use the dot kernels directly.
    DBuf<double> pt;
    pt.ensure((size_t)dot_part_blocks(n_) + 8);
    for (int s : act) {
      const int nb = b32_ ? dot_partials_f32(n_, colf(s, 0), colf(s, 0), pt.p, st)
                          : dot_partials(n_, cold(s, 0), cold(s, 0), pt.p, st);
      reduce_partials(pt.p, nb, 1, pt.p + nb, st);
      PV_HIP(hipMemcpyAsync(pinned_, pt.p + nb, sizeof(double), hipMemcpyDeviceToHost, st));
      PV_HIP(hipStreamSynchronize(st));
      if (b32_)
        scale_vec_f32(n_, colf(s, 0), 1.0 / sqrt(pinned_[0]), st);
      else
        scale_vec(n_, cold(s, 0), nullptr, 1.0 / sqrt(pinned_[0]), st);
    }
  }
  for (int k = 0; !act.empty(); k++) {
    // ---- z_{k+1} = p(A) z_k for every active slice
    ev_f.start(st);
    filter_step(act, k);
    ev_f.stop(st);
    // ---- three-term update and full reorthogonalization
    ev_r.start(st);
    lanczos_update(act, k);
    ev_c.start(st);
    cgs(act, k);
    ev_c.stop(st);
    ev_r.stop(st);
    for (int s = 0; s < ns_; s++) {
      pinned_[s] = 0;
      pinned_[ns_ + s] = 0;
    }
    PV_HIP(hipMemcpyAsync(pinned_, L_.alpha, sizeof(double) * ns_, hipMemcpyDeviceToHost, st));
    PV_HIP(hipMemcpyAsync(pinned_ + ns_, L_.nrm2b, sizeof(double) * ns_,
                          hipMemcpyDeviceToHost, st));
    PV_HIP(hipMemcpyAsync(pinned_i_, L_.pass2, sizeof(int) * ns_, hipMemcpyDeviceToHost, st));
    PV_HIP(hipStreamSynchronize(st));
    t_filter += ev_f.seconds();
    t_reorth += ev_r.seconds();
    t_cgs += ev_c.seconds();
    for (int s : act) {
      if (pinned_i_[s]) {  // second DGKS pass ran
        const double pass = 2.0 * (k + 1) * (double)vb_ * n_ + 3.0 * (double)vb_ * n_;
        c_.kstats.reorth_bytes += pass;
        c_.kstats.cgs_bytes += pass;
        c_.kstats.cgs_passes++;
      }
    }
    std::vector<int> brk;
    std::vector<double> &betas = betas_h_;
    betas.assign((size_t)ns_, 0.0);
    for (int s : act) {
      Slice &sl = S_[(size_t)s];
      const double alpha = pinned_[s];
      const double beta = sqrt(pinned_[ns_ + s]);
      sl.dT[(size_t)k] = alpha;
      sl.wn += fabs(alpha);
      sl.wn += 2.0 * beta;
      sl.nwn += 3;
      betas[(size_t)s] = beta;
      if (beta * sl.nwn < orthTol * sl.wn) brk.push_back(s);
    }
    // regular normalization for the slices without breakdown
    PV_HIP(hipMemcpyAsync(L_.beta, betas.data(), sizeof(double) * ns_, hipMemcpyHostToDevice,
                          st));
    std::vector<int> reg;
    for (int s : act)
      if (std::find(brk.begin(), brk.end(), s) == brk.end()) reg.push_back(s);
    if (!reg.empty()) {
      set_active(reg);
      lan_scale_new(L_, k, st);
    }
    for (int s : brk) {
      if (fstats) fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", k, betas[(size_t)s]);
      breakdown(s, k);
      betas[(size_t)s] = 0.0;
    }
    std::vector<int> tests;
    for (int s : act) {
      Slice &sl = S_[(size_t)s];
      sl.eT[(size_t)k] = betas[(size_t)s];
      sl.its = k + 1;
      if ((k < kNtest || (k - kNtest) % kCycle != 0) && k != sl.maxit - 1) continue;
      tests.push_back(s);
    }
    convergence_tests(tests, k);
    if (fstats) {
      for (int s : tests) {
        fprintf(fstats, "  k %4d:   nconv %4d  tr1 %21.15e\n", k, S_[(size_t)s].nconv,
                S_[(size_t)s].tr1);
      }
      fflush(fstats);
    }
    // retire converged slices: Ritz pairs, then release their basis
    std::vector<int> still;
    for (int s : act) {
      Slice &sl = S_[(size_t)s];
      if (!sl.done) {
        still.push_back(s);
        continue;
      }
      if (k == sl.maxit - 1 && fstats)
        fprintf(fstats, "The max number of iterations [%d] has been reached.\n", sl.maxit);
      if (c_.ritz_refine == 1 ||
          (c_.ritz_refine < 0 && c_.filter_prec == PEVSL_PREC_FP32))
        ritz_rr(s, tol);  // releases the basis itself
      else
        ritz(s, tol);
      free_basis(s);
    }
    act.swap(still);
    if (!act.empty()) {
      // next step writes z_{k+2}; every active slice has beta set on device
    }
  }
  p_->stats->t_polAv += t_filter;
  p_->stats->t_reorth += t_reorth;
  // products with A: every filter degree and every Ritz residual
  // (pEVSL_MatvecA's count, INC/pevsl_int.h:83-89); filter applications:
  // the start vector, one per step and one per breakdown (pEVSL_ChebAv)
  for (int s = 0; s < ns_; s++) {
    const Slice &sl = S_[(size_t)s];
    if (c_.A) p_->stats->n_mvA += (size_t)sl.nmatvec;  // apply_A counts the callback's
    if (sl.its > 0) p_->stats->n_polAv += (size_t)(1 + sl.its + sl.nbreak);
  }
  // time of the products: the fused filter kernels' HIP-event time
  p_->stats->t_mvA += c_.kt.seconds_matching("k_") - t_kern0;
  c_.kstats.t_filter += t_filter;
  c_.kstats.t_reorth += t_reorth;
  c_.kstats.t_cgs += t_cgs;
  p_->stats->t_iter += wtime() - tall;
  c_.info.assign((size_t)ns_, pevsl_LanInfo{});
  for (int s = 0; s < ns_; s++) {
    Slice &sl = S_[(size_t)s];
    pevsl_LanInfo &in = c_.info[(size_t)s];
    in.its = sl.its;
    in.kdim = sl.kdim;
    in.nbreak = sl.nbreak;
    in.nconv = sl.nconv;
    in.nev = (int)sl.lam.size();
    in.deg = sl.pol ? sl.pol->deg : 0;
    in.nmatvec = sl.nmatvec;
    in.tr = sl.tr1;
    in.t_total = wtime() - tall;
    in.tol_accept = accept_tol(s, tol);
  }
  PV_HIP(hipHostFree(pinned_));
  PV_HIP(hipHostFree(pinned_i_));
  pinned_ = nullptr;
  pinned_i_ = nullptr;
  int rc = 0;
  for (int s = 0; s < ns_; s++)
    if (S_[(size_t)s].rc) rc = S_[(size_t)s].rc;
  return rc;
}

// Timed steps of a batched solve at a fixed basis size: every slice holds
// k + 2 basis columns (zero, with a random unit column k as the filter's
// input), and each step runs what one Lanczos iteration of the solve runs --
// the filter groups, the three-term update and the Gram-Schmidt pass against
// k + 1 columns -- followed by the solve's per-step synchronization.  The
// instance's kernel statistics and per-kernel timers are reset after the
// warm-up steps, so that they cover the timed steps only.  With graph_reps,
// the first filter group's launch sequence (pack, degree steps, unpack) is
// also captured in a HIP graph and replayed, against as many eager calls.
int BatchSolver::bench(double *intvs, pevsl_Polparams *pols, int k, int warmup, int reps,
                       int graph_reps, pevsl_StepBench *out) {
  n_ = p_->n;
  const hipStream_t st = c_.stream;
  std::vector<int> maxits((size_t)ns_, k + 2);
  if (prepare(intvs, maxits.data(), pols, wtime()) < 0) return -1;
  std::vector<int> act;
  for (int s = 0; s < ns_; s++)
    if (!S_[(size_t)s].done) act.push_back(s);
  if (act.empty()) return -2;
  double *tmp = c_.cb_y.ensure((size_t)n_);
  DBuf<double> pt;
  pt.ensure((size_t)dot_part_blocks(n_) + 8);
  for (int s : act) {
    ensure_column(s, k + 1);
    rand_fill((size_t)n_, tmp, 0x5eedULL + (uint64_t)s, 0, 0, st);
    const int nb = dot_partials(n_, tmp, tmp, pt.p, st);
    reduce_partials(pt.p, nb, 1, pt.p + nb, st);
    PV_HIP(hipMemcpyAsync(pinned_, pt.p + nb, sizeof(double), hipMemcpyDeviceToHost, st));
    PV_HIP(hipStreamSynchronize(st));
    scale_vec(n_, tmp, nullptr, 1.0 / sqrt(pinned_[0]), st);
    if (b32_)
      copy_f64_to_f32((size_t)n_, tmp, colf(s, k), st);
    else
      PV_HIP(hipMemcpyAsync(cold(s, k), tmp, sizeof(double) * n_, hipMemcpyDeviceToDevice, st));
  }
  EvTimer ev_all, ev_f, ev_r, ev_c;
  double t_all = 0, t_f = 0, t_r = 0, t_c = 0, wall = 0;
  for (int r = 0; r < warmup + reps; r++) {
    if (r == warmup) {
      PV_HIP(hipStreamSynchronize(st));
      memset(&c_.kstats, 0, sizeof(c_.kstats));
      c_.kt.reset();
    }
    const double w0 = wtime();
    ev_all.start(st);
    ev_f.start(st);
    filter_step(act, k);
    ev_f.stop(st);
    ev_r.start(st);
    lanczos_update(act, k);
    ev_c.start(st);
    cgs(act, k);
    ev_c.stop(st);
    ev_r.stop(st);
    ev_all.stop(st);
    PV_HIP(hipStreamSynchronize(st));
    if (r >= warmup) {
      wall += wtime() - w0;
      t_all += ev_all.seconds();
      t_f += ev_f.seconds();
      t_r += ev_r.seconds();
      t_c += ev_c.seconds();
    }
  }
  memset(out, 0, sizeof(*out));
  out->nslices = (int)act.size();
  out->k = k;
  out->reps = reps;
  if (reps > 0) {
    out->sec_step = t_all / reps;
    out->wall_step = wall / reps;
    out->sec_filter = t_f / reps;
    out->sec_reorth = t_r / reps;
    out->sec_cgs = t_c / reps;
    out->filter_bytes = c_.kstats.filter_bytes / reps;
    out->cgs_bytes = c_.kstats.cgs_bytes / reps;
    out->filter_launches = c_.kstats.filter_launches / reps;
  }
  const std::vector<std::vector<int>> groups = filter_groups(act);
  if (graph_reps > 0 && act.size() > 1 && c_.A && c_.block_width != 1 && groups[0].size() > 1) {
    const std::vector<int> &grp = groups[0];
    const pevsl_KernelStats ks = c_.kstats;
    filter_block(grp, k);  // stages the group's coefficients and pointers
    PV_HIP(hipStreamSynchronize(st));
    double w0 = wtime();
    for (int r = 0; r < graph_reps; r++) filter_block(grp, k);
    PV_HIP(hipStreamSynchronize(st));
    out->eager_group_sec = (wtime() - w0) / graph_reps;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    capture_ = true;
    PV_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    filter_block(grp, k);
    PV_HIP(hipStreamEndCapture(st, &graph));
    capture_ = false;
    PV_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    PV_HIP(hipGraphLaunch(exec, st));  // first replay (uploads the graph)
    PV_HIP(hipStreamSynchronize(st));
    w0 = wtime();
    for (int r = 0; r < graph_reps; r++) PV_HIP(hipGraphLaunch(exec, st));
    PV_HIP(hipStreamSynchronize(st));
    out->graph_group_sec = (wtime() - w0) / graph_reps;
    size_t nn = 0;
    PV_HIP(hipGraphGetNodes(graph, nullptr, &nn));
    out->graph_nodes = (int)nn;
    out->graph_group_width = (int)grp.size();
    PV_HIP(hipGraphExecDestroy(exec));
    PV_HIP(hipGraphDestroy(graph));
    c_.kstats = ks;
  }
  for (int s : act) free_basis(s);
  PV_HIP(hipHostFree(pinned_));
  PV_HIP(hipHostFree(pinned_i_));
  pinned_ = nullptr;
  pinned_i_ = nullptr;
  return 0;
}

void BatchSolver::outputs(int *nevOut, double **lamo, pevsl_Parvecs **Wo, double **reso,
                          pevsl_Parvec *vinit) {
  for (int s = 0; s < ns_; s++) {
    Slice &sl = S_[(size_t)s];
    const int nev = (int)sl.lam.size();
    nevOut[s] = nev;
    const size_t cap = (size_t)std::max(sl.maxit, std::max(nev, 1));
    double *lam, *res;
    PEVSL_MALLOC(lam, cap, double);
    PEVSL_MALLOC(res, cap, double);
    for (int i = 0; i < nev; i++) {
      lam[i] = sl.lam[(size_t)i];
      res[i] = sl.res[(size_t)i];
    }
    lamo[s] = lam;
    reso[s] = res;
    if (Wo) {
      if (c_.keep_vectors) {
        pevsl_Parvecs *W;
        PEVSL_MALLOC(W, 1, pevsl_Parvecs);
        pEVSL_ParvecsDuplParvec(vinit, std::max(nev, 1), vinit->n_local, W);
        W->n_vecs = nev;
        if (nev) memcpy(W->data, sl.vecs.data(), sizeof(double) * (size_t)nev * n_);
        Wo[s] = W;
      } else {
        Wo[s] = NULL;
      }
    }
  }
}

}  // namespace
}  // namespace pevsl

using namespace pevsl;

extern "C" {

int pEVSL_ChebLanNrBatch(pevsl_Data *pevsl, int nslices, double *intvs, int *maxits,
                         double tol, pevsl_Parvec *vinit, pevsl_Polparams *pols,
                         int *nevOut, double **lamo, pevsl_Parvecs **Wo, double **reso,
                         FILE *fstats) {
  Ctx &c = ctx_of(pevsl);
  if (nslices <= 0) return 0;
  if (!c.A && !c.host_mv) return -3;
  if (pevsl->n != vinit->n_local) return -4;
  if (!c.A && (int)c.h_x.size() != pevsl->n) {
    c.h_x.assign((size_t)pevsl->n, 0.0);
    c.h_y.assign((size_t)pevsl->n, 0.0);
  }
  if (pevsl->ifGenEv) {
    // generalized problem: one slice after the other on the device
    c.info.clear();
    for (int s = 0; s < nslices; s++) {
      double *iv = intvs + 4 * (size_t)s;
      pevsl_Parvecs *W = NULL;
      if (iv[0] >= iv[1] || iv[0] >= iv[3] || iv[1] <= iv[2]) {
        nevOut[s] = 0;
        lamo[s] = reso[s] = NULL;
        if (Wo) Wo[s] = NULL;
        c.info.push_back(pevsl_LanInfo{});
        continue;
      }
      const int rc = cheblan_gen(c, pevsl, iv, maxits[s], tol, vinit, &pols[s], &nevOut[s],
                                 &lamo[s], Wo ? &W : NULL, &reso[s], fstats);
      if (Wo) Wo[s] = W;
      if (rc) return rc;
    }
    return 0;
  }
  // outputs of every slice released and set empty (an error return)
  auto clear_outputs = [&](bool release) {
    for (int s = 0; s < nslices; s++) {
      if (release) {
        free(lamo[s]);
        free(reso[s]);
        if (Wo && Wo[s]) {
          pEVSL_ParvecsFree(Wo[s]);
          free(Wo[s]);
        }
      }
      nevOut[s] = 0;
      lamo[s] = reso[s] = NULL;
      if (Wo) Wo[s] = NULL;
    }
  };
  clear_outputs(false);
  const int w = c.batch_wave;
  try {
    if (w <= 0 || nslices <= w) {
      BatchSolver bs(c, pevsl, nslices);
      const int rc = bs.run(intvs, maxits, tol, vinit, pols, fstats);
      if (rc == -1) return rc;
      bs.outputs(nevOut, lamo, Wo, reso, vinit);
      return rc;
    }
    // waves of w slices in order of filter degree, so that the slices of one
    // wave share filter groups of similar degree; the remainder (nslices % w)
    // goes first, on the lowest degrees, where its narrower filter groups
    // cost least (100^3 Laplacian, 20 slices in waves of 8: 4 + 8 + 8 took
    // 89.7 s of filter against 91.0 s for 8 + 8 + 4)
    std::vector<int> order((size_t)nslices);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return pols[a].deg < pols[b].deg; });
    const int nw = (nslices + w - 1) / w;
    std::vector<pevsl_LanInfo> info((size_t)nslices, pevsl_LanInfo{});
    int rc = 0, pos = 0;
    const int rem = nslices % w;
    for (int wi = 0; wi < nw; wi++) {
      const int m = (wi == 0 && rem) ? rem : w;
      std::vector<double> iv((size_t)4 * m);
      std::vector<int> mi((size_t)m), nev((size_t)m);
      std::vector<pevsl_Polparams> pp((size_t)m);
      std::vector<double *> lam((size_t)m), res((size_t)m);
      std::vector<pevsl_Parvecs *> W((size_t)m, nullptr);
      for (int j = 0; j < m; j++) {
        const int sidx = order[(size_t)(pos + j)];
        for (int q = 0; q < 4; q++) iv[(size_t)4 * j + q] = intvs[(size_t)4 * sidx + q];
        mi[(size_t)j] = maxits[sidx];
        pp[(size_t)j] = pols[sidx];
      }
      int r;
      {
        BatchSolver bs(c, pevsl, m);
        r = bs.run(iv.data(), mi.data(), tol, vinit, pp.data(), fstats);
        if (r == -1) {
          clear_outputs(true);
          return r;
        }
        bs.outputs(nev.data(), lam.data(), Wo ? W.data() : nullptr, res.data(), vinit);
      }
      for (int j = 0; j < m; j++) {
        const int sidx = order[(size_t)(pos + j)];
        nevOut[sidx] = nev[(size_t)j];
        lamo[sidx] = lam[(size_t)j];
        reso[sidx] = res[(size_t)j];
        if (Wo) Wo[sidx] = W[(size_t)j];
        info[(size_t)sidx] = c.info[(size_t)j];
      }
      if (r) rc = r;
      pos += m;
    }
    c.info = info;
    return rc;
  } catch (const DeviceOom &e) {
    fprintf(stderr,
            "pEVSL_ChebLanNrBatch: a Lanczos basis chunk of %zu bytes does not fit in device "
            "memory; solve fewer slices at once (pEVSL_SetBatchWave)\n",
            e.bytes);
    clear_outputs(true);
    c.info.assign((size_t)nslices, pevsl_LanInfo{});
    return PEVSL_ERR_DEVICE_OOM;
  }
}

int pEVSL_SetGramKernel(pevsl_Data *pevsl, int kernel) {
  if (kernel < -1 || kernel > 1) return -1;
  ctx_of(pevsl).gram_kernel = kernel;
  return 0;
}

int pEVSL_BenchGram(pevsl_Data *pevsl, int n, int m, int reps, int upper, double *sec_blas,
                    double *sec_mfma, double *maxdiff) {
  Ctx &c = ctx_of(pevsl);
  if (n < 1 || m < 1 || reps < 1 || !sec_blas || !sec_mfma || !maxdiff) return -1;
  const bool up = upper != 0;
  const size_t ld = ((size_t)n + 63) / 64 * 64;
  DBuf<double> X, Y, C0, C1, work;
  X.ensure(ld * m);
  Y.ensure(ld * m);
  rand_fill(ld * m, X.p, 11, 0, 0, c.stream);
  rand_fill(ld * m, Y.p, 13, 0, 0, c.stream);
  C0.ensure((size_t)m * m);
  C1.ensure((size_t)m * m);
  double *out[2] = {C0.p, C1.p};
  double *sec[2] = {sec_blas, sec_mfma};
  // rocBLAS forms the whole product; the MFMA kernel the whole product or
  // (upper) the upper triangle, returned symmetric
  for (int kern = 0; kern < 2; kern++) {
    const bool u = up && kern == 1;
    gram_tn(c.blas, c.stream, n, m, X.p, (int)ld, m, Y.p, (int)ld, out[kern], work, kern, u);
    EvTimer ev;
    ev.start(c.stream);
    for (int r = 0; r < reps; r++)
      gram_tn(c.blas, c.stream, n, m, X.p, (int)ld, m, Y.p, (int)ld, out[kern], work, kern, u);
    ev.stop(c.stream);
    *sec[kern] = ev.seconds() / reps;
  }
  std::vector<double> h0((size_t)m * m), h1((size_t)m * m);
  PV_HIP(hipMemcpy(h0.data(), C0.p, sizeof(double) * h0.size(), hipMemcpyDeviceToHost));
  PV_HIP(hipMemcpy(h1.data(), C1.p, sizeof(double) * h1.size(), hipMemcpyDeviceToHost));
  double dmax = 0.0, cmax = 0.0;
  for (int cc = 0; cc < m; cc++)
    for (int r = 0; r < m; r++) {
      const size_t i = (size_t)cc * m + r;
      if (up && r > cc) {
        // the lower triangle mirrors the upper one exactly
        if (h1[i] != h1[(size_t)r * m + cc]) dmax = HUGE_VAL;
        continue;
      }
      dmax = std::max(dmax, std::fabs(h0[i] - h1[i]));
      cmax = std::max(cmax, std::fabs(h0[i]));
    }
  *maxdiff = cmax > 0 ? dmax / cmax : dmax;
  return 0;
}

int pEVSL_BenchBatchStep(pevsl_Data *pevsl, int nslices, double *intvs, pevsl_Polparams *pols,
                         int k, int warmup, int reps, int graph_reps, pevsl_StepBench *out) {
  Ctx &c = ctx_of(pevsl);
  if (!c.A || pevsl->ifGenEv || nslices <= 0 || k < 1 || k + 2 > pevsl->N || warmup < 0 ||
      reps < 0 || graph_reps < 0 || !out)
    return -1;
  try {
    BatchSolver bs(c, pevsl, nslices);
    return bs.bench(intvs, pols, k, warmup, reps, graph_reps, out);
  } catch (const DeviceOom &e) {
    fprintf(stderr, "pEVSL_BenchBatchStep: %d bases of %d columns do not fit in device memory\n",
            nslices, k + 2);
    return PEVSL_ERR_DEVICE_OOM;
  }
}

int pEVSL_ChebAv(pevsl_Data *pevsl, pevsl_Polparams *pol, pevsl_Parvec *v,
                 pevsl_Parvec *y, pevsl_Parvec *w) {
  (void)w;
  Ctx &c = ctx_of(pevsl);
  const int n = pevsl->n;
  if (v->n_local != n || y->n_local != n) return -4;
  if (!c.A && (int)c.h_x.size() != n) {
    c.h_x.assign((size_t)n, 0.0);
    c.h_y.assign((size_t)n, 0.0);
  }
  double *d = c.cb_x.ensure((size_t)2 * n);
  PV_HIP(hipMemcpyAsync(d, v->data, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
  const double t0 = wtime();
  if (pevsl->ifGenEv)
    cheb_filter_gen(c, pevsl, pol, d, d + n);
  else
    cheb_filter_col(c, pol, d, d + n);
  PV_HIP(hipMemcpyAsync(y->data, d + n, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
  PV_HIP(hipStreamSynchronize(c.stream));
  pevsl->stats->t_polAv += wtime() - t0;
  pevsl->stats->n_polAv++;
  if (c.A && !pevsl->ifGenEv) pevsl->stats->n_mvA += (size_t)pol->deg;
  return 0;
}

int pEVSL_ChebLanNr(pevsl_Data *pevsl, double *intv, int maxit, double tol,
                    pevsl_Parvec *vinit, pevsl_Polparams *pol, int *nevOut,
                    double **lamo, pevsl_Parvecs **Wo, double **reso, FILE *fstats) {
  // invalid interval: nothing computed, return 0 (SRC/cheblanNr.c:96-100)
  if (intv[0] >= intv[1] || intv[0] >= intv[3] || intv[1] <= intv[2]) {
    *nevOut = 0;
    *lamo = NULL;
    *Wo = NULL;
    *reso = NULL;
    return 0;
  }
  if (fstats) {
    fprintf(fstats, " intv:[%e, %e, %e, %e]\n", intv[0], intv[1], intv[2], intv[3]);
    fprintf(fstats, " ** Cheb Poly of deg = %d, gam = %.15e, bar: %.15e\n", pol->deg,
            pol->gam, pol->bar);
    fprintf(fstats, " ** Cheb-LanNr \n");
  }
  pevsl_Parvecs *W = NULL;
  const int rc = pEVSL_ChebLanNrBatch(pevsl, 1, intv, &maxit, tol, vinit, pol, nevOut, lamo,
                                      &W, reso, fstats);
  *Wo = W;
  return rc;
}

}  // extern "C"
