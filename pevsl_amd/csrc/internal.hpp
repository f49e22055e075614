// internal.hpp -- host-side state shared by the C-ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "kernels.hpp"

extern "C" {
#include "pevsl.h"
#include "pevsl_amd.h"
#include "pevsl_itsol.h"
}

namespace pevsl {

[[noreturn]] void fatal(const char *what, const char *file, int line);

#define PV_HIP(call)                                                     \
  do {                                                                   \
    hipError_t e__ = (call);                                             \
    if (e__ != hipSuccess) ::pevsl::fatal(hipGetErrorString(e__),        \
                                          __FILE__, __LINE__);           \
  } while (0)

#define PV_BLAS(call)                                                    \
  do {                                                                   \
    rocblas_status s__ = (call);                                         \
    if (s__ != rocblas_status_success)                                   \
      ::pevsl::fatal(rocblas_status_to_string(s__), __FILE__, __LINE__); \
  } while (0)

// Owned device allocation.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // grows to at least cnt elements; contents are not preserved
  T *ensure(size_t cnt) {
    if (cnt <= n && p) return p;
    release();
    PV_HIP(hipMalloc(&p, sizeof(T) * (cnt ? cnt : 1)));
    n = cnt;
    return p;
  }
};

// Device copy of a CSR matrix (attached to pevsl_Parcsr.dev).
struct DevMatrix {
  DevCsr csr;
  DBuf<int> ia, ja, rowblk;
  DBuf<int> rowblk_q[5];  // partitions of the nnz-balanced kernels, Q >= 2
  DBuf<double> a64;
  DBuf<float> a32;
  DBuf<short> jo16;
};

// Live per-kernel timing on an instance's stream: HIP events around a run of
// back-to-back launches of one kernel (nothing else enqueued in between),
// resolved lazily once the events have completed.  Accumulates launches,
// seconds and algorithmic bytes per kernel name (pEVSL_GetKernelTimes).
struct KernTimers {
  struct Acc {
    std::string name;
    long long launches = 0;
    double sec = 0.0, bytes = 0.0;
  };
  struct Pend {
    size_t acc;
    hipEvent_t a, b;
    long long launches;
    double bytes;
  };
  std::vector<Acc> acc;
  std::vector<Pend> pend;
  std::vector<hipEvent_t> pool;
  KernTimers() = default;
  KernTimers(const KernTimers &) = delete;
  KernTimers &operator=(const KernTimers &) = delete;
  ~KernTimers();
  // records the start event; returns a token for end()
  long begin(hipStream_t s, const char *name);
  // records the end event of the run of `launches` launches
  void end(hipStream_t s, long tok, long long launches, double bytes);
  // waits for the pending runs and accumulates them
  void flush();
  void reset();
  double seconds_matching(const char *prefix);

 private:
  hipEvent_t take();
};

// Rows longer than this make the batched filter default to the nnz-balanced
// block kernels.
constexpr int kUnevenRow = 64;

// Highest filter degree a batched FP32 group runs in Clenshaw form; above it
// the group runs the three-term form (same vector passes with y updated every
// other degree, smaller rounding error at high degree).
constexpr int kClenMaxDeg = 400;

// Per-instance device context (attached to pevsl_Data.dev).
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  const DevCsr *A = nullptr;       // set by SetAParcsr: &A_view
  const DevCsr *A_src = nullptr;   // the matrix's device copy
  DevCsr A_view;                   // A_src, without jo16 when col16 == 0
  int col16 = 1;                   // 16-bit column offsets where the matrix has them
  MVFunc host_mv = nullptr;        // set by SetAMatvec (host callback)
  void *host_mv_data = nullptr;
  int filter_prec = PEVSL_PREC_FP64;
  int basis_prec = PEVSL_PREC_FP64;  // Lanczos basis storage (FP32 only with an FP32 filter)
  double mixed_tol = 1e-6;         // FP32-filter path: residual threshold / spectral radius
  int ritz_refine = -1;            // FP64 Rayleigh-Ritz refinement: 1 on, 0 off, -1 FP32 only
  int rand_mode = PEVSL_RAND_LIBC;
  unsigned long long rand_seed = 0;
  unsigned long long rand_offset = 0;
  int lanes = 0;                   // lanes per row of the single-column SpMV
  int use_stream = -1;             // CSR-stream filter kernel: 1 yes, 0 no, -1 auto
  int block_width = 0;             // columns per block SpMM (0 = auto)
  int blk_rowkernel = -1;          // block SpMM: 1 row-per-thread, 0 lane per (row, col), -1 auto
  int keep_vectors = 1;
  int chunk_cols = 512;            // Lanczos basis allocation granularity
  int batch_wave = 0;              // ChebLanNrBatch: slices per wave (0: all at once)
  int graph_filter = 0;            // batched filter groups replayed from hipGraphs
  int gram_kernel = -1;            // Rayleigh-Ritz Gram products: 0 rocBLAS, 1 MFMA kernel,
                                   // -1 by size
  std::vector<pevsl_LanInfo> info; // per slice of the last (batched) solve
  pevsl_KernelStats kstats{};      // accumulated device timings
  KernTimers kt;                   // per-kernel HIP-event timings
  pevsl_Stat *stats = nullptr;     // the instance's pevsl_Stat (n_mvA, t_mvA)
  // staging for host callbacks / small copies
  DBuf<double> cb_x, cb_y;
  std::vector<double> h_x, h_y;
  // scratch
  DBuf<double> scratch64;
  // generalized problem: device B (SetBParcsr) and work space
  const DevCsr *B = nullptr;
  std::vector<double> hb_x, hb_y;
  DBuf<double> gen_lsw, gen_dot, gen_cgs, gen_fw, gen_tri;
};

Ctx &ctx_of(pevsl_Data *p);
const DevCsr *dev_matrix_of(pevsl_Parcsr *A);

// lanes per row for a matrix with mean row length r (power of two, <= 64)
int auto_lanes(double mean_row);

// Wall-clock + device-event timing helper: measures the device time of the
// work enqueued between start() and stop() on one stream.
struct EvTimer {
  hipEvent_t a = nullptr, b = nullptr;
  EvTimer() {
    PV_HIP(hipEventCreate(&a));
    PV_HIP(hipEventCreate(&b));
  }
  ~EvTimer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  void start(hipStream_t s) { PV_HIP(hipEventRecord(a, s)); }
  void stop(hipStream_t s) { PV_HIP(hipEventRecord(b, s)); }
  double seconds() {
    PV_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    PV_HIP(hipEventElapsedTime(&ms, a, b));
    return 1e-3 * ms;
  }
};

double wtime();

// y = A x on device through the installed operator (device CSR or the host
// callback); used by the unfused paths.
void apply_A(Ctx &c, const double *dx, double *dy);

// Random vectors: libc rand() stream on the host (reference parity,
// SRC/parvec.c:65-72) or the device counter generator.
void rand_vectors(Ctx &c, int n, int nvec, double *d_out, size_t ld);

// Chebyshev filter y = p((A-cc)/dd) v on the device for one FP64 column
// (v, y in HBM).  Work vectors come from the context, or, when `work` is
// given, from that device buffer of 3n doubles (device matrix only); the call
// then allocates nothing and records no per-kernel timers.
void cheb_filter_col(Ctx &c, const pevsl_Polparams *pol, const double *dv,
                     double *dy, double *work = nullptr);
// The same for FP32 columns (FP32 filter arithmetic; device matrix only).
void cheb_filter_col32(Ctx &c, const pevsl_Polparams *pol, const float *dv, float *dy);

// ---- generalized problem (gen.cpp) -----------------------------------------
// y = B x, x = B \ b, x = L^{-T} b on device vectors through the installed
// operator / solve (device CSR, the library's LSPol solve on the device, or a
// host callback with host copies).
void apply_B(Ctx &c, pevsl_Data *p, const double *dx, double *dy);
void solve_B(Ctx &c, pevsl_Data *p, const double *dx, double *dy);
void solve_LT(Ctx &c, pevsl_Data *p, const double *dx, double *dy);
double gen_dot(Ctx &c, int n, const double *x, const double *y);
// ngs passes of v -= N(:,0:k-1) (T(:,0:k-1)^T v)
void cgs2(Ctx &c, int n, int k, int ngs, const double *T, const double *N, size_t ld,
          double *v);
// generalized filter y = p(A B^{-1}) v (device vectors)
void cheb_filter_gen(Ctx &c, pevsl_Data *p, const pevsl_Polparams *pol, const double *dv,
                     double *dy);
// generalized KPM dot products dots[m*Mdeg + k]
void kpm_moments_gen(Ctx &c, pevsl_Data *p, int Mdeg, int nvec, double ctr, double wid,
                     std::vector<double> &dots);
// y = f(x) on device vectors when f is one of the library's Chebyshev-
// iteration solves (itsol.cpp); false otherwise.  wk: 3 n doubles.
bool chebiter_device_solve(SVFunc f, void *data, const double *dx, double *dy, double *wk,
                           hipStream_t s);
int chebiter_size(void *data);
// tridiagonal matrices of the generalized Lanczos DOS, dT/eT[m*maxit + k]
void landos_gen(Ctx &c, pevsl_Data *p, int nvec, int maxit, std::vector<double> &dT,
                std::vector<double> &eT);
// generalized ChebLanNr for one slice (outputs as pEVSL_ChebLanNr)
int cheblan_gen(Ctx &c, pevsl_Data *p, double *intv, int maxit, double tol,
                pevsl_Parvec *vinit, pevsl_Polparams *pol, int *nevOut, double **lamo,
                pevsl_Parvecs **Wo, double **reso, FILE *fstats);

}  // namespace pevsl
