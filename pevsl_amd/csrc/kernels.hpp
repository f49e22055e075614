// kernels.hpp -- launchers of the HIP kernels (kernels.hip) used by the host
// side of the library.  All launchers enqueue on the given stream and never
// synchronize; shapes are validated by the callers (see check_* in dev.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace pevsl {

// Device CSR: int32 row pointers / column indices, values in FP64 and FP32.
struct DevCsr {
  int n = 0;       // rows == cols
  int nnz = 0;
  int *ia = nullptr;
  int *ja = nullptr;
  double *a64 = nullptr;
  float *a32 = nullptr;
  int max_row = 0;  // longest row
  // row blocks of the CSR-stream kernels: consecutive rows with at most
  // kStreamRows rows and kStreamNnz nonzeros (a longer row is a block alone)
  int *rowblk = nullptr;
  int nrowblk = 0;
  // row blocks of the nnz-balanced block kernels, one partition per number
  // Q = 2^i of 16-byte pieces in a B-wide row: at most kStreamRows/Q rows and
  // kStreamNnz/Q nonzeros per block (rowblk_q[0] == rowblk)
  int *rowblk_q[5] = {};
  int nrowblk_q[5] = {};
  // ja as 16-bit offsets from the first row of each nonzero's block in
  // `rowblk` (null when an offset does not fit in a short): the CSR-stream
  // and row-per-thread block kernels then read 2 index bytes per nonzero
  short *jo16 = nullptr;
};

// index bytes per nonzero read by the CSR-stream and row-block kernels
inline int index_bytes(const DevCsr &A) { return A.jo16 ? 2 : 4; }

constexpr int kStreamRows = 256;
constexpr int kStreamNnz = 2048;

// host: row-block partition with at most max_rows rows and max_nnz nonzeros
// per block, a longer row being a block alone (returns nrowblk+1 offsets);
// the CSR-stream kernels use (kStreamRows, kStreamNnz)
void build_row_blocks(int n, const int *ia, std::vector<int> &blk,
                      int max_rows = kStreamRows, int max_nnz = kStreamNnz);

// ---- sparse products -------------------------------------------------------
// y = A x, one column.  lanes = lanes per row in {1,2,4,8,16,32,64}.
void spmv_f64(const DevCsr &A, const double *x, double *y, int lanes,
              hipStream_t s);
// Y(:,c) = A X(:,c) for ncol column-major columns with leading dim ld.
void spmv_cols_f64(const DevCsr &A, const double *X, double *Y, int ncol,
                   size_t ld, int lanes, hipStream_t s);

// One Chebyshev degree step of the filter for one column (T = float/double):
//   first:  vkp1 = t (A vk - cc vk);        y = mu0 vk + muk vkp1
//   else:   vkp1 = t (A vk - cc vk) - vkm1; y = y + muk vkp1
template <typename T>
void cheb_step_vec(const DevCsr &A, const T *vk, const T *vkm1, T *vkp1, T *y,
                   double cc, double t, double mu0, double muk, bool first,
                   int lanes, hipStream_t s);

// Same step with the CSR-stream kernel (one block per row block: coalesced
// products staged in LDS, then one sequential sum per row).
template <typename T>
void cheb_step_stream(const DevCsr &A, const T *vk, const T *vkm1, T *vkp1,
                      T *y, double cc, double t, double mu0, double muk,
                      bool first, hipStream_t s);

// Same for B interleaved columns (row-major n x B block, column fastest).
// mus: device array [(deg_max+1) x B] of per-column coefficients; k = degree
// index of this step (1..); columns c >= nact are left untouched.
template <typename T>
void cheb_step_blk(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                   T *y, double cc, double t, const double *mus, int k,
                   int nact, hipStream_t s);

// y-update modes of the three-term block steps (cheb_step_blkrow,
// cheb_step_bal; k >= 2 for kYPair):
//   kYPlain  y += mu_k vkp1 (k == 1: y = mu_0 vk + mu_1 vkp1)
//   kYSkip   y is neither read nor written (its term waits for the next step)
//   kYPair   y += mu_{k-1} vk, then y += mu_k vkp1: the updates of a skipped
//            step and this one, in their order, for columns c < nact (active
//            at degree k-1; mus is zero past each column's degree)
// Skip and pair steps alternate, so y costs one read and one write per two
// degrees, with the results of plain steps bit for bit.
constexpr int kYPlain = 0, kYSkip = 2, kYPair = 1;

// Same with one thread per row (row blocks of the CSR-stream partition,
// a/ja staged in LDS, B-wide 16-byte row gathers); B <= 32.
template <typename T>
void cheb_step_blkrow(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                      T *y, double cc, double t, const double *mus, int k,
                      int nact, hipStream_t s, int ymode = kYPlain);

// Clenshaw step on B interleaved columns (row per thread, B <= 32):
//   out = mu_k v + t (A b1 - cc b1) - b2,  mu_k per column = mus[k*B + c];
// b2 == nullptr reads as zero; b1 == nullptr gives the top term out = mu_k v.
template <typename T>
void clen_step_blkrow(const DevCsr &A, int B, const T *b1, const T *b2, const T *v, T *out,
                      double cc, double t, const double *mus, int k, hipStream_t s);

// Clenshaw step on the active 16-byte pieces (SEG = 16/sizeof(T) columns
// each) of an interleaved block of Bp columns (Bp a multiple of SEG, at most
// kMaxSegs pieces), one thread per (row, piece): pieces 0..nseg-1 are active
// (sorted by top degree, descending); piece s writes the top term mu_d v at
// d == tops.top[s], reads b2 as zero at d == top - 1, and otherwise computes
// out = mu_d v + t (A b1 - cc b1) - b2 with mus[d*Bp + c], as
// clen_step_blkrow does per column.  Pieces >= nseg are not touched.
constexpr int kMaxSegs = 16;
struct SegTops {
  int top[kMaxSegs];
};
template <typename T>
void clen_step_seg(const DevCsr &A, int Bp, const T *b1, const T *b2, const T *v, T *out,
                   double cc, double t, const double *mus, int d, int nseg, const SegTops &tops,
                   hipStream_t s);

// nnz-balanced forms of cheb_step_blkrow and clen_step_blkrow (same
// arguments and results; products formed in parallel over nonzeros and
// staged in LDS, long rows reduced by a whole workgroup); B <= 32.
template <typename T>
void cheb_step_bal(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                   T *y, double cc, double t, const double *mus, int k,
                   int nact, hipStream_t s, int ymode = kYPlain);
template <typename T>
void clen_step_bal(const DevCsr &A, int B, const T *b1, const T *b2, const T *v, T *out,
                   double cc, double t, const double *mus, int k, hipStream_t s);
// partition index of the balanced kernels for element size elem, width B
int bal_level(int elem, int B);

// KPM step on B interleaved columns:
//   vkp1 = scal (A vk - ctr vk) - vkm1,   part[blk][c] = sum_rows vkp1*v0
// returns the number of row blocks (partials per column) written.
template <typename T>
int kpm_step_blk(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                 const T *v0, double ctr, double scal, double *part,
                 hipStream_t s);
int kpm_part_blocks(int n, int B);
// The same step with the nnz-balanced products (uneven rows, B <= 32):
// partials per row block of the balanced partition, returns their number
// (0 if B has no partition).
template <typename T>
int kpm_step_bal(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1, const T *v0,
                 double ctr, double scal, double *part, hipStream_t s);

// ---- dense / vector kernels (FP64 unless stated) --------------------------
// out[c] = sum_b part[b*ncol + c], b < nblk (fixed order); out may be strided.
void reduce_partials(const double *part, int nblk, int ncol, double *out,
                     hipStream_t s);
// y = x for n doubles (n even, 16-byte aligned): plain copy kernel
void copy_f64(size_t n, const double *x, double *y, hipStream_t s);
// part[p] (m1 x m2, column-major) = X(rows p*chunk .. , :)^T Y(same rows, :)
// for p < nparts (the last part ends at row n), FP64 MFMA (hand-written
// v_mfma_f64_16x16x4 tiles); X and Y column-major with leading dims ldx, ldy.
// upper (m1 == m2): only the 64 x 64 tiles on and above the diagonal.
void gram_tn_mfma(int n, int m1, const double *X, size_t ldx, int m2, const double *Y,
                  size_t ldy, int chunk, int nparts, bool upper, double *part, hipStream_t s);
// C (m x m) = sum over p < nparts of the upper triangles of part[p], in part
// order, stored in both triangles (a symmetric C)
void sum_parts_upper(const double *part, int nparts, int m, double *C, hipStream_t s);
// out[e] = sum_p part[p*nel + e] over p = 0..nparts-1 in order
void sum_parts(const double *part, int nparts, size_t nel, double *out, hipStream_t s);
// part[blk] = sum x_i*y_i over the block's rows; returns nblk
int dot_partials(int n, const double *x, const double *y, double *part,
                 hipStream_t s);
int dot_part_blocks(int n);
void scale_vec(int n, double *x, const double *inv_scale_dev, double hostscale,
               hipStream_t s);
void axpy_vec(int n, double a, const double *x, double *y, hipStream_t s);
// vkp1 = t (vkp1 - cc vk) - vkm1 (vkm1 may be null);  y += mu vkp1 (y may be
// null): the vector part of a Chebyshev step with a separately formed product
void cheb_update(int n, const double *vk, const double *vkm1, double *vkp1, double *y,
                 double cc, double t, double mu, hipStream_t s);
void copy_f64_to_f32(size_t n, const double *x, float *y, hipStream_t s);
void copy_f32_to_f64(size_t n, const float *x, double *y, hipStream_t s);
void fill_zero_f64(size_t n, double *x, hipStream_t s);

// Pack/unpack between one column-major vector of type TV per slice and an
// interleaved n x B block of type T.  cols[c] points to the n-vector of
// column c (device array of pointers), c < ncol.
template <typename T, typename TV = double>
void pack_cols(int n, int B, int ncol, const TV *const *cols, T *blk,
               hipStream_t s);
template <typename T, typename TV = double>
void unpack_cols(int n, int B, int ncol, const T *blk, TV *const *cols,
                 hipStream_t s);
// FP32 vector helpers (FP64 accumulation)
int dot_partials_f32(int n, const float *x, const float *y, double *part,
                     hipStream_t s);
void scale_vec_f32(int n, float *x, double hostscale, hipStream_t s);

// ---- Lanczos on a batch of slices ----------------------------------------
// Slices are independent; slice s owns V_s (chunked columns, leading dim ld
// a multiple of 64, chunks of 2^chunk_shift >= 32 columns whose rows
// [n, ld) are zero) and per-step scalars.  Column j of slice s is
//   vtab[s*maxchunks + (j >> chunk_shift)] + (size_t)(j & chunk_mask)*ld
// elements of vbytes bytes: FP64 (8) or FP32 (4) storage.  All arithmetic
// (dots, updates, norms) is FP64 either way.
struct LanBatch {
  int n = 0;
  size_t ld = 0;
  int nslice = 0;
  int chunk_shift = 9;
  int maxchunks = 0;
  int vbytes = 8;           // basis element size: 8 (double) or 4 (float)
  void **vtab = nullptr;    // device [nslice*maxchunks]
  // device scalars per slice
  double *alpha = nullptr;  // [nslice]
  double *beta = nullptr;   // [nslice] (beta of the previous step)
  double *nrm2a = nullptr;  // [nslice] |v|^2 before reorth pass
  double *nrm2b = nullptr;  // [nslice] |v|^2 after a reorth pass
  int *pass2 = nullptr;     // [nslice] 1 if the 2nd DGKS pass is needed
  int *active = nullptr;    // [nslice] 1 if the slice takes part
  double *h = nullptr;      // [nslice * hcap]
  int hcap = 0;
  double *part = nullptr;   // partial sums workspace
  size_t part_cap = 0;      // doubles
};

// doubles of LanBatch::part needed for n rows, nslice slices and hcap
// reorthogonalization coefficients per slice
size_t lan_part_doubles(int n, int nslice, int hcap);

// znew_s -= beta_s zold_s (k > 0);  alpha_s = <z_s, znew_s>
void lan_alpha(const LanBatch &L, int k, hipStream_t s);
// znew_s -= alpha_s z_s;  nrm2a_s = |znew_s|^2
void lan_subtract_alpha(const LanBatch &L, int k, hipStream_t s);
// One classical Gram-Schmidt pass against V_s(:,0:k): h = V^T znew,
// znew -= V h, nrm2b = |znew|^2.  pass = 0 always runs; pass = 1 runs only
// for slices whose pass2 flag is set (decided on the device after pass 0).
void lan_cgs_pass(const LanBatch &L, int k, int pass, hipStream_t s);
// znew_s /= beta_s for active slices (beta read from the device array).
void lan_scale_new(const LanBatch &L, int k, hipStream_t s);

// ---- tridiagonal (FP64) ---------------------------------------------------
// Restricted trace of T (diag d[0..m-1], sub-diag e[0..m-2]): eigenvalues
// >= thr by Sturm bisection.  out[0] = sum, out[1] = count (as double).
// Workspace: lam (m doubles).
void tridiag_trace_above(int m, const double *d, const double *e, double thr,
                         double *lam, double *out, hipStream_t s);
// The same for np matrices of equal order m at once: T_p in d/e + p*ld,
// thresholds thr[p] (device), eigenvalues in lam + p*ld, 4 workspace values
// per problem in ws, out[2p] = sum and out[2p+1] = count.
void tridiag_trace_above_batch(int np, int m, size_t ld, const double *d, const double *e,
                               const double *thr, double *lam, double *ws, double *out,
                               hipStream_t s);
// All eigenvalues of T by bisection (ascending), for tests and small m.
void tridiag_eigvals(int m, const double *d, const double *e, double *lam,
                     hipStream_t s);

// ---- column-major multi-vectors (Ritz stage) ------------------------------
// part must hold ncol * col_part_blocks(n) doubles.
int col_part_blocks(int n);
// out[c] = <X(:,c), Y(:,c)>
void col_dots(int n, int ncol, const double *X, size_t ldx, const double *Y,
              size_t ldy, double *part, double *out, hipStream_t s);
// X(:,c) *= f[c]
void col_scale(int n, int ncol, double *X, size_t ld, const double *f,
               hipStream_t s);
// Y(:,c) += a[c] X(:,c)
void col_axpy(int n, int ncol, const double *a, const double *X, size_t ldx,
              double *Y, size_t ldy, hipStream_t s);

// ---- misc -----------------------------------------------------------------
// y[j] = integral of the KPM DOS at x[j] (closed form, per-point sequential
// sum over degrees as the reference does).
void kpm_intchx(int Mdeg, const double *mu, int npts, const double *th,
                double th0, double *yi, hipStream_t s);
// Rademacher (+-1) or uniform [-1,1] counter-based random numbers.
void rand_fill(size_t n, double *x, uint64_t seed, uint64_t offset,
               int rademacher, hipStream_t s);

}  // namespace pevsl
