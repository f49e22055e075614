// kernels.hip -- hand-written HIP kernels for gfx950 (CDNA4, wave64) behind
// the pEVSL C API.
//
// Hot path (HBM-bound, ~1 flop/byte).  Every filter kernel fuses one degree
// of the Chebyshev filter (SRC/chebpol.c:538-568) into the sparse product:
// the SpMV/SpMM, the recurrence and the coefficient accumulation are one pass
// over the matrix and the vectors.
//   * k_cheb_stream : one column, CSR-stream form.  A row block (<= 256 rows,
//     <= 2048 nonzeros) multiplies its nonzeros with the gathered x values in
//     coalesced order into LDS, then each thread sums one row sequentially
//     (the reference's order); a longer row is one block, reduced by all of
//     its threads.  The default single-slice filter.
//   * k_cheb_vec : one column with L lanes per row (tuning alternative).
//   * k_cheb_blkrow / k_clen_blkrow : B slices as interleaved columns
//     (row-major n x B), a/ja staged in LDS, one thread per row gathering
//     whole B-wide rows with 16-byte loads; three-term form with an
//     accumulator, or the Clenshaw form (four vector passes instead of five).
//   * k_clen_seg : the Clenshaw step with one thread per (row, 16-byte piece
//     of columns), up to 64 FP32 columns in one pass over the matrix, each
//     piece active from its own top degree (uneven degrees cost their own
//     steps only, at 4-column granularity).
//   * k_blk_bal (epilogues EpiCheb / EpiClen) and k_kpm_bal : the same steps
//     for matrices with uneven rows, parallel over (nonzero, 16-byte piece)
//     with the products staged in LDS and summed per row in column order.
//   * k_cheb_blk : B columns with one lane per (row, column).
//   * k_kpm_blk : one degree of the KPM recurrence (SRC/kpmdos.c:129-156) for
//     all random vectors at once, fused with the per-column dot product
//     against v0 (block partial sums, deterministic second stage).
//   * k_lan_* : Lanczos vector updates and classical Gram-Schmidt with the
//     DGKS test (SRC/cheblanNr.c:219-246, SRC/miscla.c:211-240) as batched
//     tall-skinny GEMV-T / GEMV-N over chunked Lanczos bases (FP64 or FP32
//     storage, FP64 sums).
// Filter launches map blockIdx.x to row blocks XCD by XCD (xcd_block), so
// that each of the 8 XCDs, with its own L2, owns one contiguous range of rows.
// Support: Sturm-multisection restricted trace of the Lanczos tridiagonal
// matrix, the closed-form DOS integral of the slicer, random fills, the
// vector part of a generalized filter step (k_cheb_update), a plain copy
// kernel (the bandwidth and counter reference).
//
// Every reduction is two-stage with a fixed order, so results are
// deterministic for a given launch shape.  The filter kernels disable FMA
// contraction so that, with one lane per row, each element follows the same
// rounding sequence as the reference's scalar loops.
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <algorithm>
#include <type_traits>

namespace pevsl {

namespace {

constexpr int kThreads = 256;
// MI355X Infinity Cache (memory-side last-level cache)
constexpr double kInfinityCacheBytes = 256.0 * 1024 * 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the lanes of a group of L consecutive lanes.
template <int L>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
  return v;
}
template <int L>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
  return v;
}

typedef double dvec2 __attribute__((ext_vector_type(2)));

// streaming loads of the Lanczos basis: no reuse, keep the caches for the
// matrix and the filter vectors
template <bool NT>
__device__ __forceinline__ double2 ld_basis(const double2 *p) {
  if (NT) {
    const dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p));
    return make_double2(v.x, v.y);
  }
  return *p;
}

typedef float fvec4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 ld_basis_nt(const float4 *p) {
  const fvec4 v = __builtin_nontemporal_load(reinterpret_cast<const fvec4 *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// 16-byte row groups of the Lanczos basis: R rows of type TV per load, kept
// in the loaded type and widened to FP64 where they are used.  Launch shapes
// per element type keep the rows per block (and so the grids) of the two
// storage types equal.
template <typename TV>
struct Row16;
template <>
struct Row16<double> {
  using vec = double2;
  static constexpr int R = 2;
  static constexpr int kTiters = 16;   // GEMV-T row groups per thread
  static constexpr int kGnPairs = 2;   // GEMV-N row groups per thread
  static constexpr int kGnUnroll = 8;  // GEMV-N columns per load batch
  template <bool NT>
  __device__ static __forceinline__ vec ld(const vec *p) { return ld_basis<NT>(p); }
  __device__ static __forceinline__ double get(const vec &v, int r) { return r ? v.y : v.x; }
  __device__ static __forceinline__ vec zero() { return make_double2(0.0, 0.0); }
  __device__ static __forceinline__ void store(vec *p, const double (&o)[R]) {
    *p = make_double2(o[0], o[1]);
  }
};
template <>
struct Row16<float> {
  using vec = float4;
  static constexpr int R = 4;
  static constexpr int kTiters = 8;
  static constexpr int kGnPairs = 1;
  static constexpr int kGnUnroll = 16;
  template <bool NT>
  __device__ static __forceinline__ vec ld(const vec *p) { return NT ? ld_basis_nt(p) : *p; }
  __device__ static __forceinline__ double get(const vec &v, int r) {
    return r == 0 ? v.x : (r == 1 ? v.y : (r == 2 ? v.z : v.w));
  }
  __device__ static __forceinline__ vec zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
  __device__ static __forceinline__ void store(vec *p, const double (&o)[R]) {
    *p = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
  }
};

// Row-block index of this workgroup such that each XCD works on one
// contiguous range of row blocks.  Workgroups are dispatched to the 8 XCDs
// round-robin (workgroup i to XCD i mod 8), and each XCD has its own L2: with
// contiguous ranges the neighbour rows that a stencil-like matrix gathers
// (+-1, +-nx, +-nx*ny) mostly sit in the same L2.  A bijection of [0, nb).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int bid, int nb) {
  const int q = nb / kXcds, r = nb % kXcds;
  const int x = bid % kXcds, j = bid / kXcds;
  return x * q + min(x, r) + j;
}

// Column of nonzero j in a row block that starts at row r0: the 32-bit index,
// or (J16) r0 plus the nonzero's 16-bit offset from r0 (DevCsr::jo16), which
// saves 2 of the 4 index bytes per nonzero.
template <bool J16>
__device__ __forceinline__ int block_col(const int *__restrict__ ja, const short *__restrict__ jo,
                                         int r0, int j) {
  if constexpr (J16)
    return r0 + (int)jo[j];
  else
    return ja[j];
}

inline int grid_for(size_t work, int threads = kThreads) {
  return (int)((work + threads - 1) / threads);
}

// ---------------------------------------------------------------------------
// Sparse products
// ---------------------------------------------------------------------------

// y = A x, L lanes per row (rows of a wave: 64/L).
template <int L>
__global__ void __launch_bounds__(kThreads)
k_spmv_f64(int n, const int *__restrict__ ia, const int *__restrict__ ja,
           const double *__restrict__ a, const double *__restrict__ x,
           double *__restrict__ y, size_t ld, int ncol) {
#pragma clang fp contract(off)
  const int col = blockIdx.y;
  if (col >= ncol) return;
  x += (size_t)col * ld;
  y += (size_t)col * ld;
  const size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const int row = (int)(g / L);
  const int lane = (int)(g % L);
  double acc = 0.0;
  if (row < n) {
    const int e = ia[row + 1];
    for (int j = ia[row] + lane; j < e; j += L) acc += a[j] * x[ja[j]];
  }
  if (L > 1) acc = group_sum<L>(acc);
  if (row < n && lane == 0) y[row] = acc;
}

template <typename T, int L>
__global__ void __launch_bounds__(kThreads)
k_cheb_vec(int n, const int *__restrict__ ia, const int *__restrict__ ja,
           const T *__restrict__ a, const T *__restrict__ vk,
           const T *__restrict__ vkm1, T *__restrict__ vkp1,
           T *__restrict__ y, T cc, T t, T mu0, T muk, int first) {
#pragma clang fp contract(off)
  const size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const int row = (int)(g / L);
  const int lane = (int)(g % L);
  // all lanes of a row group exit together, so the group shuffles below
  // only involve active lanes
  if (row >= n) return;
  const int b = ia[row], e = ia[row + 1];
  // operands of the update are loaded before the row sum so that their
  // latency overlaps the gather
  const T v = vk[row];
  const T vm = first ? (T)0 : vkm1[row];
  const T yv = first ? (T)0 : y[row];
  T acc = 0;
  for (int j = b + lane; j < e; j += L) acc += a[j] * vk[ja[j]];
  if (L > 1) acc = group_sum<L>(acc);
  if (lane == 0) {
    T w = acc + (-cc) * v;  // A vk - cc vk
    w = w * t;
    if (first) {
      vkp1[row] = w;
      y[row] = v * mu0 + muk * w;
    } else {
      w = w + (T)(-1) * vm;
      vkp1[row] = w;
      y[row] = yv + muk * w;
    }
  }
}

// CSR-stream step: the block's nonzeros are multiplied with the gathered
// x values in coalesced order and staged in LDS; each thread then sums one
// row sequentially (the reference's summation order) and applies the
// recurrence.  A row longer than kStreamNnz forms a block alone and is
// reduced by the whole block.
template <typename T, bool J16>
__global__ void __launch_bounds__(kThreads)
k_cheb_stream(const int *__restrict__ rowblk, const int *__restrict__ ia,
              const int *__restrict__ ja, const short *__restrict__ jo, const T *__restrict__ a,
              const T *__restrict__ vk, const T *__restrict__ vkm1,
              T *__restrict__ vkp1, T *__restrict__ y, T cc, T t, T mu0, T muk,
              int first) {
#pragma clang fp contract(off)
  __shared__ T prod[kStreamNnz];
  const int rb = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = rowblk[rb], r1 = rowblk[rb + 1];
  const int j0 = ia[r0], nz = ia[r1] - j0;
  T acc = 0;
  int r = r0 + threadIdx.x;
  // the row's own operands and row pointers are loaded before the gathers,
  // so that their latency overlaps the staging (100^3 Laplacian, 16-bit
  // offsets: FP32 15.56 against 16.27 us per step, FP64 19.66 against 19.69)
  T v = 0, vm = 0, yv = 0;
  int qs = 0, qe = 0;
  if (r < r1) {
    v = vk[r];
    if (!first) {
      vm = vkm1[r];
      yv = y[r];
    }
    qs = ia[r] - j0;
    qe = ia[r + 1] - j0;
  }
  if (nz <= kStreamNnz) {
    for (int q = threadIdx.x; q < nz; q += kThreads)
      prod[q] = a[j0 + q] * vk[block_col<J16>(ja, jo, r0, j0 + q)];
    __syncthreads();
    if (r >= r1) return;
    for (int q = qs; q < qe; q++) acc += prod[q];
  } else {
    // single long row
    for (int q = threadIdx.x; q < nz; q += kThreads)
      acc += a[j0 + q] * vk[block_col<J16>(ja, jo, r0, j0 + q)];
    __shared__ T red[kThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x != 0) return;
    acc = red[0];  // thread 0's operands are those of the block's one row r0
  }
  T w = acc + (-cc) * v;
  w = w * t;
  if (first) {
    vkp1[r] = w;
    y[r] = v * mu0 + muk * w;
  } else {
    w = w + (T)(-1) * vm;
    vkp1[r] = w;
    y[r] = yv + muk * w;
  }
}

// B interleaved columns; lane = (row_in_block, column).
template <typename T, int B>
__global__ void __launch_bounds__(kThreads)
k_cheb_blk(int n, const int *__restrict__ ia, const int *__restrict__ ja,
           const T *__restrict__ a, const T *__restrict__ vk,
           const T *__restrict__ vkm1, T *__restrict__ vkp1,
           T *__restrict__ y, T cc, T t, const double *__restrict__ mus,
           int k, int nact) {
#pragma clang fp contract(off)
  const size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const int row = (int)(g / B);
  const int c = (int)(g % B);
  if (row >= n || c >= nact) return;
  T acc = 0;
  const int e = ia[row + 1];
  for (int j = ia[row]; j < e; j++) acc += a[j] * vk[(size_t)ja[j] * B + c];
  const size_t idx = (size_t)row * B + c;
  const T v = vk[idx];
  T w = acc + (-cc) * v;
  w = w * t;
  const T muk = (T)mus[(size_t)k * B + c];
  if (k == 1) {
    vkp1[idx] = w;
    y[idx] = v * (T)mus[c] + muk * w;
  } else {
    w = w + (T)(-1) * vkm1[idx];
    vkp1[idx] = w;
    y[idx] = y[idx] + muk * w;
  }
}

// Row-per-thread block step for B interleaved columns: the row block's
// a/ja are staged in LDS (coalesced), then each thread owns one row and
// gathers whole B-wide rows of vk with 16-byte loads, so one memory
// instruction moves B values instead of one.  Columns >= nact keep y.
template <typename T, int B>
__device__ __forceinline__ void load_row(const T *__restrict__ p, T (&x)[B]) {
  if constexpr ((B * sizeof(T)) % 16 == 0) {
    const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
    for (int i = 0; i < (int)(B * sizeof(T) / 16); i++)
      reinterpret_cast<float4 *>(x)[i] = q[i];
  } else {
#pragma unroll
    for (int c = 0; c < B; c++) x[c] = p[c];
  }
}
template <typename T, int B>
__device__ __forceinline__ void store_row(T *__restrict__ p, const T (&x)[B]) {
  if constexpr ((B * sizeof(T)) % 16 == 0) {
    float4 *q = reinterpret_cast<float4 *>(p);
#pragma unroll
    for (int i = 0; i < (int)(B * sizeof(T) / 16); i++)
      q[i] = reinterpret_cast<const float4 *>(x)[i];
  } else {
#pragma unroll
    for (int c = 0; c < B; c++) p[c] = x[c];
  }
}

// 16-byte pieces that are read or written once per launch (matrix, v, b2,
// out): nontemporal, so that they do not evict the gathered rows of b1 from
// L2 between their reuses (neighbours nx*ny rows apart on a 3-D grid)
template <typename T, int W>
__device__ __forceinline__ void load_row_nt(const T *__restrict__ p, T (&x)[W]) {
  if constexpr ((W * sizeof(T)) % 16 == 0) {
#pragma unroll
    for (int i = 0; i < (int)(W * sizeof(T) / 16); i++) {
      const fvec4 v = __builtin_nontemporal_load(reinterpret_cast<const fvec4 *>(p) + i);
      reinterpret_cast<fvec4 *>(x)[i] = v;
    }
  } else {
#pragma unroll
    for (int c = 0; c < W; c++) x[c] = __builtin_nontemporal_load(p + c);
  }
}
template <typename T, int W>
__device__ __forceinline__ void store_row_nt(T *__restrict__ p, const T (&x)[W]) {
  if constexpr ((W * sizeof(T)) % 16 == 0) {
#pragma unroll
    for (int i = 0; i < (int)(W * sizeof(T) / 16); i++)
      __builtin_nontemporal_store(reinterpret_cast<const fvec4 *>(x)[i],
                                  reinterpret_cast<fvec4 *>(p) + i);
  } else {
#pragma unroll
    for (int c = 0; c < W; c++) __builtin_nontemporal_store(x[c], p + c);
  }
}

template <typename T, int B, bool J16>
__global__ void __launch_bounds__(kThreads)
k_cheb_blkrow(const int *__restrict__ rowblk, const int *__restrict__ ia,
              const int *__restrict__ ja, const short *__restrict__ jo, const T *__restrict__ a,
              const T *__restrict__ vk, const T *__restrict__ vkm1,
              T *__restrict__ vkp1, T *__restrict__ y, T cc, T t,
              const double *__restrict__ mus, int k, int nact, int ymode) {
#pragma clang fp contract(off)
  __shared__ T as[kStreamNnz];
  __shared__ int js[kStreamNnz];
  const int rb = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = rowblk[rb], r1 = rowblk[rb + 1];
  const int j0 = ia[r0], nz = ia[r1] - j0;
  const bool staged = nz <= kStreamNnz;
  if (staged) {
    for (int q = threadIdx.x; q < nz; q += kThreads) {
      as[q] = a[j0 + q];
      js[q] = block_col<J16>(ja, jo, r0, j0 + q);
    }
    __syncthreads();
  } else if (threadIdx.x != 0) {
    return;  // one long row: handled by thread 0 from global memory
  }
  const int r = r0 + threadIdx.x;
  if (r >= r1) return;
  T acc[B];
#pragma unroll
  for (int c = 0; c < B; c++) acc[c] = 0;
  const int qs = ia[r] - j0, qe = ia[r + 1] - j0;
  for (int q = qs; q < qe; q++) {
    const T av = staged ? as[q] : a[j0 + q];
    const int col = staged ? js[q] : block_col<J16>(ja, jo, r0, j0 + q);
    T x[B];
    load_row<T, B>(vk + (size_t)col * B, x);
#pragma unroll
    for (int c = 0; c < B; c++) acc[c] += av * x[c];
  }
  T v[B], vm[B], yy[B], w[B];
  load_row<T, B>(vk + (size_t)r * B, v);
  if (ymode != kYSkip) load_row<T, B>(y + (size_t)r * B, yy);
  if (k == 1) {
#pragma unroll
    for (int c = 0; c < B; c++) {
      w[c] = acc[c] + (-cc) * v[c];
      w[c] = w[c] * t;
      if (c < nact) yy[c] = v[c] * (T)mus[c] + (T)mus[(size_t)B + c] * w[c];
    }
  } else {
    load_row<T, B>(vkm1 + (size_t)r * B, vm);
#pragma unroll
    for (int c = 0; c < B; c++) {
      w[c] = acc[c] + (-cc) * v[c];
      w[c] = w[c] * t;
      w[c] = w[c] + (T)(-1) * vm[c];
      if (c < nact) {
        // kYPair: the term of degree k-1 (v, this step's input) that the
        // previous step skipped, then degree k: the two updates of two plain
        // steps, in their order
        if (ymode == kYPair) yy[c] = yy[c] + (T)mus[(size_t)(k - 1) * B + c] * v[c];
        if (ymode != kYSkip) yy[c] = yy[c] + (T)mus[(size_t)k * B + c] * w[c];
      }
    }
  }
  store_row<T, B>(vkp1 + (size_t)r * B, w);
  if (ymode != kYSkip) store_row<T, B>(y + (size_t)r * B, yy);
}

// One Clenshaw step of the filter on B interleaved columns (row per thread,
// a/ja staged in LDS as in k_cheb_blkrow):
//   out = mu_k v + t (A b1 - cc b1) - b2      (b2 may be null: zero)
// with per-column mu_k = mus[k*B + c].  Run from the top degree down with
// t = 2/dd it produces the Clenshaw sums b_k of p((A-cc)/dd) v =
// sum_k mu_k T_k((A-cc)/dd) v; the last step (k = 0, t = 1/dd) gives the
// filtered vector itself.  Per step it gathers b1 and streams b2, v and out:
// four vector passes where the three-term form with an accumulator needs
// five.
template <typename T, int B, bool J16>
__global__ void __launch_bounds__(kThreads)
k_clen_blkrow(const int *__restrict__ rowblk, const int *__restrict__ ia,
              const int *__restrict__ ja, const short *__restrict__ jo, const T *__restrict__ a,
              const T *__restrict__ b1, const T *__restrict__ b2,
              const T *__restrict__ v, T *__restrict__ out, T cc, T t,
              const double *__restrict__ mus, int k) {
#pragma clang fp contract(off)
  __shared__ T as[kStreamNnz];
  __shared__ int js[kStreamNnz];
  const int rb = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = rowblk[rb], r1 = rowblk[rb + 1];
  const int j0 = ia[r0], nz = ia[r1] - j0;
  const bool staged = nz <= kStreamNnz;
  if (staged) {
    for (int q = threadIdx.x; q < nz; q += kThreads) {
      as[q] = a[j0 + q];
      js[q] = block_col<J16>(ja, jo, r0, j0 + q);
    }
    __syncthreads();
  } else if (threadIdx.x != 0) {
    return;  // one long row: handled by thread 0 from global memory
  }
  const int r = r0 + threadIdx.x;
  if (r >= r1) return;
  T acc[B];
#pragma unroll
  for (int c = 0; c < B; c++) acc[c] = 0;
  const int qs = ia[r] - j0, qe = ia[r + 1] - j0;
  for (int q = qs; q < qe; q++) {
    const T av = staged ? as[q] : a[j0 + q];
    const int col = staged ? js[q] : block_col<J16>(ja, jo, r0, j0 + q);
    T x[B];
    load_row<T, B>(b1 + (size_t)col * B, x);
#pragma unroll
    for (int c = 0; c < B; c++) acc[c] += av * x[c];
  }
  T bb[B], vv[B], w[B], b2r[B];
  load_row<T, B>(b1 + (size_t)r * B, bb);
  load_row<T, B>(v + (size_t)r * B, vv);
  if (b2) {
    load_row<T, B>(b2 + (size_t)r * B, b2r);
  } else {
#pragma unroll
    for (int c = 0; c < B; c++) b2r[c] = 0;
  }
#pragma unroll
  for (int c = 0; c < B; c++) {
    T z = acc[c] + (-cc) * bb[c];
    z = z * t;
    z = z + (T)mus[(size_t)k * B + c] * vv[c];
    w[c] = z + (T)(-1) * b2r[c];
  }
  store_row<T, B>(out + (size_t)r * B, w);
}

// One Clenshaw step on the active pieces of an interleaved block of Bp
// columns (row stride Bp), one thread per (row, 16-byte piece of SEG
// columns): the columns of a wide batch share one pass over the matrix while
// each piece carries its own degree range.  Piece s is active while
// d <= tops.top[s] (its highest degree); at d == top it writes the top term
// mu_d v (no product), at d == top - 1 it reads b2 as zero, below that
//   out = mu_d v + t (A b1 - cc b1) - b2
// exactly as k_clen_blkrow computes each column (same operations and order,
// hence the same bits).  Inactive pieces are neither read nor written: the
// host launches the active prefix (pieces sorted by top degree, descending)
// over a row partition with at most 256/Q rows and 2048/Q nonzeros per
// block, Q >= nseg, so that every thread owns at most one (row, piece).
// NT: the matrix, v, b2 and out pass through as nontemporal accesses, which
// keeps L2 for the gathered rows of b1 when a launch's data exceeds the
// Infinity Cache (measured on the 100^3 Laplacian, 32 columns: 149.6 ->
// 132.1 us; at 8 columns, whose 187 MB stay cached from launch to launch,
// the hint costs 30%: 32.6 -> 42.5 us).
template <typename T, bool NT>
__global__ void __launch_bounds__(kThreads)
k_clen_seg(const int *__restrict__ rowblk, const int *__restrict__ ia,
           const int *__restrict__ ja, const T *__restrict__ a,
           const T *__restrict__ b1, const T *__restrict__ b2,
           const T *__restrict__ v, T *__restrict__ out, T cc, T t,
           const double *__restrict__ mus, int d, int Bp, int nseg, int maxnz,
           SegTops tops) {
#pragma clang fp contract(off)
  constexpr int SEG = 16 / (int)sizeof(T);
  __shared__ T as[kStreamNnz];
  __shared__ int js[kStreamNnz];
  const int rb = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = rowblk[rb], r1 = rowblk[rb + 1];
  const int j0 = ia[r0], nz = ia[r1] - j0;
  // pieces that need the product this step (top > d: piece 0 has the
  // highest top)
  const bool gather = tops.top[0] > d;
  const bool staged = gather && nz <= maxnz;
  if (staged) {
    for (int q = threadIdx.x; q < nz; q += kThreads) {
      as[q] = NT ? __builtin_nontemporal_load(a + j0 + q) : a[j0 + q];
      js[q] = NT ? __builtin_nontemporal_load(ja + j0 + q) : ja[j0 + q];
    }
    __syncthreads();
  }
  const int lr = (int)threadIdx.x / nseg;
  const int sg = (int)threadIdx.x - lr * nseg;
  const int r = r0 + lr;
  if (r >= r1) return;
  const int top = tops.top[sg];
  const int c0 = sg * SEG;
  const size_t off = (size_t)r * Bp + c0;
  const double *mu = mus + (size_t)d * Bp + c0;
  T vv[SEG], w[SEG];
  if (NT)
    load_row_nt<T, SEG>(v + off, vv);
  else
    load_row<T, SEG>(v + off, vv);
  if (d == top) {
#pragma unroll
    for (int c = 0; c < SEG; c++) w[c] = (T)mu[c] * vv[c];
  } else {
    T acc[SEG];
#pragma unroll
    for (int c = 0; c < SEG; c++) acc[c] = 0;
    const int qs = ia[r] - j0, qe = ia[r + 1] - j0;
    for (int q = qs; q < qe; q++) {
      const T av = staged ? as[q] : a[j0 + q];
      const int col = staged ? js[q] : ja[j0 + q];
      T x[SEG];
      load_row<T, SEG>(b1 + (size_t)col * Bp + c0, x);
#pragma unroll
      for (int c = 0; c < SEG; c++) acc[c] += av * x[c];
    }
    T bb[SEG], b2r[SEG];
    load_row<T, SEG>(b1 + off, bb);
    if (b2 && d < top - 1) {
      if (NT)
        load_row_nt<T, SEG>(b2 + off, b2r);
      else
        load_row<T, SEG>(b2 + off, b2r);
    } else {
#pragma unroll
      for (int c = 0; c < SEG; c++) b2r[c] = 0;
    }
#pragma unroll
    for (int c = 0; c < SEG; c++) {
      T z = acc[c] + (-cc) * bb[c];
      z = z * t;
      z = z + (T)mu[c] * vv[c];
      w[c] = z + (T)(-1) * b2r[c];
    }
  }
  if (NT)
    store_row_nt<T, SEG>(out + off, w);
  else
    store_row<T, SEG>(out + off, w);
}

// out[r*B + c] = mus[k*B + c] * v[r*B + c]  (top Clenshaw term)
template <typename T>
__global__ void k_blk_colscale(size_t n, int B, const T *__restrict__ v, T *__restrict__ out,
                               const double *__restrict__ mus, int k) {
  const size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (g >= n * (size_t)B) return;
  out[g] = (T)mus[(size_t)k * B + (int)(g % B)] * v[g];
}

// ---- nnz-balanced block products -------------------------------------------
// Block SpMM for B interleaved columns on matrices whose row lengths are very
// uneven (power-law graphs), where a thread that walks a whole row stalls its
// wave for the longest row.  Per row block:
//   phase 1: one thread per (nonzero, 16-byte piece of the B-wide row of x):
//            every product a_ij * x_j(piece) is formed in parallel, with
//            independent gathers, and staged in LDS;
//   phase 2: one thread per (row, piece) sums the row's staged products in
//            column order and runs the fused epilogue on its piece.
// A row with more nonzeros than a block stages is a block alone and is reduced
// by the whole workgroup (shuffles, then LDS across waves).  Products are
// rounded before they are added and staged rows are summed in column order,
// so those rows round exactly as in the row-per-thread kernels.
//
// Budget: NZ = kStreamNnz / Q staged products of B values (32 KiB of LDS when
// a piece is 16 bytes) and at most kStreamRows / Q rows per block, Q = pieces
// per row; the host builds one row partition per Q (DevCsr::rowblk_q).
template <typename T, int B>
struct Pieces {
  static constexpr int P = (B * (int)sizeof(T) >= 16) ? 16 / (int)sizeof(T) : B;
  static constexpr int Q = B / P;
  static constexpr int NZ = kStreamNnz / Q;
};

template <typename T, int P>
__device__ __forceinline__ void ld_piece(const T *p, T (&x)[P]) {
  if constexpr (P * sizeof(T) == 16) {
    *reinterpret_cast<float4 *>(x) = *reinterpret_cast<const float4 *>(p);
  } else if constexpr (P * sizeof(T) == 8) {
    *reinterpret_cast<float2 *>(x) = *reinterpret_cast<const float2 *>(p);
  } else {
#pragma unroll
    for (int e = 0; e < P; e++) x[e] = p[e];
  }
}
template <typename T, int P>
__device__ __forceinline__ void st_piece(T *p, const T (&x)[P]) {
  if constexpr (P * sizeof(T) == 16) {
    *reinterpret_cast<float4 *>(p) = *reinterpret_cast<const float4 *>(x);
  } else if constexpr (P * sizeof(T) == 8) {
    *reinterpret_cast<float2 *>(p) = *reinterpret_cast<const float2 *>(x);
  } else {
#pragma unroll
    for (int e = 0; e < P; e++) p[e] = x[e];
  }
}

// Epilogue of the three-term step (same operations as k_cheb_blkrow) on
// columns [piece*P, piece*P + P) of row r.
template <typename T, int B>
struct EpiCheb {
  const T *vk, *vkm1;
  T *vkp1, *y;
  T cc, t;
  const double *mus;
  int k, nact, ymode;
  __device__ __forceinline__ void operator()(int r, int piece,
                                             const T (&acc)[Pieces<T, B>::P]) const {
#pragma clang fp contract(off)
    constexpr int P = Pieces<T, B>::P;
    const size_t o = (size_t)r * B + piece * P;
    T v[P], yy[P], w[P];
    ld_piece<T, P>(vk + o, v);
    if (ymode != kYSkip) ld_piece<T, P>(y + o, yy);
    if (k == 1) {
#pragma unroll
      for (int e = 0; e < P; e++) {
        const int c = piece * P + e;
        w[e] = acc[e] + (-cc) * v[e];
        w[e] = w[e] * t;
        if (c < nact) yy[e] = v[e] * (T)mus[c] + (T)mus[(size_t)B + c] * w[e];
      }
    } else {
      T vm[P];
      ld_piece<T, P>(vkm1 + o, vm);
#pragma unroll
      for (int e = 0; e < P; e++) {
        const int c = piece * P + e;
        w[e] = acc[e] + (-cc) * v[e];
        w[e] = w[e] * t;
        w[e] = w[e] + (T)(-1) * vm[e];
        if (c < nact) {
          if (ymode == kYPair) yy[e] = yy[e] + (T)mus[(size_t)(k - 1) * B + c] * v[e];
          if (ymode != kYSkip) yy[e] = yy[e] + (T)mus[(size_t)k * B + c] * w[e];
        }
      }
    }
    st_piece<T, P>(vkp1 + o, w);
    if (ymode != kYSkip) st_piece<T, P>(y + o, yy);
  }
};

// Epilogue of the Clenshaw step (same operations as k_clen_blkrow); b1 is the
// gathered block.
template <typename T, int B>
struct EpiClen {
  const T *b1, *b2, *v;
  T *out;
  T cc, t;
  const double *mus;
  int k;
  __device__ __forceinline__ void operator()(int r, int piece,
                                             const T (&acc)[Pieces<T, B>::P]) const {
#pragma clang fp contract(off)
    constexpr int P = Pieces<T, B>::P;
    const size_t o = (size_t)r * B + piece * P;
    T bb[P], vv[P], b2r[P], w[P];
    ld_piece<T, P>(b1 + o, bb);
    ld_piece<T, P>(v + o, vv);
    if (b2) {
      ld_piece<T, P>(b2 + o, b2r);
    } else {
#pragma unroll
      for (int e = 0; e < P; e++) b2r[e] = 0;
    }
#pragma unroll
    for (int e = 0; e < P; e++) {
      T z = acc[e] + (-cc) * bb[e];
      z = z * t;
      z = z + (T)mus[(size_t)k * B + piece * P + e] * vv[e];
      w[e] = z + (T)(-1) * b2r[e];
    }
    st_piece<T, P>(out + o, w);
  }
};

// U = NZ / S = 8 products per thread and block: each thread first issues
// its 8 index and value loads (nontemporal: the matrix streams once per step
// and should not evict the gathered block from the caches), then its 8
// gathers, so that 8 gathers per thread are in flight at once.
constexpr int kBalUnroll = 8;

// Both phases for row block rb: calls f(row, piece, acc) once per (row,
// piece) of the block with the row's product sums for that piece, from the
// thread threadIdx.x with threadIdx.x % Q == piece (phase 2 covers each
// (row, piece) with one thread, the long-row path uses threads 0..Q-1).
// prod: the block's LDS staging (NZ * B values).
template <typename T, int B, class F>
__device__ __forceinline__ void bal_rows(int rb, const int *__restrict__ rowblk,
                                         const int *__restrict__ ia, const int *__restrict__ ja,
                                         const T *__restrict__ a, const T *__restrict__ x,
                                         T *prod, F &&f) {
#pragma clang fp contract(off)
  using PC = Pieces<T, B>;
  constexpr int P = PC::P, Q = PC::Q, S = kThreads / Q, U = kBalUnroll;
  static_assert(PC::NZ == U * S, "one unrolled batch covers a staged block");
  const int r0 = rowblk[rb], r1 = rowblk[rb + 1];
  const int j0 = ia[r0], nz = ia[r1] - j0;
  const int piece = threadIdx.x % Q, slot = threadIdx.x / Q;
  // products q = q0 + slot + i*S, i < U, of the nonzeros j0 + q < j0 + nz
  auto batch = [&](int q0, T (&xv)[U][P], T (&av)[U]) {
    int col[U];
#pragma unroll
    for (int i = 0; i < U; i++) {
      const int q = q0 + slot + i * S;
      col[i] = 0;
      av[i] = 0;
      if (q < nz) {
        col[i] = __builtin_nontemporal_load(ja + j0 + q);
        av[i] = __builtin_nontemporal_load(a + j0 + q);
      }
    }
#pragma unroll
    for (int i = 0; i < U; i++) {
      if (q0 + slot + i * S < nz) ld_piece<T, P>(x + (size_t)col[i] * B + piece * P, xv[i]);
    }
  };
  if (nz > PC::NZ) {
    // one long row (the host partition never puts it with others)
    T acc[P];
#pragma unroll
    for (int e = 0; e < P; e++) acc[e] = 0;
    for (int q0 = 0; q0 < nz; q0 += U * S) {
      T xv[U][P], av[U];
      batch(q0, xv, av);
#pragma unroll
      for (int i = 0; i < U; i++) {
        if (q0 + slot + i * S < nz) {
#pragma unroll
          for (int e = 0; e < P; e++) acc[e] += av[i] * xv[i][e];
        }
      }
    }
    // lanes of a wave that hold the same piece: lane % Q == piece
#pragma unroll
    for (int off = 32; off >= Q; off >>= 1) {
#pragma unroll
      for (int e = 0; e < P; e++) acc[e] += __shfl_xor(acc[e], off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < Q) {
#pragma unroll
      for (int e = 0; e < P; e++) prod[wave * B + piece * P + e] = acc[e];
    }
    __syncthreads();
    if (threadIdx.x < Q) {
#pragma unroll
      for (int e = 0; e < P; e++) {
        T sum = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; w++) sum += prod[w * B + piece * P + e];
        acc[e] = sum;
      }
      f(r0, piece, acc);
    }
    return;
  }
  {
    T xv[U][P], av[U];
    batch(0, xv, av);
#pragma unroll
    for (int i = 0; i < U; i++) {
      const int q = slot + i * S;
      if (q < nz) {
#pragma unroll
        for (int e = 0; e < P; e++) xv[i][e] = av[i] * xv[i][e];
        st_piece<T, P>(prod + (size_t)q * B + piece * P, xv[i]);
      }
    }
  }
  __syncthreads();
  const int nrq = (r1 - r0) * Q;
  for (int u = threadIdx.x; u < nrq; u += kThreads) {
    const int r = r0 + u / Q, pc = u % Q;
    const int qs = ia[r] - j0, qe = ia[r + 1] - j0;
    T acc[P];
#pragma unroll
    for (int e = 0; e < P; e++) acc[e] = 0;
    for (int q = qs; q < qe; q++) {
      T pv[P];
      ld_piece<T, P>(prod + (size_t)q * B + pc * P, pv);
#pragma unroll
      for (int e = 0; e < P; e++) acc[e] += pv[e];
    }
    f(r, pc, acc);
  }
}

template <typename T, int B, class Epi>
__global__ void __launch_bounds__(kThreads)
k_blk_bal(const int *__restrict__ rowblk, const int *__restrict__ ia,
          const int *__restrict__ ja, const T *__restrict__ a, const T *__restrict__ x,
          Epi epi) {
  __shared__ T prod[Pieces<T, B>::NZ * B];
  bal_rows<T, B>(xcd_block(blockIdx.x, gridDim.x), rowblk, ia, ja, a, x, prod,
                 [&](int r, int pc, const T (&acc)[Pieces<T, B>::P]) { epi(r, pc, acc); });
}

// KPM step on B interleaved vectors with the nnz-balanced products (uneven
// rows): vkp1 = scal (A vk - ctr vk) - vkm1 and, per column, the partial dot
// of vkp1 with v0 over the row block, part[rb * B + c] (rb = row block, so
// the second-stage sum runs in row order whatever the dispatch order).
template <typename T, int B>
__global__ void __launch_bounds__(kThreads)
k_kpm_bal(const int *__restrict__ rowblk, const int *__restrict__ ia,
          const int *__restrict__ ja, const T *__restrict__ a, const T *__restrict__ vk,
          const T *__restrict__ vkm1, T *__restrict__ vkp1, const T *__restrict__ v0, T ctr,
          T scal, double *__restrict__ part) {
#pragma clang fp contract(off)
  using PC = Pieces<T, B>;
  constexpr int P = PC::P, Q = PC::Q;
  __shared__ T prod[PC::NZ * B];
  const int rb = xcd_block(blockIdx.x, gridDim.x);
  double dsum[P];
#pragma unroll
  for (int e = 0; e < P; e++) dsum[e] = 0.0;
  bal_rows<T, B>(rb, rowblk, ia, ja, a, vk, prod, [&](int r, int pc, const T (&acc)[P]) {
    const size_t o = (size_t)r * B + pc * P;
    T v[P], vm[P], z[P], w[P];
    ld_piece<T, P>(vk + o, v);
    ld_piece<T, P>(vkm1 + o, vm);
    ld_piece<T, P>(v0 + o, z);
#pragma unroll
    for (int e = 0; e < P; e++) {
      w[e] = acc[e] + (-ctr) * v[e];
      w[e] = w[e] * scal;
      w[e] = w[e] + (T)(-1) * vm[e];
      dsum[e] += (double)w[e] * (double)z[e];
    }
    st_piece<T, P>(vkp1 + o, w);
  });
  // per column: sum over the threads holding its piece, in thread order
  __syncthreads();
  double *red = reinterpret_cast<double *>(prod);
  static_assert(kThreads * P * sizeof(double) <= PC::NZ * B * sizeof(T), "LDS reuse");
#pragma unroll
  for (int e = 0; e < P; e++) red[threadIdx.x * P + e] = dsum[e];
  __syncthreads();
  if (threadIdx.x < B) {
    const int c = threadIdx.x, pc = c / P, e = c % P;
    double sum = 0.0;
    for (int t = pc; t < kThreads; t += Q) sum += red[t * P + e];
    part[(size_t)rb * B + c] = sum;
  }
}

// KPM step + per-column partial dot with v0 over the block's rows.
template <typename T, int B>
__global__ void __launch_bounds__(kThreads)
k_kpm_blk(int n, const int *__restrict__ ia, const int *__restrict__ ja,
          const T *__restrict__ a, const T *__restrict__ vk,
          const T *__restrict__ vkm1, T *__restrict__ vkp1,
          const T *__restrict__ v0, T ctr, T scal, double *__restrict__ part,
          int rows_per_block) {
#pragma clang fp contract(off)
  constexpr int RPW = 64 / B;  // rows per wave per sweep
  __shared__ double red[kThreads / 64][B];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane % B;
  const int rsub = lane / B;
  const int row0 = blockIdx.x * rows_per_block;
  const int row_end = min(n, row0 + rows_per_block);
  double dsum = 0.0;
  for (int row = row0 + wave * RPW + rsub; row < row_end;
       row += (kThreads / 64) * RPW) {
    T acc = 0;
    const int e = ia[row + 1];
    for (int j = ia[row]; j < e; j++) acc += a[j] * vk[(size_t)ja[j] * B + c];
    const size_t idx = (size_t)row * B + c;
    T w = acc + (-ctr) * vk[idx];
    w = w * scal;
    w = w + (T)(-1) * vkm1[idx];
    vkp1[idx] = w;
    dsum += (double)w * (double)v0[idx];
  }
  // reduce over the rows held by this wave (lanes with equal c)
#pragma unroll
  for (int off = 32; off >= B; off >>= 1) dsum += __shfl_xor(dsum, off, 64);
  if (rsub == 0) red[wave][c] = dsum;
  __syncthreads();
  if (threadIdx.x < B) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) s += red[w][threadIdx.x];
    part[(size_t)blockIdx.x * B + threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// Vector kernels
// ---------------------------------------------------------------------------

constexpr int kDotBlocksMax = 1024;

// out[c] = sum_b part[b*ncol + c]; one block per column, fixed order
__global__ void __launch_bounds__(kThreads)
k_reduce_partials(const double *__restrict__ part, int nblk, int ncol,
                  double *__restrict__ out) {
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kThreads) s += part[(size_t)b * ncol + c];
  __shared__ double red[kThreads / 64];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; w++) t += red[w];
    out[c] = t;
  }
}

// Block-level sum of v into part[blockIdx.x] (all threads participate).
__device__ __forceinline__ void block_sum_store(double v, double *dst) {
  __shared__ double red[kThreads / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; w++) s += red[w];
    *dst = s;
  }
}

__global__ void __launch_bounds__(kThreads)
k_dot_part(int n, const double *__restrict__ x, const double *__restrict__ y,
           double *__restrict__ part) {
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < (size_t)n;
       i += (size_t)gridDim.x * kThreads)
    acc += x[i] * y[i];
  block_sum_store(acc, part + blockIdx.x);
}

// y = x over n2 16-byte pairs (a plain copy: reference kernel for bandwidth
// and for calibrating hardware byte counters)
__global__ void k_copy2(size_t n2, const double2 *__restrict__ x, double2 *__restrict__ y) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n2) y[i] = x[i];
}

// out[e] = sum_p part[p*nel + e], p in order (one thread per element)
__global__ void k_sum_parts(int nparts, size_t nel, const double *__restrict__ part,
                            double *__restrict__ out) {
  const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= nel) return;
  double v = 0.0;
  for (int p = 0; p < nparts; p++) v += part[(size_t)p * nel + e];
  out[e] = v;
}

__global__ void __launch_bounds__(kThreads)
k_dot_part_f32(int n, const float *__restrict__ x, const float *__restrict__ y,
               double *__restrict__ part) {
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < (size_t)n;
       i += (size_t)gridDim.x * kThreads)
    acc += (double)x[i] * (double)y[i];
  block_sum_store(acc, part + blockIdx.x);
}

// v_{k+1} = t (v_{k+1} - cc v_k) [- v_{k-1}];  y += mu v_{k+1} (y may be
// null): the vector part of a Chebyshev step whose product was formed
// separately (the reference's AXPY/SCAL sequence, SRC/chebpol.c:550-555)
__global__ void k_cheb_update(int n, const double *__restrict__ vk,
                              const double *__restrict__ vkm1, double *__restrict__ vkp1,
                              double *__restrict__ y, double cc, double t, double mu) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  double w = vkp1[i] + (-cc) * vk[i];
  w = w * t;
  if (vkm1) w = w + (-1.0) * vkm1[i];
  vkp1[i] = w;
  if (y) y[i] = y[i] + mu * w;
}

__global__ void k_scale_f32(int n, float *__restrict__ x, double hs) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < (size_t)n) x[i] = (float)((double)x[i] * hs);
}

__global__ void k_scale(int n, double *__restrict__ x,
                        const double *__restrict__ inv_dev, double hs) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= (size_t)n) return;
  const double s = inv_dev ? (1.0 / *inv_dev) : hs;
  x[i] = x[i] * s;
}

__global__ void k_axpy(int n, double a, const double *__restrict__ x,
                       double *__restrict__ y) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < (size_t)n) y[i] = y[i] + a * x[i];
}

template <typename A, typename B>
__global__ void k_convert(size_t n, const A *__restrict__ x, B *__restrict__ y) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) y[i] = (B)x[i];
}

__global__ void k_zero(size_t n, double *__restrict__ x) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

template <typename T, typename TV>
__global__ void k_pack(int n, int B, int ncol, const TV *const *cols,
                       T *__restrict__ blk) {
  const size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const size_t row = g / B;
  const int c = (int)(g % B);
  if (row >= (size_t)n) return;
  blk[g] = c < ncol ? (T)cols[c][row] : (T)0;
}

template <typename T, typename TV>
__global__ void k_unpack(int n, int B, int ncol, const T *__restrict__ blk,
                         TV *const *cols) {
  const size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const size_t row = g / B;
  const int c = (int)(g % B);
  if (row >= (size_t)n || c >= ncol) return;
  cols[c][row] = (TV)blk[g];
}

// ---------------------------------------------------------------------------
// Tall-skinny Gram products on the FP64 matrix cores
// ---------------------------------------------------------------------------

typedef double v4d __attribute__((ext_vector_type(4)));

// part[p] (m1 x m2, column-major) = X(rows of part p, :)^T Y(rows of part p, :)
// for the row parts [p*chunk, min(n, (p+1)*chunk)).  One workgroup per
// (T x T tile of C, row part), T = 32 WB.  KR rows of the tile's T columns of
// X and of Y are staged in LDS per step (rows padded by one double, so that
// the column-wise stores and the row-wise MFMA reads fall in distinct
// banks); the next step's rows are loaded into registers while the matrix
// cores work on the current ones.  Each of the 4 waves accumulates a
// (T/2) x (T/2) quarter of the tile as WB x WB blocks of v_mfma_f64_16x16x4:
// lane l supplies A(l % 16, l / 16) = X(r + l/16, i + l%16) and
// B(l / 16, l % 16) = Y(r + l/16, j + l%16), and holds C(l/16 + 4q, l%16) of
// its block in accumulator q.  The rows of a part are summed in order, so the
// partial products are deterministic.
template <int WB, int KR>
__global__ void __launch_bounds__(kThreads)
k_gram_mfma(int n, int m1, int m2, const double *__restrict__ X, size_t ldx,
            const double *__restrict__ Y, size_t ldy, int chunk, int upper,
            double *__restrict__ part) {
  constexpr int T = 32 * WB;
  constexpr int PER = KR * T / kThreads;  // staged elements per thread and matrix
  __shared__ double Xs[KR][T + 1];
  __shared__ double Ys[KR][T + 1];
  int ti, tj;
  if (upper) {
    // tiles on and above the diagonal, column by column: x = tj(tj+1)/2 + ti
    const int x = (int)blockIdx.x;
    tj = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
    while ((tj + 1) * (tj + 2) / 2 <= x) tj++;
    while (tj * (tj + 1) / 2 > x) tj--;
    ti = x - tj * (tj + 1) / 2;
  } else {
    const int tiles1 = (m1 + T - 1) / T;
    ti = (int)(blockIdx.x % tiles1);
    tj = (int)(blockIdx.x / tiles1);
  }
  const int i0 = ti * T, j0 = tj * T;
  const int r0 = (int)blockIdx.y * chunk;
  const int r1 = min(n, r0 + chunk);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wi = (T / 2) * (wave & 1), wj = (T / 2) * (wave >> 1);
  v4d acc[WB][WB];
#pragma unroll
  for (int a = 0; a < WB; a++)
#pragma unroll
    for (int b = 0; b < WB; b++) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  // element e = t + q*kThreads of a step: row e % KR, column e / KR (KR
  // consecutive threads read KR consecutive rows of one column)
  double xr[PER], yr[PER];
  auto fetch = [&](int r) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = t + q * kThreads, rr = e % KR, cc = e / KR, row = r + rr;
      xr[q] = (row < r1 && i0 + cc < m1) ? X[(size_t)(i0 + cc) * ldx + row] : 0.0;
      yr[q] = (row < r1 && j0 + cc < m2) ? Y[(size_t)(j0 + cc) * ldy + row] : 0.0;
    }
  };
  if (r0 < r1) fetch(r0);
  for (int r = r0; r < r1; r += KR) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = t + q * kThreads;
      Xs[e % KR][e / KR] = xr[q];
      Ys[e % KR][e / KR] = yr[q];
    }
    __syncthreads();
    if (r + KR < r1) fetch(r + KR);  // in flight while the matrix cores run
#pragma unroll
    for (int kk = 0; kk < KR; kk += 4) {
      const int kr = kk + lane / 16, c16 = lane % 16;
      double av[WB], bv[WB];
#pragma unroll
      for (int a = 0; a < WB; a++) av[a] = Xs[kr][wi + 16 * a + c16];
#pragma unroll
      for (int b = 0; b < WB; b++) bv[b] = Ys[kr][wj + 16 * b + c16];
#pragma unroll
      for (int a = 0; a < WB; a++)
#pragma unroll
        for (int b = 0; b < WB; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  double *P = part + (size_t)blockIdx.y * m1 * m2;
#pragma unroll
  for (int a = 0; a < WB; a++)
#pragma unroll
    for (int b = 0; b < WB; b++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = i0 + wi + 16 * a + lane / 16 + 4 * q;
        const int col = j0 + wj + 16 * b + lane % 16;
        if (row < m1 && col < m2) P[(size_t)col * m1 + row] = acc[a][b][q];
      }
}

// ---------------------------------------------------------------------------
// Lanczos (batched over slices; blockIdx.y = slice)
// ---------------------------------------------------------------------------

template <typename TV>
struct LanArgs {
  int n;
  size_t ld;
  int chunk_shift;
  int maxchunks;
  TV **vtab;
  double *alpha, *beta, *nrm2a, *nrm2b;
  int *pass2, *active;
  double *h;
  int hcap;
  double *part;
};

template <typename TV>
LanArgs<TV> make_args(const LanBatch &L) {
  LanArgs<TV> a;
  a.n = L.n;
  a.ld = L.ld;
  a.chunk_shift = L.chunk_shift;
  a.maxchunks = L.maxchunks;
  a.vtab = reinterpret_cast<TV **>(L.vtab);
  a.alpha = L.alpha;
  a.beta = L.beta;
  a.nrm2a = L.nrm2a;
  a.nrm2b = L.nrm2b;
  a.pass2 = L.pass2;
  a.active = L.active;
  a.h = L.h;
  a.hcap = L.hcap;
  a.part = L.part;
  return a;
}

template <typename TV>
__device__ __forceinline__ TV *lan_col(const LanArgs<TV> &L, int s, int j) {
  const int mask = (1 << L.chunk_shift) - 1;
  return L.vtab[(size_t)s * L.maxchunks + (j >> L.chunk_shift)] +
         (size_t)(j & mask) * L.ld;
}

// number of row blocks used by the per-slice vector reductions
inline int lan_vec_blocks(int n) {
  int b = (n + kThreads * 4 - 1) / (kThreads * 4);
  return b < 1 ? 1 : (b > kDotBlocksMax ? kDotBlocksMax : b);
}

// znew -= beta zold (k > 0); partial <z, znew>
template <typename TV>
__global__ void __launch_bounds__(kThreads)
k_lan_alpha(LanArgs<TV> L, int k) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  if (!L.active[s]) return;
  TV *znew = lan_col(L, s, k + 1);
  const TV *z = lan_col(L, s, k);
  const TV *zold = k > 0 ? lan_col(L, s, k - 1) : nullptr;
  const double mb = -L.beta[s];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < (size_t)L.n;
       i += (size_t)gridDim.x * kThreads) {
    double w = znew[i];
    if (zold) {
      w = w + mb * (double)zold[i];
      znew[i] = (TV)w;
    }
    acc += (double)z[i] * w;
  }
  block_sum_store(acc, L.part + (size_t)s * gridDim.x + blockIdx.x);
}

// znew -= alpha z; partial |znew|^2
template <typename TV>
__global__ void __launch_bounds__(kThreads)
k_lan_sub_alpha(LanArgs<TV> L, int k) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  if (!L.active[s]) return;
  TV *znew = lan_col(L, s, k + 1);
  const TV *z = lan_col(L, s, k);
  const double ma = -L.alpha[s];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < (size_t)L.n;
       i += (size_t)gridDim.x * kThreads) {
    const TV w = (TV)((double)znew[i] + ma * (double)z[i]);
    znew[i] = w;
    acc += (double)w * (double)w;
  }
  block_sum_store(acc, L.part + (size_t)s * gridDim.x + blockIdx.x);
}

// per-slice sums of the vector partials into one of the scalar arrays
// which: 0 alpha, 1 nrm2a, 2 nrm2b (pass 0, sets pass2), 3 nrm2b (pass 1)
template <typename TV>
__global__ void k_lan_reduce(LanArgs<TV> L, int nblk, int which) {
  const int s = blockIdx.x;
  if (!L.active[s]) return;
  if (which == 3 && !L.pass2[s]) return;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64)
    v += L.part[(size_t)s * nblk + b];
  v = wave_sum(v);
  if (threadIdx.x != 0) return;
  if (which == 0) {
    L.alpha[s] = v;
  } else if (which == 1) {
    L.nrm2a[s] = v;
  } else {
    L.nrm2b[s] = v;
    if (which == 2) {
      const double eta = 1.0 / sqrt(2.0);
      L.pass2[s] = (sqrt(v) > eta * sqrt(L.nrm2a[s])) ? 0 : 1;
    }
  }
}

// GEMV-T partials: part[(s*hcap + j)*nrb + rb] = sum over row block rb of
// V_s(:,j) * znew_s, for the kTcols columns of this block.  Each lane keeps
// one accumulator per column and streams 16-byte row groups (2 FP64 or 4
// FP32 rows, one independent load per column per iteration); the cross-lane
// reduction runs once per column at the end.  Rows in [n, ld) are zero in
// every basis column, so whole row groups up to ld are read without masks.
constexpr int kTcols = 32;               // columns per block
template <typename TV>
constexpr int gt_rows() {  // rows per block: 8192 for both storage types
  return kThreads * Row16<TV>::R * Row16<TV>::kTiters;
}

template <bool NT, typename TV>
__global__ void __launch_bounds__(kThreads)
k_lan_gemvt(LanArgs<TV> L, int k, int pass, int nrb) {
  using RV = Row16<TV>;
  using vec = typename RV::vec;
  constexpr int R = RV::R;
  const int s = blockIdx.z;
  if (!L.active[s]) return;
  if (pass && !L.pass2[s]) return;
  const int ncols = k + 1;
  const int j0 = blockIdx.y * kTcols;
  if (j0 >= ncols) return;
  const int jn = min(kTcols, ncols - j0);
  const int rb = blockIdx.x;
  const vec *z = reinterpret_cast<const vec *>(lan_col(L, s, k + 1));
  // the kTcols columns of a block lie in one basis chunk (chunks hold a
  // power of two >= kTcols columns), kTcols-aligned: column c is at
  // base + c*ld
  const vec *base = reinterpret_cast<const vec *>(lan_col(L, s, j0));
  const size_t cstride = L.ld / R;
  double acc[kTcols];
#pragma unroll
  for (int c = 0; c < kTcols; c++) acc[c] = 0.0;
  const size_t ngrp = L.ld / R;
  const size_t p0 = (size_t)rb * (gt_rows<TV>() / R) + threadIdx.x;
  // full column groups load without guards; the last group of a basis
  // guards its columns
  auto sweep = [&](auto full) {
    for (int i = 0; i < RV::kTiters; i++) {
      const size_t p = p0 + (size_t)i * kThreads;
      if (p >= ngrp) break;
      const vec zv = z[p];
      double zz[R];
#pragma unroll
      for (int r = 0; r < R; r++) zz[r] = RV::get(zv, r);
      vec v[kTcols];
#pragma unroll
      for (int c = 0; c < kTcols; c++)
        v[c] = (decltype(full)::value || c < jn) ? RV::template ld<NT>(base + c * cstride + p)
                                                 : RV::zero();
#pragma unroll
      for (int c = 0; c < kTcols; c++) {
        double t = RV::get(v[c], 0) * zz[0];
#pragma unroll
        for (int r = 1; r < R; r++) t += RV::get(v[c], r) * zz[r];
        acc[c] += t;
      }
    }
  };
  if (jn == kTcols)
    sweep(std::true_type{});
  else
    sweep(std::false_type{});
  __shared__ double red[kThreads / 64][kTcols];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kTcols; c++) {
    const double t = wave_sum(acc[c]);
    if (lane == 0) red[wave][c] = t;
  }
  __syncthreads();
  if (threadIdx.x < jn) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) t += red[w][threadIdx.x];
    L.part[((size_t)s * L.hcap + j0 + threadIdx.x) * nrb + rb] = t;
  }
}

// h_s[j] = sum_rb part[(s*hcap + j)*nrb + rb]
template <typename TV>
__global__ void __launch_bounds__(kThreads)
k_lan_reduce_h(LanArgs<TV> L, int k, int pass, int nrb) {
  const int s = blockIdx.y;
  if (!L.active[s]) return;
  if (pass && !L.pass2[s]) return;
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j > k) return;
  const double *p = L.part + ((size_t)s * L.hcap + j) * nrb;
  double v = 0.0;
  for (int b = 0; b < nrb; b++) v += p[b];
  L.h[(size_t)s * L.hcap + j] = v;
}

// znew -= V(:,0:k) h, column by column in the order of DGEMV 'N'
// (y_i += (-h_j) V_ij for j = 0..k), with the block partial of |znew|^2.
// Each thread owns kGnPairs 16-byte row groups and walks the columns
// kGnUnroll at a time so that 16 loads are in flight per thread.  The
// running sums are FP64; an FP32 basis is rounded once, when znew is stored.
template <typename TV>
constexpr int gn_rows() {  // rows per block: 1024 for both storage types
  return kThreads * Row16<TV>::R * Row16<TV>::kGnPairs;
}

template <bool NT, typename TV>
__global__ void __launch_bounds__(kThreads)
k_lan_gemvn(LanArgs<TV> L, int k, int pass, int nblk) {
#pragma clang fp contract(off)
  using RV = Row16<TV>;
  using vec = typename RV::vec;
  constexpr int R = RV::R;
  constexpr int P = RV::kGnPairs;
  constexpr int U = RV::kGnUnroll;
  const int s = blockIdx.y;
  if (!L.active[s]) return;
  if (pass && !L.pass2[s]) return;
  vec *z = reinterpret_cast<vec *>(lan_col(L, s, k + 1));
  const double *h = L.h + (size_t)s * L.hcap;
  const size_t ngrp = L.ld / R;
  size_t pr[P];
  double acc[P][R];
  bool ok[P];
#pragma unroll
  for (int i = 0; i < P; i++) {
    pr[i] = (size_t)blockIdx.x * (gn_rows<TV>() / R) + (size_t)i * kThreads + threadIdx.x;
    ok[i] = pr[i] < ngrp;
    const vec zv = ok[i] ? z[pr[i]] : RV::zero();
#pragma unroll
    for (int r = 0; r < R; r++) acc[i][r] = RV::get(zv, r);
  }
  int j = 0;
  const size_t cstride = L.ld / R;
  for (; j + U <= k + 1; j += U) {
    vec v[U][P];
    double mh[U];
    // j is a multiple of U (<= 32): the group lies in one basis chunk
    const vec *base = reinterpret_cast<const vec *>(lan_col(L, s, j));
#pragma unroll
    for (int u = 0; u < U; u++) {
      const vec *col = base + u * cstride;
      mh[u] = -h[j + u];
#pragma unroll
      for (int i = 0; i < P; i++) v[u][i] = ok[i] ? RV::template ld<NT>(col + pr[i]) : RV::zero();
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int i = 0; i < P; i++)
#pragma unroll
        for (int r = 0; r < R; r++) acc[i][r] = acc[i][r] + mh[u] * RV::get(v[u][i], r);
  }
  for (; j <= k; j++) {
    const vec *col = reinterpret_cast<const vec *>(lan_col(L, s, j));
    const double m = -h[j];
#pragma unroll
    for (int i = 0; i < P; i++) {
      if (ok[i]) {
        const vec v = RV::template ld<NT>(col + pr[i]);
#pragma unroll
        for (int r = 0; r < R; r++) acc[i][r] = acc[i][r] + m * RV::get(v, r);
      }
    }
  }
  double sq = 0.0;
#pragma unroll
  for (int i = 0; i < P; i++) {
    if (ok[i]) {
      RV::store(z + pr[i], acc[i]);
      // the norm of the stored vector (FP32: after rounding)
      double st[R];
#pragma unroll
      for (int r = 0; r < R; r++)
        st[r] = std::is_same<TV, float>::value ? (double)(float)acc[i][r] : acc[i][r];
      double t = st[0] * st[0];
#pragma unroll
      for (int r = 1; r < R; r++) t = t + st[r] * st[r];
      sq += t;
    }
  }
  block_sum_store(sq, L.part + (size_t)s * nblk + blockIdx.x);
}

template <typename TV>
__global__ void k_lan_scale(LanArgs<TV> L, int k) {
  const int s = blockIdx.y;
  if (!L.active[s]) return;
  TV *z = lan_col(L, s, k + 1);
  const double t = 1.0 / L.beta[s];
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < (size_t)L.n;
       i += (size_t)gridDim.x * kThreads)
    z[i] = (TV)((double)z[i] * t);
}

// ---------------------------------------------------------------------------
// Symmetric tridiagonal: Sturm counts and bisection (LAPACK DLAEBZ-style
// pivot guard, W. Kahan's method)
// ---------------------------------------------------------------------------

__device__ int sturm_less(int m, const double *__restrict__ d,
                          const double *__restrict__ e, double x,
                          double pivmin) {
  // q_i = (d_i - x) - e_{i-1}^2 / q_{i-1}; counts eigenvalues <= x
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q <= 0.0) cnt++;
  for (int i = 1; i < m; i++) {
    const double ei = e[i - 1];
    q = (d[i] - x) - (ei * ei) / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q <= 0.0) cnt++;
  }
  return cnt;
}

// ws: [0] lower bound, [1] upper bound, [2] pivmin, [3] first index >= thr
// Gershgorin bounds by a block reduction, then one Sturm count at thr.
__global__ void __launch_bounds__(kThreads)
k_tri_prep(int m, size_t ld, const double *__restrict__ d, const double *__restrict__ e,
           const double *__restrict__ thr_arr, double thr, double *__restrict__ ws) {
  // problem blockIdx.x of a batch: T_p in d/e + p*ld, 4 workspace values each
  const int pb = blockIdx.x;
  d += (size_t)pb * ld;
  e += (size_t)pb * ld;
  ws += 4 * (size_t)pb;
  if (thr_arr) thr = thr_arr[pb];
  double lo = HUGE_VAL, hi = -HUGE_VAL, emax2 = 0.0;
  for (int i = threadIdx.x; i < m; i += kThreads) {
    double r = 0.0;
    if (i > 0) r += fabs(e[i - 1]);
    if (i < m - 1) {
      r += fabs(e[i]);
      emax2 = fmax(emax2, e[i] * e[i]);
    }
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
  }
  __shared__ double red[3][kThreads];
  red[0][threadIdx.x] = lo;
  red[1][threadIdx.x] = hi;
  red[2][threadIdx.x] = emax2;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + w]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  lo = red[0][0];
  hi = red[1][0];
  const double pivmin = DBL_MIN * fmax(1.0, red[2][0]);
  const double span = fmax(fabs(lo), fabs(hi));
  lo -= 2.0 * DBL_EPSILON * span + pivmin;
  hi += 2.0 * DBL_EPSILON * span + pivmin;
  ws[0] = lo;
  ws[1] = hi;
  ws[2] = pivmin;
  ws[3] = (thr <= lo) ? 0.0 : (double)sturm_less(m, d, e, thr, pivmin);
}

// Multisection: one wave per eigenvalue index idx >= first (first = ws[3],
// or 0 when all == 1).  Each round the 64 lanes evaluate Sturm counts at 64
// interior points of [lo, hi]; the new interval is the gap between the last
// point with count <= idx and the first with count > idx, so each round
// shrinks the interval 65-fold (about 9 rounds to double precision).
__global__ void __launch_bounds__(64)
k_tri_msect(int m, size_t ld, const double *__restrict__ d, const double *__restrict__ e,
            const double *__restrict__ thr_arr, double thr, const double *__restrict__ ws,
            double *__restrict__ lam, int all) {
  const int idx = blockIdx.x;
  // problem blockIdx.y of a batch
  const int pb = blockIdx.y;
  d += (size_t)pb * ld;
  e += (size_t)pb * ld;
  lam += (size_t)pb * ld;
  ws += 4 * (size_t)pb;
  if (thr_arr) thr = thr_arr[pb];
  const int first = all ? 0 : (int)ws[3];
  if (idx < first || idx >= m) return;
  const int l = threadIdx.x;
  double lo = (all || thr <= ws[0]) ? ws[0] : thr, hi = ws[1];
  const double pivmin = ws[2];
  for (int round = 0; round < 40; round++) {
    if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double h = (hi - lo) / 65.0;
    const double x = lo + h * (l + 1);
    const bool above = sturm_less(m, d, e, x, pivmin) > idx;
    const unsigned long long mask = __ballot(above);
    double nlo, nhi;
    if (mask == 0ull) {
      nlo = lo + h * 64;
      nhi = hi;
    } else {
      const int f = __ffsll((long long)mask) - 1;
      nhi = lo + h * (f + 1);
      nlo = f == 0 ? lo : lo + h * f;
    }
    if (!(nlo > lo || nhi < hi)) break;  // no progress at this precision
    lo = nlo;
    hi = nhi;
  }
  if (l == 0) lam[idx] = 0.5 * (lo + hi);
}

__global__ void k_tri_sum(int m, size_t ld, const double *__restrict__ ws,
                          const double *__restrict__ lam,
                          double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  const int pb = blockIdx.x;
  ws += 4 * (size_t)pb;
  lam += (size_t)pb * ld;
  out += 2 * (size_t)pb;
  const int first = (int)ws[3];
  double s = 0.0;
  for (int i = first; i < m; i++) s += lam[i];
  out[0] = s;
  out[1] = (double)(m - first);
}

// ---------------------------------------------------------------------------
// Misc
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kThreads)
k_intchx(int Mdeg, const double *__restrict__ mu, int npts,
         const double *__restrict__ th, double th0, double *__restrict__ yi) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= npts) return;
  const double t = th[j];
  double y = mu[0] * (th0 - t);
  for (int k = 1; k <= Mdeg; k++) {
    const double val0 = sin(k * th0) / k;
    y += mu[k] * (val0 - sin(k * t) / k);
  }
  yi[j] = y;
}

__global__ void __launch_bounds__(kThreads)
k_col_dots(int n, const double *__restrict__ X, size_t ldx,
           const double *__restrict__ Y, size_t ldy, double *__restrict__ part) {
  const int c = blockIdx.y;
  const double *x = X + (size_t)c * ldx;
  const double *y = Y + (size_t)c * ldy;
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < (size_t)n;
       i += (size_t)gridDim.x * kThreads)
    acc += x[i] * y[i];
  block_sum_store(acc, part + (size_t)c * gridDim.x + blockIdx.x);
}

__global__ void k_col_reduce(int ncol, int nblk, const double *__restrict__ part,
                             double *__restrict__ out) {
  const int c = blockIdx.x;
  if (c >= ncol) return;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) v += part[(size_t)c * nblk + b];
  v = wave_sum(v);
  if (threadIdx.x == 0) out[c] = v;
}

__global__ void k_col_scale(int n, double *__restrict__ X, size_t ld,
                            const double *__restrict__ f) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= (size_t)n) return;
  X[(size_t)blockIdx.y * ld + i] *= f[blockIdx.y];
}

__global__ void k_col_axpy(int n, const double *__restrict__ a,
                           const double *__restrict__ X, size_t ldx,
                           double *__restrict__ Y, size_t ldy) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= (size_t)n) return;
  const int c = blockIdx.y;
  Y[(size_t)c * ldy + i] = Y[(size_t)c * ldy + i] + a[c] * X[(size_t)c * ldx + i];
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_rand(size_t n, double *__restrict__ x, uint64_t seed,
                       uint64_t offset, int rademacher) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = splitmix64(seed ^ splitmix64(offset + i));
  if (rademacher)
    x[i] = (r >> 63) ? 1.0 : -1.0;
  else
    x[i] = (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

}  // namespace

// ===========================================================================
// Launchers
// ===========================================================================

void spmv_cols_f64(const DevCsr &A, const double *X, double *Y, int ncol,
                   size_t ld, int lanes, hipStream_t s) {
  if (A.n == 0 || ncol == 0) return;
  dim3 grid(grid_for((size_t)A.n * lanes), ncol);
  switch (lanes) {
#define CASE(L)                                                              \
  case L:                                                                    \
    hipLaunchKernelGGL(k_spmv_f64<L>, grid, dim3(kThreads), 0, s, A.n, A.ia, \
                       A.ja, A.a64, X, Y, ld, ncol);                         \
    break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default:
      hipLaunchKernelGGL(k_spmv_f64<1>, grid, dim3(kThreads), 0, s, A.n, A.ia,
                         A.ja, A.a64, X, Y, ld, ncol);
  }
}

void spmv_f64(const DevCsr &A, const double *x, double *y, int lanes,
              hipStream_t s) {
  spmv_cols_f64(A, x, y, 1, (size_t)A.n, lanes, s);
}

template <typename T>
void cheb_step_vec(const DevCsr &A, const T *vk, const T *vkm1, T *vkp1, T *y,
                   double cc, double t, double mu0, double muk, bool first,
                   int lanes, hipStream_t s) {
  if (A.n == 0) return;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64
                                              : (const T *)A.a32;
  dim3 grid(grid_for((size_t)A.n * lanes));
  switch (lanes) {
#define CASE(L)                                                                \
  case L:                                                                      \
    hipLaunchKernelGGL((k_cheb_vec<T, L>), grid, dim3(kThreads), 0, s, A.n,    \
                       A.ia, A.ja, a, vk, vkm1, vkp1, y, (T)cc, (T)t, (T)mu0,  \
                       (T)muk, first ? 1 : 0);                                 \
    break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default:
      hipLaunchKernelGGL((k_cheb_vec<T, 1>), grid, dim3(kThreads), 0, s, A.n,
                         A.ia, A.ja, a, vk, vkm1, vkp1, y, (T)cc, (T)t, (T)mu0,
                         (T)muk, first ? 1 : 0);
  }
}
void build_row_blocks(int n, const int *ia, std::vector<int> &blk, int max_rows, int max_nnz) {
  blk.clear();
  blk.push_back(0);
  int r = 0;
  while (r < n) {
    const int r0 = r;
    if (ia[r + 1] - ia[r] > max_nnz) {
      r++;
    } else {
      while (r < n && r - r0 < max_rows && ia[r + 1] - ia[r0] <= max_nnz) r++;
    }
    blk.push_back(r);
  }
}

int bal_level(int elem, int B) {
  int q = B * elem / 16, l = 0;
  while ((1 << l) < q) l++;
  return l;
}

template <typename T>
void cheb_step_bal(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                   T *y, double cc, double t, const double *mus, int k,
                   int nact, hipStream_t s, int ymode) {
  if (A.n == 0 || nact <= 0) return;
  const int lv = bal_level((int)sizeof(T), B);
  if (lv > 4 || !A.rowblk_q[lv] || A.nrowblk_q[lv] == 0) return;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64 : (const T *)A.a32;
  switch (B) {
#define CASE(BB)                                                                         \
  case BB: {                                                                             \
    const EpiCheb<T, BB> epi{vk, vkm1, vkp1, y, (T)cc, (T)t, mus, k, nact, ymode};       \
    hipLaunchKernelGGL((k_blk_bal<T, BB, EpiCheb<T, BB>>), dim3(A.nrowblk_q[lv]),        \
                       dim3(kThreads), 0, s, A.rowblk_q[lv], A.ia, A.ja, a, vk, epi);    \
    break;                                                                               \
  }
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32)
#undef CASE
    default:
      break;
  }
}
template <typename T>
void clen_step_bal(const DevCsr &A, int B, const T *b1, const T *b2, const T *v, T *out,
                   double cc, double t, const double *mus, int k, hipStream_t s) {
  if (A.n == 0) return;
  if (!b1) {  // top term: no product
    hipLaunchKernelGGL((k_blk_colscale<T>), dim3(grid_for((size_t)A.n * B)), dim3(kThreads), 0,
                       s, (size_t)A.n, B, v, out, mus, k);
    return;
  }
  const int lv = bal_level((int)sizeof(T), B);
  if (lv > 4 || !A.rowblk_q[lv] || A.nrowblk_q[lv] == 0) return;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64 : (const T *)A.a32;
  switch (B) {
#define CASE(BB)                                                                         \
  case BB: {                                                                             \
    const EpiClen<T, BB> epi{b1, b2, v, out, (T)cc, (T)t, mus, k};                       \
    hipLaunchKernelGGL((k_blk_bal<T, BB, EpiClen<T, BB>>), dim3(A.nrowblk_q[lv]),        \
                       dim3(kThreads), 0, s, A.rowblk_q[lv], A.ia, A.ja, a, b1, epi);    \
    break;                                                                               \
  }
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32)
#undef CASE
    default:
      break;
  }
}
template void cheb_step_bal<double>(const DevCsr &, int, const double *, const double *,
                                    double *, double *, double, double, const double *, int,
                                    int, hipStream_t, int);
template void cheb_step_bal<float>(const DevCsr &, int, const float *, const float *, float *,
                                   float *, double, double, const double *, int, int,
                                   hipStream_t, int);
template void clen_step_bal<double>(const DevCsr &, int, const double *, const double *,
                                    const double *, double *, double, double, const double *,
                                    int, hipStream_t);
template void clen_step_bal<float>(const DevCsr &, int, const float *, const float *,
                                   const float *, float *, double, double, const double *, int,
                                   hipStream_t);

template <typename T>
void cheb_step_stream(const DevCsr &A, const T *vk, const T *vkm1, T *vkp1,
                      T *y, double cc, double t, double mu0, double muk,
                      bool first, hipStream_t s) {
  if (A.n == 0 || A.nrowblk == 0) return;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64
                                              : (const T *)A.a32;
  if (A.jo16)
    hipLaunchKernelGGL((k_cheb_stream<T, true>), dim3(A.nrowblk), dim3(kThreads), 0, s,
                       A.rowblk, A.ia, A.ja, A.jo16, a, vk, vkm1, vkp1, y, (T)cc, (T)t,
                       (T)mu0, (T)muk, first ? 1 : 0);
  else
    hipLaunchKernelGGL((k_cheb_stream<T, false>), dim3(A.nrowblk), dim3(kThreads), 0, s,
                       A.rowblk, A.ia, A.ja, A.jo16, a, vk, vkm1, vkp1, y, (T)cc, (T)t,
                       (T)mu0, (T)muk, first ? 1 : 0);
}
template void cheb_step_stream<double>(const DevCsr &, const double *,
                                       const double *, double *, double *,
                                       double, double, double, double, bool,
                                       hipStream_t);
template void cheb_step_stream<float>(const DevCsr &, const float *,
                                      const float *, float *, float *, double,
                                      double, double, double, bool,
                                      hipStream_t);

template void cheb_step_vec<double>(const DevCsr &, const double *,
                                    const double *, double *, double *, double,
                                    double, double, double, bool, int,
                                    hipStream_t);
template void cheb_step_vec<float>(const DevCsr &, const float *, const float *,
                                   float *, float *, double, double, double,
                                   double, bool, int, hipStream_t);

template <typename T>
void cheb_step_blk(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                   T *y, double cc, double t, const double *mus, int k,
                   int nact, hipStream_t s) {
  if (A.n == 0 || nact <= 0) return;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64
                                              : (const T *)A.a32;
  dim3 grid(grid_for((size_t)A.n * B));
  switch (B) {
#define CASE(BB)                                                              \
  case BB:                                                                    \
    hipLaunchKernelGGL((k_cheb_blk<T, BB>), grid, dim3(kThreads), 0, s, A.n,  \
                       A.ia, A.ja, a, vk, vkm1, vkp1, y, (T)cc, (T)t, mus, k, \
                       nact);                                                 \
    break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default:
      break;
  }
}
template <typename T>
void cheb_step_blkrow(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                      T *y, double cc, double t, const double *mus, int k,
                      int nact, hipStream_t s, int ymode) {
  if (A.n == 0 || nact <= 0 || A.nrowblk == 0) return;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64
                                              : (const T *)A.a32;
  switch (B) {
#define CASE(BB)                                                                        \
  case BB:                                                                              \
    if (A.jo16)                                                                         \
      hipLaunchKernelGGL((k_cheb_blkrow<T, BB, true>), dim3(A.nrowblk), dim3(kThreads), 0, \
                         s, A.rowblk, A.ia, A.ja, A.jo16, a, vk, vkm1, vkp1, y, (T)cc,   \
                         (T)t, mus, k, nact, ymode);                                    \
    else                                                                                \
      hipLaunchKernelGGL((k_cheb_blkrow<T, BB, false>), dim3(A.nrowblk), dim3(kThreads), 0, \
                         s, A.rowblk, A.ia, A.ja, A.jo16, a, vk, vkm1, vkp1, y, (T)cc,   \
                         (T)t, mus, k, nact, ymode);                                    \
    break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32)
#undef CASE
    default:
      break;
  }
}
template <typename T>
void clen_step_blkrow(const DevCsr &A, int B, const T *b1, const T *b2, const T *v, T *out,
                      double cc, double t, const double *mus, int k, hipStream_t s) {
  if (A.n == 0 || A.nrowblk == 0) return;
  if (!b1) {  // top term: no product
    hipLaunchKernelGGL((k_blk_colscale<T>), dim3(grid_for((size_t)A.n * B)), dim3(kThreads), 0,
                       s, (size_t)A.n, B, v, out, mus, k);
    return;
  }
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64 : (const T *)A.a32;
  switch (B) {
#define CASE(BB)                                                                        \
  case BB:                                                                              \
    if (A.jo16)                                                                         \
      hipLaunchKernelGGL((k_clen_blkrow<T, BB, true>), dim3(A.nrowblk), dim3(kThreads), 0, \
                         s, A.rowblk, A.ia, A.ja, A.jo16, a, b1, b2, v, out, (T)cc, (T)t, \
                         mus, k);                                                       \
    else                                                                                \
      hipLaunchKernelGGL((k_clen_blkrow<T, BB, false>), dim3(A.nrowblk), dim3(kThreads), 0, \
                         s, A.rowblk, A.ia, A.ja, A.jo16, a, b1, b2, v, out, (T)cc, (T)t, \
                         mus, k);                                                       \
    break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32)
#undef CASE
    default:
      break;
  }
}
template void clen_step_blkrow<float>(const DevCsr &, int, const float *, const float *,
                                      const float *, float *, double, double, const double *,
                                      int, hipStream_t);
template void clen_step_blkrow<double>(const DevCsr &, int, const double *, const double *,
                                       const double *, double *, double, double,
                                       const double *, int, hipStream_t);

template <typename T>
void clen_step_seg(const DevCsr &A, int Bp, const T *b1, const T *b2, const T *v, T *out,
                   double cc, double t, const double *mus, int d, int nseg, const SegTops &tops,
                   hipStream_t s) {
  if (A.n == 0 || nseg <= 0) return;
  int l = 0;
  while ((1 << l) < nseg) l++;
  const int *rb = l == 0 ? A.rowblk : A.rowblk_q[l];
  const int nb = l == 0 ? A.nrowblk : A.nrowblk_q[l];
  if (!rb || nb == 0) return;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64 : (const T *)A.a32;
  // data of one full-width launch: the matrix and four Bp-wide vector passes
  const double bytes = (double)A.nnz * (sizeof(T) + 4) + 4.0 * (double)Bp * A.n * sizeof(T);
  if (bytes > kInfinityCacheBytes)
    hipLaunchKernelGGL((k_clen_seg<T, true>), dim3(nb), dim3(kThreads), 0, s, rb, A.ia, A.ja,
                       a, b1, b2, v, out, (T)cc, (T)t, mus, d, Bp, nseg, kStreamNnz >> l, tops);
  else
    hipLaunchKernelGGL((k_clen_seg<T, false>), dim3(nb), dim3(kThreads), 0, s, rb, A.ia, A.ja,
                       a, b1, b2, v, out, (T)cc, (T)t, mus, d, Bp, nseg, kStreamNnz >> l, tops);
}
template void clen_step_seg<float>(const DevCsr &, int, const float *, const float *,
                                   const float *, float *, double, double, const double *, int,
                                   int, const SegTops &, hipStream_t);
template void clen_step_seg<double>(const DevCsr &, int, const double *, const double *,
                                    const double *, double *, double, double, const double *,
                                    int, int, const SegTops &, hipStream_t);

template void cheb_step_blkrow<double>(const DevCsr &, int, const double *,
                                       const double *, double *, double *, double,
                                       double, const double *, int, int, hipStream_t, int);
template void cheb_step_blkrow<float>(const DevCsr &, int, const float *, const float *,
                                      float *, float *, double, double, const double *,
                                      int, int, hipStream_t, int);

template void cheb_step_blk<double>(const DevCsr &, int, const double *,
                                    const double *, double *, double *, double,
                                    double, const double *, int, int,
                                    hipStream_t);
template void cheb_step_blk<float>(const DevCsr &, int, const float *,
                                   const float *, float *, float *, double,
                                   double, const double *, int, int,
                                   hipStream_t);

// rows per block of the KPM kernel: enough rows that every block does a few
// sweeps, and at most ~4096 blocks.
static int kpm_rows_per_block(int n, int B) {
  const int rpw = 64 / B;
  const int sweep = (kThreads / 64) * rpw;
  int rpb = sweep * 4;
  while ((n + rpb - 1) / rpb > 4096) rpb *= 2;
  return rpb;
}

int kpm_part_blocks(int n, int B) {
  const int rpb = kpm_rows_per_block(n, B);
  return (n + rpb - 1) / rpb;
}

template <typename T>
int kpm_step_blk(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1,
                 const T *v0, double ctr, double scal, double *part,
                 hipStream_t s) {
  const int rpb = kpm_rows_per_block(A.n, B);
  const int nb = (A.n + rpb - 1) / rpb;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64
                                              : (const T *)A.a32;
  switch (B) {
#define CASE(BB)                                                             \
  case BB:                                                                   \
    hipLaunchKernelGGL((k_kpm_blk<T, BB>), dim3(nb), dim3(kThreads), 0, s,   \
                       A.n, A.ia, A.ja, a, vk, vkm1, vkp1, v0, (T)ctr,       \
                       (T)scal, part, rpb);                                  \
    break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default:
      return 0;
  }
  return nb;
}
template <typename T>
int kpm_step_bal(const DevCsr &A, int B, const T *vk, const T *vkm1, T *vkp1, const T *v0,
                 double ctr, double scal, double *part, hipStream_t s) {
  const int lv = bal_level((int)sizeof(T), B);
  if (lv > 4 || !A.rowblk_q[lv] || A.nrowblk_q[lv] == 0) return 0;
  const T *a = std::is_same<T, double>::value ? (const T *)A.a64 : (const T *)A.a32;
  switch (B) {
#define CASE(BB)                                                                          \
  case BB:                                                                                \
    hipLaunchKernelGGL((k_kpm_bal<T, BB>), dim3(A.nrowblk_q[lv]), dim3(kThreads), 0, s,   \
                       A.rowblk_q[lv], A.ia, A.ja, a, vk, vkm1, vkp1, v0, (T)ctr, (T)scal, \
                       part);                                                             \
    break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32)
#undef CASE
    default:
      return 0;
  }
  return A.nrowblk_q[lv];
}
template int kpm_step_bal<double>(const DevCsr &, int, const double *, const double *, double *,
                                  const double *, double, double, double *, hipStream_t);
template int kpm_step_bal<float>(const DevCsr &, int, const float *, const float *, float *,
                                 const float *, double, double, double *, hipStream_t);

template int kpm_step_blk<double>(const DevCsr &, int, const double *,
                                  const double *, double *, const double *,
                                  double, double, double *, hipStream_t);
template int kpm_step_blk<float>(const DevCsr &, int, const float *,
                                 const float *, float *, const float *, double,
                                 double, double *, hipStream_t);

void reduce_partials(const double *part, int nblk, int ncol, double *out,
                     hipStream_t s) {
  if (ncol <= 0) return;
  hipLaunchKernelGGL(k_reduce_partials, dim3(ncol), dim3(kThreads), 0, s, part,
                     nblk, ncol, out);
}

void copy_f64(size_t n, const double *x, double *y, hipStream_t s) {
  const size_t n2 = n / 2;
  if (n2)
    hipLaunchKernelGGL(k_copy2, dim3(grid_for(n2)), dim3(kThreads), 0, s, n2,
                       reinterpret_cast<const double2 *>(x), reinterpret_cast<double2 *>(y));
}

void sum_parts(const double *part, int nparts, size_t nel, double *out, hipStream_t s) {
  if (nel == 0) return;
  hipLaunchKernelGGL(k_sum_parts, dim3(grid_for(nel)), dim3(kThreads), 0, s, nparts, nel, part,
                     out);
}

int dot_part_blocks(int n) { return lan_vec_blocks(n); }

int dot_partials(int n, const double *x, const double *y, double *part,
                 hipStream_t s) {
  const int nb = lan_vec_blocks(n);
  hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(kThreads), 0, s, n, x, y,
                     part);
  return nb;
}

void gram_tn_mfma(int n, int m1, const double *X, size_t ldx, int m2, const double *Y,
                  size_t ldy, int chunk, int nparts, bool upper, double *part, hipStream_t s) {
  if (n <= 0 || m1 <= 0 || m2 <= 0 || nparts <= 0) return;
  if (upper && m1 != m2) return;
  // 64 x 64 workgroup tiles, 64-row steps: measured on n = 1e6 rows, m = 800
  // (FP64), 29.0 ms against 30.0 / 31.0 ms for 32- / 16-row steps and
  // 33.3 ms for 128 x 128 tiles
  constexpr int T = 64;
  const int t1 = (m1 + T - 1) / T, t2 = (m2 + T - 1) / T;
  const int tiles = upper ? t1 * (t1 + 1) / 2 : t1 * t2;
  hipLaunchKernelGGL((k_gram_mfma<2, 64>), dim3(tiles, nparts), dim3(kThreads), 0, s, n, m1, m2,
                     X, ldx, Y, ldy, chunk, upper ? 1 : 0, part);
}

// C = sum of the upper triangles of the m x m parts, in part order (as
// k_sum_parts sums each element), written to both triangles of C
__global__ void k_sum_parts_upper(int nparts, int m, const double *__restrict__ part,
                                  double *__restrict__ C) {
  const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const size_t nel = (size_t)m * m;
  if (e >= nel) return;
  const int r = (int)(e % m), c = (int)(e / m);
  if (r > c) return;
  double v = 0.0;
  for (int p = 0; p < nparts; p++) v += part[(size_t)p * nel + e];
  C[e] = v;
  C[(size_t)r * m + c] = v;
}

void sum_parts_upper(const double *part, int nparts, int m, double *C, hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_sum_parts_upper, dim3(grid_for((size_t)m * m)), dim3(kThreads), 0, s,
                     nparts, m, part, C);
}

void scale_vec(int n, double *x, const double *inv_scale_dev, double hostscale,
               hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x,
                     inv_scale_dev, hostscale);
}

void cheb_update(int n, const double *vk, const double *vkm1, double *vkp1, double *y,
                 double cc, double t, double mu, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cheb_update, dim3(grid_for((size_t)n)), dim3(kThreads), 0, s, n, vk,
                     vkm1, vkp1, y, cc, t, mu);
}

void axpy_vec(int n, double a, const double *x, double *y, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(kThreads), 0, s, n, a, x,
                     y);
}

void copy_f64_to_f32(size_t n, const double *x, float *y, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)),
                     dim3(kThreads), 0, s, n, x, y);
}
void copy_f32_to_f64(size_t n, const float *x, double *y, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)),
                     dim3(kThreads), 0, s, n, x, y);
}
void fill_zero_f64(size_t n, double *x, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x);
}

template <typename T, typename TV>
void pack_cols(int n, int B, int ncol, const TV *const *cols, T *blk, hipStream_t s) {
  hipLaunchKernelGGL((k_pack<T, TV>), dim3(grid_for((size_t)n * B)), dim3(kThreads), 0, s,
                     n, B, ncol, cols, blk);
}
template <typename T, typename TV>
void unpack_cols(int n, int B, int ncol, const T *blk, TV *const *cols, hipStream_t s) {
  hipLaunchKernelGGL((k_unpack<T, TV>), dim3(grid_for((size_t)n * B)), dim3(kThreads), 0, s,
                     n, B, ncol, blk, cols);
}
template void pack_cols<double, double>(int, int, int, const double *const *, double *,
                                        hipStream_t);
template void pack_cols<float, double>(int, int, int, const double *const *, float *,
                                       hipStream_t);
template void pack_cols<float, float>(int, int, int, const float *const *, float *,
                                      hipStream_t);
template void unpack_cols<double, double>(int, int, int, const double *, double *const *,
                                          hipStream_t);
template void unpack_cols<float, double>(int, int, int, const float *, double *const *,
                                         hipStream_t);
template void unpack_cols<float, float>(int, int, int, const float *, float *const *,
                                        hipStream_t);

int dot_partials_f32(int n, const float *x, const float *y, double *part, hipStream_t s) {
  const int nb = lan_vec_blocks(n);
  hipLaunchKernelGGL(k_dot_part_f32, dim3(nb), dim3(kThreads), 0, s, n, x, y, part);
  return nb;
}

void scale_vec_f32(int n, float *x, double hostscale, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scale_f32, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x, hostscale);
}

size_t lan_part_doubles(int n, int nslice, int hcap) {
  // sized for FP64 storage, the larger of the two row-block counts
  const size_t ld = ((size_t)n + 63) / 64 * 64;
  const size_t nrb = (ld + gt_rows<double>() - 1) / gt_rows<double>();
  const size_t nbn = (ld + gn_rows<double>() - 1) / gn_rows<double>();
  const size_t per = std::max(std::max((size_t)hcap * nrb, nbn),
                              (size_t)lan_vec_blocks(n));
  return (size_t)nslice * per + 64;
}

namespace {
template <typename TV>
void lan_alpha_t(const LanBatch &L, int k, hipStream_t s) {
  const int nb = lan_vec_blocks(L.n);
  LanArgs<TV> a = make_args<TV>(L);
  hipLaunchKernelGGL(k_lan_alpha<TV>, dim3(nb, L.nslice), dim3(kThreads), 0, s, a, k);
  hipLaunchKernelGGL(k_lan_reduce<TV>, dim3(L.nslice), dim3(64), 0, s, a, nb, 0);
}

template <typename TV>
void lan_subtract_alpha_t(const LanBatch &L, int k, hipStream_t s) {
  const int nb = lan_vec_blocks(L.n);
  LanArgs<TV> a = make_args<TV>(L);
  hipLaunchKernelGGL(k_lan_sub_alpha<TV>, dim3(nb, L.nslice), dim3(kThreads), 0, s, a, k);
  hipLaunchKernelGGL(k_lan_reduce<TV>, dim3(L.nslice), dim3(64), 0, s, a, nb, 1);
}

template <typename TV>
void lan_cgs_pass_t(const LanBatch &L, int k, int pass, hipStream_t s) {
  LanArgs<TV> a = make_args<TV>(L);
  const int nrb = (int)((L.ld + gt_rows<TV>() - 1) / gt_rows<TV>());
  const int ncb = (k + 1 + kTcols - 1) / kTcols;
  // nontemporal basis loads: measured 6.5 vs 5.9 TB/s at k = 4096 (C2)
  hipLaunchKernelGGL((k_lan_gemvt<true, TV>), dim3(nrb, ncb, L.nslice), dim3(kThreads), 0,
                     s, a, k, pass, nrb);
  hipLaunchKernelGGL(k_lan_reduce_h<TV>, dim3(grid_for(k + 1), L.nslice), dim3(kThreads), 0,
                     s, a, k, pass, nrb);
  const int nbn = (int)((L.ld + gn_rows<TV>() - 1) / gn_rows<TV>());
  hipLaunchKernelGGL((k_lan_gemvn<true, TV>), dim3(nbn, L.nslice), dim3(kThreads), 0, s, a,
                     k, pass, nbn);
  hipLaunchKernelGGL(k_lan_reduce<TV>, dim3(L.nslice), dim3(64), 0, s, a, nbn, pass ? 3 : 2);
}

template <typename TV>
void lan_scale_new_t(const LanBatch &L, int k, hipStream_t s) {
  LanArgs<TV> a = make_args<TV>(L);
  const int nb = lan_vec_blocks(L.n);
  hipLaunchKernelGGL(k_lan_scale<TV>, dim3(nb, L.nslice), dim3(kThreads), 0, s, a, k);
}
}  // namespace

void lan_alpha(const LanBatch &L, int k, hipStream_t s) {
  if (L.vbytes == 4)
    lan_alpha_t<float>(L, k, s);
  else
    lan_alpha_t<double>(L, k, s);
}

void lan_subtract_alpha(const LanBatch &L, int k, hipStream_t s) {
  if (L.vbytes == 4)
    lan_subtract_alpha_t<float>(L, k, s);
  else
    lan_subtract_alpha_t<double>(L, k, s);
}

void lan_cgs_pass(const LanBatch &L, int k, int pass, hipStream_t s) {
  if (L.vbytes == 4)
    lan_cgs_pass_t<float>(L, k, pass, s);
  else
    lan_cgs_pass_t<double>(L, k, pass, s);
}

void lan_scale_new(const LanBatch &L, int k, hipStream_t s) {
  if (L.vbytes == 4)
    lan_scale_new_t<float>(L, k, s);
  else
    lan_scale_new_t<double>(L, k, s);
}

void tridiag_trace_above(int m, const double *d, const double *e, double thr,
                         double *lam, double *out, hipStream_t s) {
  // ws lives in out[2..5]
  double *ws = out + 2;
  hipLaunchKernelGGL(k_tri_prep, dim3(1), dim3(kThreads), 0, s, m, (size_t)0, d, e,
                     (const double *)nullptr, thr, ws);
  hipLaunchKernelGGL(k_tri_msect, dim3(m, 1), dim3(64), 0, s, m, (size_t)0, d, e,
                     (const double *)nullptr, thr, (const double *)ws, lam, 0);
  hipLaunchKernelGGL(k_tri_sum, dim3(1), dim3(64), 0, s, m, (size_t)0, (const double *)ws,
                     (const double *)lam, out);
}

void tridiag_trace_above_batch(int np, int m, size_t ld, const double *d, const double *e,
                               const double *thr, double *lam, double *ws, double *out,
                               hipStream_t s) {
  if (np <= 0 || m <= 0) return;
  hipLaunchKernelGGL(k_tri_prep, dim3(np), dim3(kThreads), 0, s, m, ld, d, e, thr, 0.0, ws);
  hipLaunchKernelGGL(k_tri_msect, dim3(m, np), dim3(64), 0, s, m, ld, d, e, thr, 0.0,
                     (const double *)ws, lam, 0);
  hipLaunchKernelGGL(k_tri_sum, dim3(np), dim3(64), 0, s, m, ld, (const double *)ws,
                     (const double *)lam, out);
}

void tridiag_eigvals(int m, const double *d, const double *e, double *lam,
                     hipStream_t s) {
  // uses lam[m..m+3] as workspace for the bounds: caller gives m+4 doubles
  double *ws = lam + m;
  hipLaunchKernelGGL(k_tri_prep, dim3(1), dim3(kThreads), 0, s, m, (size_t)0, d, e,
                     (const double *)nullptr, -HUGE_VAL, ws);
  hipLaunchKernelGGL(k_tri_msect, dim3(m, 1), dim3(64), 0, s, m, (size_t)0, d, e,
                     (const double *)nullptr, -HUGE_VAL, (const double *)ws, lam, 1);
}

void kpm_intchx(int Mdeg, const double *mu, int npts, const double *th,
                double th0, double *yi, hipStream_t s) {
  hipLaunchKernelGGL(k_intchx, dim3(grid_for(npts)), dim3(kThreads), 0, s,
                     Mdeg, mu, npts, th, th0, yi);
}

int col_part_blocks(int n) { return lan_vec_blocks(n); }

void col_dots(int n, int ncol, const double *X, size_t ldx, const double *Y,
              size_t ldy, double *part, double *out, hipStream_t s) {
  if (ncol <= 0) return;
  const int nb = lan_vec_blocks(n);
  hipLaunchKernelGGL(k_col_dots, dim3(nb, ncol), dim3(kThreads), 0, s, n, X,
                     ldx, Y, ldy, part);
  hipLaunchKernelGGL(k_col_reduce, dim3(ncol), dim3(64), 0, s, ncol, nb,
                     (const double *)part, out);
}

void col_scale(int n, int ncol, double *X, size_t ld, const double *f,
               hipStream_t s) {
  if (ncol <= 0 || n == 0) return;
  hipLaunchKernelGGL(k_col_scale, dim3(grid_for(n), ncol), dim3(kThreads), 0,
                     s, n, X, ld, f);
}

void col_axpy(int n, int ncol, const double *a, const double *X, size_t ldx,
              double *Y, size_t ldy, hipStream_t s) {
  if (ncol <= 0 || n == 0) return;
  hipLaunchKernelGGL(k_col_axpy, dim3(grid_for(n), ncol), dim3(kThreads), 0,
                     s, n, a, X, ldx, Y, ldy);
}

void rand_fill(size_t n, double *x, uint64_t seed, uint64_t offset,
               int rademacher, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_rand, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x,
                     seed, offset, rademacher);
}

}  // namespace pevsl
