// noalloc.hip -- pEVSL_ChebAvNoAlloc: the Chebyshev filter of pEVSL_ChebAv,
// y = p((A - cc)/dd) v, on vectors that are already in HBM, for callers that
// must not allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The filter's work vectors (three n-vectors in FP64; x, y and three columns
// in FP32 on the mixed path, which fit in the same space) live in `scratch`,
// a file-scope device array of kScratchDoubles doubles that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel: the first call on a device resolves its device
// address with hipGetSymbolAddress (that device current) and keeps it per
// device; kernels receive that address.
//
// All calls on one GPU share its array and it keeps the last call's
// contents (nothing in it needs to start at zero: every work vector is
// written before it is read).  Calls that overlap on one GPU race.
#include <mutex>

#include "internal.hpp"

namespace {

// Rows of the largest matrix the entry point accepts (2^21 = 2,097,152).
constexpr size_t kMaxRows = size_t(1) << 21;
// three n-vectors of doubles; 5n floats on the FP32 filter path fit as well
constexpr size_t kScratchDoubles = 3 * kMaxRows + 64;  // 48 MiB + 512 B

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_mu;

// Device address of `scratch` on device `dev`, resolved once per device with
// that device current.  Returns nullptr if the runtime cannot resolve it.
double *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[dev]) return g_scratch[dev];
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return nullptr;
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p) {
    fprintf(stderr, "pEVSL_ChebAvNoAlloc: hipGetSymbolAddress failed on device %d: %s\n", dev,
            hipGetErrorString(e));
    return nullptr;
  }
  g_scratch[dev] = static_cast<double *>(p);
  return g_scratch[dev];
}

}  // namespace

using namespace pevsl;

extern "C" {

size_t pEVSL_ChebAvNoAllocMaxRows(void) { return kMaxRows; }

int pEVSL_ChebAvNoAlloc(pevsl_Data *pevsl, pevsl_Polparams *pol, const double *d_v,
                        double *d_y) {
  Ctx &c = ctx_of(pevsl);
  if (!c.A || pevsl->ifGenEv) {
    fprintf(stderr, "pEVSL_ChebAvNoAlloc: needs a device matrix (pEVSL_SetAParcsr) and the "
                    "standard problem\n");
    return PEVSL_NOT_IMPLEMENT;
  }
  const int n = pevsl->n;
  if (n != c.A->n || !d_v || !d_y || !pol || (pol->deg > 0 && !pol->mu)) return -4;
  if (n <= 0 || (size_t)n > kMaxRows) {
    fprintf(stderr, "pEVSL_ChebAvNoAlloc: n = %d rows exceeds the library's work array "
                    "(%zu rows)\n", n, kMaxRows);
    return -5;
  }
  double *w = scratch_on(c.device);
  if (!w) return -6;
  const double t0 = wtime();
  cheb_filter_col(c, pol, d_v, d_y, w);
  PV_HIP(hipStreamSynchronize(c.stream));
  pevsl->stats->t_polAv += wtime() - t0;
  pevsl->stats->n_polAv++;
  pevsl->stats->n_mvA += (size_t)pol->deg;
  return 0;
}

}  // extern "C"
