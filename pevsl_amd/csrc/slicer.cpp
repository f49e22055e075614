// slicer.cpp -- the learned spectrum slicer.  No counterpart in the
// reference: it replaces the outputs of pEVSL_Kpmdos (SRC/kpmdos.c:36) and
// pEVSL_SpslicerKpm (SRC/spslicer.c:62) by a network's estimate, and the
// filter degrees then follow from pEVSL_FindPol on the proposed slices as in
// the reference driver (TESTS/Lap/LapPLanN.c:211-225).
//
// Pipeline (pEVSL_SpslicerMlp):
//   1. a cheap KPM run, PEVSL_SLICER_MDEG moments from PEVSL_SLICER_NVEC
//      random vectors with Jackson damping (pEVSL_Kpmdos, on the GPU);
//   2. features and a prior from those moments (pEVSL_SlicerFeaturesFromMu);
//   3. the network (mlp.hip) turns them into the DOS of the interval on
//      PEVSL_SLICER_NBINS bins plus the log of the fraction of the spectrum
//      inside the interval; its output is added to the prior, so an untrained
//      network reproduces the cheap KPM estimate;
//   4. equal-share cuts of the binned DOS (pEVSL_SlicerFromBins).
// The network is trained on pairs whose targets are full-KPM bin masses
// (pEVSL_KpmBinMasses on the moments of a full pEVSL_Kpmdos run), and full
// KPM judges a proposal (pEVSL_SlicerJudge); a caller falls back to
// pEVSL_SpslicerKpm when the judge rejects it.
//
// DOS convention (that of pEVSL_Kpmdos): with x = (t - c)/h over the frame
// [lmin, lmax], c = (lmax + lmin)/2, h = (lmax - lmin)/2, the fraction of
// the spectrum below x is
//   F(x) = mu_0 (pi - acos x) - sum_{k>=1} mu_k sin(k acos x) / k,
// which is the closed form pEVSL_Kpmdos uses for the eigenvalue count
// (SRC/kpmdos.c:127,155,182) and SpslicerKpm for its grid (intChx,
// SRC/spslicer.c:12-38).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace pevsl {
namespace {

inline double clampx(double x) { return std::min(1.0, std::max(-1.0, x)); }

// F(x) for moments mu[0..Mdeg]
double kpm_cdf(const double *mu, int Mdeg, double x) {
  const double th = std::acos(clampx(x));
  double s = mu[0] * (M_PI - th);
  // sin(k th) by the recurrence s_{k+1} = 2 cos(th) s_k - s_{k-1}
  const double c2 = 2.0 * std::cos(th);
  double sm1 = 0.0, sk = std::sin(th);
  for (int k = 1; k <= Mdeg; k++) {
    s -= mu[k] * sk / k;
    const double sn = c2 * sk - sm1;
    sm1 = sk;
    sk = sn;
  }
  return s;
}

struct Frame {
  double c, h;
  explicit Frame(const double *intv)
      : c((intv[3] + intv[2]) / 2.0), h((intv[3] - intv[2]) / 2.0) {}
  double x(double t) const { return clampx((t - c) / h); }
};

bool bad_intv(const double *intv) {
  return !intv || !(intv[0] < intv[1]) || !(intv[2] < intv[3]) || intv[0] >= intv[3] ||
         intv[1] <= intv[2];
}

}  // namespace
}  // namespace pevsl

using namespace pevsl;

extern "C" {

int pEVSL_KpmBinMasses(const double *mu, int Mdeg, const double *intv, int nbins,
                       double *masses, double *frac) {
  if (!mu || Mdeg < 0 || nbins < 1 || bad_intv(intv)) return -1;
  const Frame f(intv);
  const double a = std::max(intv[0], intv[2]), b = std::min(intv[1], intv[3]);
  double prev = kpm_cdf(mu, Mdeg, f.x(a));
  const double f0 = prev;
  for (int i = 1; i <= nbins; i++) {
    const double t = i == nbins ? b : a + (b - a) * i / nbins;
    const double cur = kpm_cdf(mu, Mdeg, f.x(t));
    if (masses) masses[i - 1] = cur - prev;
    prev = cur;
  }
  if (frac) *frac = prev - f0;
  return 0;
}

int pEVSL_SlicerFromBins(const double *masses, int nbins, const double *intv, int n_int,
                         double *sli) {
  if (!masses || nbins < 1 || n_int < 1 || !sli || !intv || !(intv[0] < intv[1])) return -1;
  const double a = intv[0], b = intv[1], w = (b - a) / nbins;
  double tot = 0.0;
  for (int i = 0; i < nbins; i++) tot += std::max(masses[i], 0.0);
  sli[0] = a;
  sli[n_int] = b;
  if (!(tot > 0.0)) {  // no mass: uniform cuts
    for (int j = 1; j < n_int; j++) sli[j] = a + (b - a) * j / n_int;
    return 1;
  }
  int bin = 0;
  double below = 0.0;  // mass of bins [0, bin)
  for (int j = 1; j < n_int; j++) {
    const double target = tot * j / n_int;
    while (bin < nbins - 1 && below + std::max(masses[bin], 0.0) < target) {
      below += std::max(masses[bin], 0.0);
      bin++;
    }
    const double m = std::max(masses[bin], 0.0);
    const double fr = m > 0.0 ? std::min(1.0, std::max(0.0, (target - below) / m)) : 0.0;
    sli[j] = std::max(sli[j - 1], a + w * (bin + fr));
  }
  return 0;
}

int pEVSL_SlicerFeaturesFromMu(const double *mu, int Mdeg, int n, const double *intv,
                               float *feat, float *prior) {
  if (!mu || Mdeg < PEVSL_SLICER_MDEG || n < 1 || bad_intv(intv) || !feat) return -1;
  const Frame f(intv);
  const double a = std::max(intv[0], intv[2]), b = std::min(intv[1], intv[3]);
  const double xa = f.x(a), xb = f.x(b);
  int k = 0;
  for (int j = 1; j <= PEVSL_SLICER_MDEG; j++) feat[k++] = (float)(M_PI * mu[j]);
  feat[k++] = (float)xa;
  feat[k++] = (float)xb;
  feat[k++] = (float)(std::acos(xa) / M_PI);
  feat[k++] = (float)(std::acos(xb) / M_PI);
  feat[k++] = (float)(10.0 * (std::acos(xa) - std::acos(xb)) / M_PI);
  feat[k++] = (float)(std::log10((double)n) / 6.0);
  // cheap-KPM fraction of the spectrum in [a, b] and the shape of its CDF
  const double fa = kpm_cdf(mu, PEVSL_SLICER_MDEG, xa), fb = kpm_cdf(mu, PEVSL_SLICER_MDEG, xb);
  const double fr = std::max(fb - fa, 1e-12);
  feat[k++] = (float)(std::log(fr) / 10.0);
  for (int j = 1; j <= 15; j++) {
    const double t = a + (b - a) * j / 16.0;
    const double cj = (kpm_cdf(mu, PEVSL_SLICER_MDEG, f.x(t)) - fa) / fr;
    feat[k++] = (float)(cj - j / 16.0);
  }
  // Fourier features of both ends in the Chebyshev angle: the DOS is a
  // cosine series in that angle, so these let the network place structure
  // (band edges, van Hove points) relative to the interval
  const double tha = std::acos(xa), thb = std::acos(xb);
  for (int j = 1; j <= PEVSL_SLICER_NFREQ; j++) {
    feat[k++] = (float)std::cos(j * tha);
    feat[k++] = (float)std::sin(j * tha);
    feat[k++] = (float)std::cos(j * thb);
    feat[k++] = (float)std::sin(j * thb);
  }
  if (prior) {
    std::vector<double> m((size_t)PEVSL_SLICER_NBINS);
    double fr2 = 0.0;
    pEVSL_KpmBinMasses(mu, PEVSL_SLICER_MDEG, intv, PEVSL_SLICER_NBINS, m.data(), &fr2);
    double tot = 0.0;
    for (double v : m) tot += std::max(v, 0.0);
    const double floor_ = 1e-9 * std::max(tot, 1e-300);
    for (int i = 0; i < PEVSL_SLICER_NBINS; i++)
      prior[i] = (float)std::log(std::max(m[(size_t)i], floor_) / std::max(tot, 1e-300));
    prior[PEVSL_SLICER_NBINS] = (float)std::log(std::max(fr2, 1e-12));
  }
  return 0;
}

int pEVSL_SlicerFeatures(pevsl_Data *pevsl, const double *intv, float *feat, float *prior) {
  if (!pevsl || bad_intv(intv)) return -1;
  double mu[PEVSL_SLICER_MDEG + 1], ecnt = 0.0;
  double iv[4] = {intv[0], intv[1], intv[2], intv[3]};
  const int rc = pEVSL_Kpmdos(pevsl, PEVSL_SLICER_MDEG, 1, PEVSL_SLICER_NVEC, iv, 1, 0,
                              MPI_COMM_WORLD, mu, &ecnt);
  if (rc) return rc;
  return pEVSL_SlicerFeaturesFromMu(mu, PEVSL_SLICER_MDEG, pevsl->N, intv, feat, prior);
}

int pEVSL_SpslicerMlp(pevsl_Data *pevsl, pevsl_Mlp *mlp, const double *intv, int n_int,
                      double *sli, double *ecnt) {
  if (!pevsl || !mlp || n_int < 1 || !sli || bad_intv(intv)) return -1;
  int nin = 0, nout = 0, nsoft = 0;
  pEVSL_MlpShape(mlp, &nin, nullptr, nullptr, &nout, &nsoft, nullptr);
  if (nin != PEVSL_SLICER_NIN || nout != PEVSL_SLICER_NBINS + 1 ||
      nsoft != PEVSL_SLICER_NBINS)
    return -4;  // not a slicer network
  std::vector<float> feat((size_t)nin), prior((size_t)nout), y((size_t)nout);
  int rc = pEVSL_SlicerFeatures(pevsl, intv, feat.data(), prior.data());
  if (rc) return rc;
  rc = pEVSL_MlpForward(mlp, feat.data(), prior.data(), 1, y.data());
  if (rc) return rc;
  std::vector<double> m((size_t)PEVSL_SLICER_NBINS);
  for (int i = 0; i < PEVSL_SLICER_NBINS; i++) m[(size_t)i] = y[(size_t)i];
  double iv[2] = {std::max(intv[0], intv[2]), std::min(intv[1], intv[3])};
  rc = pEVSL_SlicerFromBins(m.data(), PEVSL_SLICER_NBINS, iv, n_int, sli);
  sli[0] = intv[0];
  sli[n_int] = intv[1];
  if (ecnt) *ecnt = (double)pevsl->N * std::exp((double)y[PEVSL_SLICER_NBINS]);
  return rc < 0 ? rc : 0;
}

int pEVSL_SlicerJudge(const double *mu, int Mdeg, const double *intv, double n, int n_int,
                      const double *sli, double *counts, double *imbalance) {
  if (!mu || Mdeg < 0 || n_int < 1 || !sli || bad_intv(intv)) return -1;
  const Frame f(intv);
  double prev = kpm_cdf(mu, Mdeg, f.x(sli[0])), mx = 0.0, tot = 0.0;
  for (int j = 0; j < n_int; j++) {
    const double cur = kpm_cdf(mu, Mdeg, f.x(sli[j + 1]));
    const double cj = n * (cur - prev);
    if (counts) counts[j] = cj;
    mx = std::max(mx, cj);
    tot += cj;
    prev = cur;
  }
  if (imbalance) *imbalance = tot > 0.0 ? mx / (tot / n_int) - 1.0 : 0.0;
  return 0;
}

}  // extern "C"
