// chebpol.cpp -- design of the Chebyshev "delta" filter of one slice.
//
// Host code: the design is scalar work of O(deg^2) at most and runs once per
// slice.  The algorithm follows SRC/chebpol.c:10-493:
//   * damping factors (Jackson / Lanczos sigma), SRC/chebpol.c:34-61;
//   * end slices: a scaled Chebyshev polynomial of growing degree, expressed
//     in the T_k basis of [-1,1], until its value at the inner edge drops
//     below thresh_ext (SRC/chebpol.c:166-255);
//   * interior slices: for growing degree m, balance the centre theta_c of a
//     damped delta expansion by Newton/bisection so that p(a) == p(b)
//     (SRC/chebpol.c:284-332), accept the first m whose edge values are below
//     thresh_int * p(gamma) (SRC/chebpol.c:433-478).
// The filter is applied on the GPU (kernels.hip: cheb_step_*).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace {

int check_intv(const double *intv) {
  if (intv[0] >= intv[1]) return -1;
  if (intv[0] >= intv[3] || intv[1] <= intv[2]) return -2;
  return 0;
}

void damping_factors(int m, int damping, std::vector<double> &jac) {
  jac.resize((size_t)m + 1);
  const double dm = (double)m;
  double thetJ = 0, thetL = 0, a1 = 0, a2 = 0;
  if (damping == 1) {
    thetJ = M_PI / (dm + 2);
    a1 = 1 / (dm + 2);
    a2 = sin(thetJ);
  } else if (damping == 2) {
    thetL = M_PI / (dm + 1);
  }
  jac[0] = 0.5;  // halves the zeroth Chebyshev term
  for (int k = 1; k <= m; k++) {
    if (damping == 0) {
      jac[(size_t)k] = 1.0;
    } else if (damping == 1) {
      const int k1 = k + 1;
      jac[(size_t)k] = a1 * sin(k1 * thetJ) / a2 + (1 - k1 * a1) * cos(k * thetJ);
    } else {
      jac[(size_t)k] = sin(k * thetL) / (k * thetL);
    }
  }
}

// p(x) for p = sum_{k<=m} mu_k T_k by the three-term recurrence
double cheb_eval(int m, const double *mu, double x) {
  double tkm1 = 0.0, tk = 1.0, y = mu[0];
  for (int k = 0; k < m; k++) {
    const double tkp1 = (k == 0 ? 1.0 : 2.0) * x * tk - tkm1;
    tkm1 = tk;
    tk = tkp1;
    y += mu[k + 1] * tk;
  }
  return y;
}

// sum_j v_j cos(j th) jac_j : difference of the expansion at the two edges
double edge_difference(const std::vector<double> &v, int m, double th,
                       const std::vector<double> &jac) {
  double f = 0.0;
  for (int j = 0; j <= m; j++) f += v[(size_t)j] * cos(j * th) * jac[(size_t)j];
  return f;
}

// Balance the centre angle; 1 if the degree is too low to balance.
int balance_center(int m, const std::vector<double> &v,
                   const std::vector<double> &jac, double tha, double thb,
                   double *mu, double *thc_out) {
  const double tolBal = fabs(tha - thb) * 1.e-13;
  double thc = 0.5 * (tha + thb);
  const double fb = edge_difference(v, m, thb, jac);
  const double fa = edge_difference(v, m, tha, jac);
  if (fa > 0 || fb < 0) return 1;
  for (int it = 0; it <= 30; it++) {
    const double fval = edge_difference(v, m, thc, jac);
    double d = 0.0;
    for (int j = 1; j <= m; j++)
      d += jac[(size_t)j] * j * sin(j * thc) * v[(size_t)j];
    double thN = thc + fval / d;
    if (fabs(fval) < tolBal || fabs(thc - thN) < DBL_EPSILON * fabs(thc)) break;
    if (fval > 0) {
      if (thN < thb || thN > tha) thN = 0.5 * (thc + tha);
      thb = thc;
    } else {
      if (thN < thb || thN > tha) thN = 0.5 * (thc + thb);
      tha = thc;
    }
    thc = thN;
  }
  for (int j = 0; j <= m; j++) mu[j] = cos(j * thc) * jac[(size_t)j];
  *thc_out = thc;
  return 0;
}

// End-interval filter: scaled Chebyshev polynomial in the T_k basis.
void end_filter(pevsl_Polparams *pol, double aIn, double bIn) {
  int max_deg = pol->max_deg;
  double thresh = pol->thresh_ext;
  if (pol->deg > 0) {  // forced degree
    thresh = -1.0;
    max_deg = pol->deg;
  }
  const double del = 0.1 * sqrt((bIn - aIn) * 0.5);
  double a, b, x, gam;
  if (bIn >= 1.0 - 1e-13) {  // right end: damp [-1, aIn - del]
    x = aIn;
    b = x - del;
    a = -1.0;
    gam = 1.0;
  } else {                   // left end: damp [bIn + del, 1]
    x = bIn;
    a = x + del;
    b = 1.0;
    gam = -1.0;
  }
  const double e = (b - a) * 0.5, c = (b + a) * 0.5;
  const double sigma1 = e / (gam - c);
  double sigma = sigma1;
  const size_t m1 = (size_t)max_deg + 1;
  std::vector<double> t0(m1 + 1, 0.0), t1(m1 + 1, 0.0), tn(m1 + 1, 0.0);
  t0[0] = 1.0;
  t1[0] = -c * (sigma1 / e);
  t1[1] = sigma1 / e;
  double g0 = 1.0, g1 = (x - c) * sigma1 / e;
  int mbest = 0;
  if (g1 < thresh) {
    mbest = 1;
  } else {
    for (int k = 2; k <= max_deg; k++) {
      const double sigma_new = 1.0 / (2.0 / sigma1 - sigma);
      const double s1 = sigma_new / e, s2 = 2 * s1 * c, s3 = sigma * sigma_new;
      // (x - c) * T-expansion: x T_j = (T_{j+1} + T_{j-1})/2 in coefficients
      for (int j = 2; j <= k; j++) tn[(size_t)j - 1] = s1 * (t1[(size_t)j] + t1[(size_t)j - 2]);
      tn[1] = tn[1] + s1 * t1[0];
      tn[0] = s1 * t1[1];
      tn[(size_t)k] = s1 * t1[(size_t)k - 1];
      for (int j = 0; j <= k; j++)
        tn[(size_t)j] = tn[(size_t)j] - s2 * t1[(size_t)j] - s3 * t0[(size_t)j];
      for (int j = 0; j <= k; j++) {
        t0[(size_t)j] = t1[(size_t)j];
        t1[(size_t)j] = tn[(size_t)j];
      }
      const double gnew = 2 * (x - c) * s1 * g1 - s3 * g0;
      g0 = g1;
      g1 = gnew;
      mbest = k;
      if (g1 < thresh) break;
      sigma = sigma_new;
    }
  }
  memcpy(pol->mu, t1.data(), sizeof(double) * (size_t)(mbest + 1));
  pol->deg = mbest;
  pol->bar = g1;
  pol->gam = gam;
}

}  // namespace

extern "C" {

void pEVSL_SetPolDef(pevsl_Polparams *pol) {
  pol->max_deg = 5000;
  pol->min_deg = 2;
  pol->damping = 2;
  pol->thresh_ext = 0.20;
  pol->thresh_int = 0.8;
  pol->mu = NULL;
  pol->deg = 0;
  pol->intvtol = 1e-9;
  pol->type = 0;
  pol->cc = pol->dd = pol->gam = pol->bar = 0.0;
}

void pEVSL_FreePol(pevsl_Polparams *pol) { PEVSL_FREE(pol->mu); }

int pEVSL_FindPol(double *intv, pevsl_Polparams *pol) {
  if (check_intv(intv) < 0) return -1;
  int max_deg = pol->max_deg;
  PEVSL_CALLOC(pol->mu, (size_t)max_deg + 1, double);
  double *mu = pol->mu;
  double aa = intv[0] > intv[2] ? intv[0] : intv[2];
  double bb = intv[1] < intv[3] ? intv[1] : intv[3];
  if (intv[0] < intv[2] || intv[1] > intv[3])
    fprintf(stdout, " warning: interval (%e, %e) is adjusted to (%e, %e)\n",
            intv[0], intv[1], aa, bb);
  const double lmin = intv[2], lmax = intv[3];
  const double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  pol->cc = cc;
  pol->dd = dd;
  // map [lmin, lmax] -> [-1, 1]
  aa = (aa - cc) / dd;
  bb = (bb - cc) / dd;
  if (aa < -1.0) aa = -1.0;
  if (bb > 1.0) bb = 1.0;
  const double thb = acos(bb), tha = acos(aa);
  const double IntTol = pol->intvtol;
  if (aa - IntTol <= -1.0) {
    pol->type = 1;
    end_filter(pol, -1.0, bb);
    return 0;
  }
  if (bb + IntTol >= 1.0) {
    pol->type = 2;
    end_filter(pol, aa, 1.0);
    return 0;
  }
  pol->type = 0;
  // The reference computes the starting degree as `2 + (int) 0.5/(bb-aa)`
  // (SRC/chebpol.c:434); the cast applies to 0.5, so the search starts at 2.
  const int min_deg = 2;
  double thresh = pol->thresh_int;
  if (pol->deg > 0) {
    thresh = -1;
    max_deg = pol->deg;
  }
  std::vector<double> v((size_t)pol->max_deg + 1, 0.0), jac;
  for (int j = 0; j < min_deg; j++) v[(size_t)j] = cos(j * thb) - cos(j * tha);
  const double itv[2] = {aa, bb};
  double vals[2] = {0.0, 0.0}, t = 0.0, gam = 0.0, thc = 0.0;
  int m;
  for (m = min_deg; m < max_deg; m++) {
    damping_factors(m, pol->damping, jac);
    v[(size_t)m] = cos(m * thb) - cos(m * tha);
    if (balance_center(m, v, jac, tha, thb, mu, &thc)) continue;
    gam = cos(thc);
    t = cheb_eval(m, mu, gam);
    vals[0] = cheb_eval(m, mu, itv[0]);
    vals[1] = cheb_eval(m, mu, itv[1]);
    if (vals[0] <= t * thresh && vals[1] <= t * thresh) {
      m++;
      break;
    }
  }
  const int mbest = m - 1;
  for (int j = 0; j <= mbest; j++) mu[j] /= t;  // p(gam) = 1
  pol->bar = (vals[0] < vals[1] ? vals[0] : vals[1]) / t;
  pol->gam = gam;
  pol->deg = mbest;
  return 0;
}

}  // extern "C"
