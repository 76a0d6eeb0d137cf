// filters.cpp -- polynomial filter construction (host scalar set-up).
//
// Same filter as SRC/chebpoly.c:15-501 of EVSL 1.1.1 (degree search,
// damping, Newton balancing of the delta expansion, end-interval scaled
// Chebyshev), arranged for throughput: the per-degree candidates of the
// middle-interval search are independent, so they are evaluated in parallel
// blocks of degrees and the smallest accepted degree is taken -- the same
// degree and coefficients the sequential search finds.
#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"

namespace evsl {

// damping factors; jac[0] = 1/2 carries the half weight of T_0
// (SRC/chebpoly.c:39-66)
int dampcf(int m, int damping, double* jac) {
  const double dm = (double)m;
  jac[0] = 0.5;
  if (damping == 1) {
    const double th = EVSL_PI / (dm + 2), a1 = 1 / (dm + 2), a2 = std::sin(th);
    for (int k = 1; k <= m; k++)
      jac[k] = a1 * std::sin((k + 1) * th) / a2 + (1 - (k + 1) * a1) * std::cos(k * th);
  } else if (damping == 2) {
    const double th = EVSL_PI / (dm + 1);
    for (int k = 1; k <= m; k++) jac[k] = std::sin(k * th) / (k * th);
  } else {
    for (int k = 1; k <= m; k++) jac[k] = 1.0;
  }
  return 0;
}

// yi = sum_k mu[k] T_k(xi) (SRC/chebpoly.c:105-135)
int chebxPltd(int m, const double* mu, int npts, const double* xi, double* yi) {
  for (int j = 0; j < npts; ++j) {
    const double x = xi[j];
    double tkm1 = 0.0, tk = 1.0, y = mu[0];
    for (int k = 0; k < m; ++k) {
      const double tkp1 = (k == 0 ? 1.0 : 2.0) * x * tk - tkm1;
      tkm1 = tk;
      tk = tkp1;
      y += mu[k + 1] * tk;
    }
    yi[j] = y;
  }
  return 0;
}

namespace {

double dif_eval(const double* v, int m, double thc, const double* jac) {
  double f = 0.0;
  for (int j = 0; j <= m; j++) f += v[j] * std::cos(j * thc) * jac[j];
  return f;
}

// balance p(a) == p(b) by safeguarded Newton on the angle (SRC/chebpoly.c:288-336)
bool balance(int m, const double* v, const double* jac, double tha, double thb, double* mu,
             double* thc_out) {
  const int max_it = 30;
  const double tol = std::fabs(tha - thb) * 1.e-13;
  double thc = 0.5 * (tha + thb);
  const double fb = dif_eval(v, m, thb, jac);
  const double fa = dif_eval(v, m, tha, jac);
  if (fa > 0 || fb < 0) return false;
  for (int it = 0; it <= max_it; it++) {
    const double f = dif_eval(v, m, thc, jac);
    double d = 0.0;
    for (int j = 1; j <= m; j++) d += jac[j] * j * std::sin(j * thc) * v[j];
    double thN = thc + f / d;
    if (std::fabs(f) < tol || std::fabs(thc - thN) < DBL_EPSILON * std::fabs(thc)) break;
    if (f > 0) {
      if (thN < thb || thN > tha) thN = 0.5 * (thc + tha);
      thb = thc;
      thc = thN;
    } else {
      if (thN < thb || thN > tha) thN = 0.5 * (thc + thb);
      tha = thc;
      thc = thN;
    }
  }
  for (int j = 0; j <= m; j++) mu[j] = std::cos(j * thc) * jac[j];
  *thc_out = thc;
  return true;
}

// end-interval filter: T_k((t-c)/e) / T_k((gam-c)/e) expanded in the
// Chebyshev basis of [-1,1] by a coefficient recurrence (SRC/chebpoly.c:170-259)
void end_filter(polparams* pol, double aIn, double bIn) {
  int max_deg = pol->max_deg;
  double thresh = pol->thresh_ext;
  if (pol->deg > 0) {
    thresh = -1.0;
    max_deg = pol->deg;
  }
  const double del = 0.1 * std::sqrt((bIn - aIn) * 0.5);
  double a, b, x, gam;
  if (bIn >= (1.0 - 1e-13)) {
    x = aIn;
    b = x - del;
    a = -1.0;
    gam = 1.0;
  } else {
    x = bIn;
    a = x + del;
    b = 1.0;
    gam = -1.0;
  }
  const double e = (b - a) * 0.5, c = (b + a) * 0.5;
  const double sigma1 = e / (gam - c);
  double sigma = sigma1;
  const int m1 = max_deg + 1;
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
      const double sn = 1.0 / (2.0 / sigma1 - sigma);
      const double s1 = sn / e, s2 = 2 * s1 * c, s3 = sigma * sn;
      for (int j = 2; j <= k; j++) tn[j - 1] = s1 * (t1[j] + t1[j - 2]);
      tn[1] = tn[1] + s1 * t1[0];
      tn[0] = s1 * t1[1];
      tn[k] = s1 * t1[k - 1];
      for (int j = 0; j <= k; j++) tn[j] = tn[j] - s2 * t1[j] - s3 * t0[j];
      for (int j = 0; j <= k; j++) {
        t0[j] = t1[j];
        t1[j] = tn[j];
      }
      const double gn = 2 * (x - c) * s1 * g1 - s3 * g0;
      g0 = g1;
      g1 = gn;
      mbest = k;
      if (g1 < thresh) break;
      sigma = sn;
    }
  }
  std::memcpy(pol->mu, t1.data(), (mbest + 1) * sizeof(double));
  pol->deg = mbest;
  pol->bar = g1;
  pol->gam = gam;
}

struct Candidate {
  bool ok = false;    // balancing succeeded
  bool acc = false;   // thresholds met
  double t = 0.0, thc = 0.0, v0 = 0.0, v1 = 0.0;
};

// evaluate degree m of the middle-interval search
Candidate eval_degree(int m, int damping, const double* v, double tha, double thb,
                      const double* itv, double thresh, std::vector<double>& jac,
                      std::vector<double>& mu) {
  Candidate c;
  dampcf(m, damping, jac.data());
  double thc;
  if (!balance(m, v, jac.data(), tha, thb, mu.data(), &thc)) return c;
  c.ok = true;
  c.thc = thc;
  const double gam = std::cos(thc);
  chebxPltd(m, mu.data(), 1, &gam, &c.t);
  double vals[2];
  chebxPltd(m, mu.data(), 2, itv, vals);
  c.v0 = vals[0];
  c.v1 = vals[1];
  c.acc = (vals[0] <= c.t * thresh && vals[1] <= c.t * thresh);
  return c;
}

}  // namespace
}  // namespace evsl

using namespace evsl;

extern "C" {

// SRC/chebpoly.c:15-24
void set_pol_def(polparams* pol) {
  pol->max_deg = 10000;
  pol->min_deg = 2;
  pol->damping = 2;
  pol->thresh_ext = 0.20;
  pol->thresh_int = 0.8;
  pol->mu = NULL;
  pol->deg = 0;
  pol->intvtol = 1e-9;
}

void free_pol(polparams* pol) {
  if (pol->mu) std::free(pol->mu);
  pol->mu = NULL;
}

// SRC/chebpoly.c:363-490
int find_pol(double* intv, polparams* pol) {
  const double a0 = intv[0], b0 = intv[1], lmin = intv[2], lmax = intv[3];
  if (a0 >= b0 || a0 >= lmax || b0 <= lmin) return -1;
  int max_deg = pol->max_deg;
  const int cap = std::max(max_deg, pol->deg) + 2;
  double* mu = (double*)std::calloc(cap, sizeof(double));
  if (!mu) return EVSL_ERR_ALLOC;
  pol->mu = mu;
  const double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  pol->cc = cc;
  pol->dd = dd;
  double aa = (std::max(a0, lmin) - cc) / dd, bb = (std::min(b0, lmax) - cc) / dd;
  aa = std::max(aa, -1.0);
  bb = std::min(bb, 1.0);
  const double thb = std::acos(bb), tha = std::acos(aa);
  const double itol = pol->intvtol;
  if (aa - itol <= -1.0) {
    pol->type = 1;
    end_filter(pol, -1.0, bb);
    return 0;
  }
  if (bb + itol >= 1.0) {
    pol->type = 2;
    end_filter(pol, aa, 1.0);
    return 0;
  }
  pol->type = 0;
  const double itv[2] = {aa, bb};
  int min_deg = (int)(2 + 0.5 / (bb - aa));
  min_deg = std::max(min_deg, 2);
  min_deg = std::min(min_deg, (max_deg + 1) / 2);
  double thresh = pol->thresh_int;
  const bool forced = pol->deg > 0;
  if (forced) {
    thresh = -1;
    max_deg = pol->deg;
  }
  std::vector<double> v(max_deg + 2, 0.0);
  for (int j = 0; j <= max_deg; j++) v[j] = std::cos(j * thb) - std::cos(j * tha);

  int best = -1;  // degree whose coefficients are kept
  Candidate cbest;
  if (forced) {
    // no threshold can pass: the sequential search ends with the highest
    // degree whose balancing succeeded
    std::vector<double> jac(max_deg + 2), mtmp(max_deg + 2);
    for (int m = max_deg - 1; m >= min_deg; --m) {
      Candidate c = eval_degree(m, pol->damping, v.data(), tha, thb, itv, thresh, jac, mtmp);
      if (c.ok) {
        best = m;
        cbest = c;
        break;
      }
    }
  } else {
    const int nthr = std::max(1, omp_get_max_threads());
    const int chunk = std::max(4, 2 * nthr);
    int last_ok = -1;
    Candidate c_last;
    for (int m0 = min_deg; m0 < max_deg && best < 0; m0 += chunk) {
      const int m1 = std::min(max_deg, m0 + chunk);
      std::vector<Candidate> res(m1 - m0);
#pragma omp parallel
      {
        std::vector<double> jac(max_deg + 2), mtmp(max_deg + 2);
#pragma omp for schedule(dynamic, 1)
        for (int m = m0; m < m1; ++m)
          res[m - m0] =
              eval_degree(m, pol->damping, v.data(), tha, thb, itv, thresh, jac, mtmp);
      }
      for (int m = m0; m < m1; ++m) {
        const Candidate& c = res[m - m0];
        if (!c.ok) continue;
        last_ok = m;
        c_last = c;
        if (c.acc) {
          best = m;
          cbest = c;
          break;
        }
      }
    }
    if (best < 0 && last_ok >= 0) {
      best = last_ok;
      cbest = c_last;
    }
  }
  int deg_out;
  if (best >= 0) {
    std::vector<double> jac(best + 2);
    dampcf(best, pol->damping, jac.data());
    for (int j = 0; j <= best; j++) mu[j] = std::cos(j * cbest.thc) * jac[j];
    deg_out = (cbest.acc ? best : max_deg - 1);
    const double t = cbest.t;
    for (int j = 0; j <= best; j++) mu[j] /= t;
    pol->bar = std::min(cbest.v0, cbest.v1) / t;
    pol->gam = std::cos(cbest.thc);
  } else {
    deg_out = max_deg - 1;
    pol->bar = 0.0;
    pol->gam = 0.0;
  }
  pol->deg = deg_out;
  return 0;
}

}  // extern "C"
