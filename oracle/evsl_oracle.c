/*
 * evsl_oracle.c -- TEST INFRASTRUCTURE ONLY (see evsl_oracle.h).
 *
 * CPU restatement of the EVSL 1.1.1 hot path, written to follow the
 * reference's operation order so that, on the same inputs and the same libc
 * rand() stream, it produces the reference's results.  Citations are
 * file:line relative to EVSL_1.1.1/.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   ora_dampcf        closed-form Jackson / Lanczos-sigma coefficients
 *   ora_chebxPltd     numpy.polynomial.chebyshev.chebval
 *   ora_find_pol      closed form of the scaled Chebyshev end filter; delta
 *                     expansion form, balancing and threshold invariants;
 *                     degree minimality re-derived with a bracketing root finder
 *   ora_ChebAv        dense eigendecomposition: y = Q p(L) Q^T v
 *   ora_kpmdos        dense eigendecomposition of the same random vectors
 *   ora_intChx        numerical quadrature of the KPM density
 *   ora_spslicer      closed-form boundaries of a density uniform in theta
 *                     (exact grid indices from arccos), invariants
 *   ora_CGS_DGKS      (I - Q Q^T) v and orthogonality
 *   ora_tridiag_eig   LAPACK (scipy.linalg.eigh_tridiagonal)
 *   ora_sym_eig       LAPACK (numpy.linalg.eigh)
 *   ora_ChebLanNr     closed-form Laplacian spectra and dense eigh counts
 *   ora_ChebLanTr     closed-form Laplacian spectra and dense eigh counts
 *   ora_LanBounds     Ritz values and residual norms of an independently
 *                     built (QR) Krylov space: lmin = theta_1 - ||r_1||
 *   ora_LanDos        Gauss quadrature of a full Krylov space: the smoothed
 *                     DOS in closed form from the eigenpairs of A (1e-12)
 *   ora_simpson       exact for cubics
 *   ora_spslicer2     closed-form boundaries of linear and constant densities
 *
 * ORA_PLANT (default 0) compiles one planted mistake into a copy of this
 * file, for the tests that show each pin above fails on a plausible error
 * (tests/test_oracle_pins.py builds those copies in a temporary directory):
 *   1 spslicer without the nearest-point snap   2 spslicer with fixed shares
 *   3 spslicer2 advancing t by the share        4 LanBounds without |beta s|
 *   5 LanBounds with the first Ritz-vector row  6 LanDos weights from the last row
 *   7 LanDos scaled without 1/nvec             8 LanDos with M = msteps (no cap)
 */
#ifndef ORA_PLANT
#define ORA_PLANT 0
#endif
#define _POSIX_C_SOURCE 199309L
#include "evsl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORTH_TOL 1e-14 /* INC/internal_header.h:21 */
#define NGS_MAX 2      /* INC/internal_header.h:24 */
#define EPSM (EVSL_DBL_EPS_MULT * DBL_EPSILON)

/* ------------------------------------------------------------------ */
/* operator                                                            */
/* ------------------------------------------------------------------ */
static int g_n = 0;
static const int *g_ia = NULL, *g_ja = NULL;
static const double *g_a = NULL;
static MVFunc g_func = NULL;
static void *g_data = NULL;
static int g_last_steps = 0;

void ora_set_csr(int n, const int *ia, const int *ja, const double *a) {
  g_n = n;
  g_ia = ia;
  g_ja = ja;
  g_a = a;
  g_func = NULL;
  g_data = NULL;
}

void ora_set_matvec(int n, MVFunc func, void *data) {
  g_n = n;
  g_func = func;
  g_data = data;
}

/* y = A x, row by row (SRC/spmat.c:257-280) */
void ora_matvec(const double *x, double *y) {
  if (g_func) {
    g_func((double *)x, y, g_data);
    return;
  }
  for (int i = 0; i < g_n; i++) {
    double r = 0.0;
    for (int j = g_ia[i]; j < g_ia[i + 1]; j++) r += x[g_ja[j]] * g_a[j];
    y[i] = r;
  }
}

/* ------------------------------------------------------------------ */
/* small BLAS-like helpers                                             */
/* ------------------------------------------------------------------ */
static double dot(int n, const double *x, const double *y) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}
static double nrm2(int n, const double *x) { return sqrt(dot(n, x, x)); }
static void scal(int n, double a, double *x) {
  for (int i = 0; i < n; i++) x[i] *= a;
}
static void axpy(int n, double a, const double *x, double *y) {
  for (int i = 0; i < n; i++) y[i] += a * x[i];
}
/* y = V(:,0:k) * c, V column-major with leading dimension ldv */
static void gemv_n(int n, int k, const double *V, size_t ldv, const double *c, double *y) {
  memset(y, 0, (size_t)n * sizeof(double));
  for (int j = 0; j < k; j++) {
    double t = c[j];
    const double *vj = V + (size_t)j * ldv;
    for (int i = 0; i < n; i++) y[i] += t * vj[i];
  }
}

/* interval validity (INC/internal_header.h:194-217) */
static int check_intv(const double *intv) {
  double a = intv[0], b = intv[1], lmin = intv[2], lmax = intv[3];
  if (a >= b) return -1;
  if (a >= lmax || b <= lmin) return -2;
  return 0;
}

/* ------------------------------------------------------------------ */
/* RNG and vectors (SRC/vect.c)                                        */
/* ------------------------------------------------------------------ */
/* U[-1,1] from libc rand() (SRC/vect.c:14-20) */
void ora_rand_double(int n, double *v) {
  double t = ((double)RAND_MAX) / 2.0;
  for (int i = 0; i < n; i++) v[i] = (rand() - t) / t;
}

/* Box-Muller keeping only the cosine branch, two rand() per entry, with the
 * reference's truncated 2*pi (SRC/vect.c:29-49) */
void ora_randn_double(int n, double *v) {
  const double two_pi = 2.0 * 3.1415926535;
  for (int i = 0; i < n; i++) {
    double u1 = rand() * (1.0 / RAND_MAX);
    double u2 = rand() * (1.0 / RAND_MAX);
    v[i] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
  }
}

/* SRC/vect.c:82-92 */
void ora_linspace(double a, double b, int num, double *arr) {
  double h = (num == 1 ? 0 : (b - a) / (num - 1));
  arr[0] = a;
  arr[num - 1] = b;
  for (int i = 1; i < num - 1; i++) arr[i] = a + i * h;
}

/* ------------------------------------------------------------------ */
/* polynomial filter construction (SRC/chebpoly.c)                     */
/* ------------------------------------------------------------------ */
/* SRC/chebpoly.c:15-24 */
void ora_set_pol_def(polparams *pol) {
  pol->max_deg = 10000;
  pol->min_deg = 2;
  pol->damping = 2;
  pol->thresh_ext = 0.20;
  pol->thresh_int = 0.8;
  pol->mu = NULL;
  pol->deg = 0;
  pol->intvtol = 1e-9;
}

/* damping factors; jac[0] carries the 1/2 of the zeroth Chebyshev term
 * (SRC/chebpoly.c:39-66) */
int ora_dampcf(int m, int damping, double *jac) {
  double thetJ = 0.0, thetL = 0.0, a1 = 0.0, a2 = 0.0, dm = (double)m;
  if (damping == 1) {
    thetJ = EVSL_PI / (dm + 2);
    a1 = 1 / (dm + 2);
    a2 = sin(thetJ);
  }
  if (damping == 2) thetL = EVSL_PI / (dm + 1);
  jac[0] = 0.5;
  for (int k = 1; k <= m; k++) {
    if (damping == 0) {
      jac[k] = 1.0;
    } else if (damping == 1) {
      int k1 = k + 1;
      jac[k] = a1 * sin(k1 * thetJ) / a2 + (1 - k1 * a1) * cos(k * thetJ);
    } else {
      jac[k] = sin(k * thetL) / (k * thetL);
    }
  }
  return 0;
}

/* sum_j v[j] cos(j thc) jac[j] (SRC/chebpoly.c:78-84) */
static double dif_eval(const double *v, int m, double thc, const double *jac) {
  double fval = 0.0;
  for (int j = 0; j <= m; j++) fval += v[j] * cos(j * thc) * jac[j];
  return fval;
}

/* yi = sum_k mu[k] T_k(xi) by the three-term recurrence (SRC/chebpoly.c:105-135) */
int ora_chebxPltd(int m, const double *mu, int npts, const double *xi, double *yi) {
  double *vkm1 = calloc(npts, sizeof(double));
  double *vk = malloc(npts * sizeof(double));
  double *vkp1 = malloc(npts * sizeof(double));
  for (int j = 0; j < npts; j++) {
    vk[j] = 1.0;
    yi[j] = mu[0];
  }
  for (int k = 0; k < m; k++) {
    double sc = (k == 0 ? 1.0 : 2.0);
    for (int j = 0; j < npts; j++) vkp1[j] = sc * xi[j] * vk[j] - vkm1[j];
    double *tmp = vkm1;
    vkm1 = vk;
    vk = vkp1;
    vkp1 = tmp;
    for (int j = 0; j < npts; j++) yi[j] += mu[k + 1] * vk[j];
  }
  free(vkm1);
  free(vk);
  free(vkp1);
  return 0;
}

/* End-interval filter: a scaled Chebyshev polynomial expressed in the
 * Chebyshev basis of [-1,1], degree chosen as the first one whose value at
 * the interval edge drops below thresh_ext (SRC/chebpoly.c:170-259). */
static void chext(polparams *pol, double aIn, double bIn) {
  int max_deg = pol->max_deg;
  double thresh = pol->thresh_ext;
  double *mu = pol->mu;
  double del = 0.1 * sqrt((bIn - aIn) * 0.5);
  double eps = 1e-13;
  double a, b, x, e, c, sigma, sigma1, sigma_new, g0, g1, gnew, s1, s2, s3, gam;
  int mbest = 0;
  if (pol->deg > 0) {
    thresh = -1.0;
    max_deg = pol->deg;
  }
  int m1 = max_deg + 1;
  if (bIn >= (1.0 - eps)) { /* right interval */
    x = aIn;
    b = x - del;
    a = -1.0;
    gam = 1.0;
  } else { /* left interval */
    x = bIn;
    a = x + del;
    b = 1.0;
    gam = -1.0;
  }
  e = (b - a) * 0.5;
  c = (b + a) * 0.5;
  sigma1 = e / (gam - c);
  sigma = sigma1;
  double *t0 = calloc(3 * (size_t)m1, sizeof(double));
  double *t1 = t0 + m1, *tnew = t1 + m1;
  t0[0] = 1.0;
  t1[0] = -c * (sigma1 / e);
  t1[1] = (sigma1 / e);
  g0 = 1.0;
  g1 = (x - c) * sigma1 / e;
  if (g1 < thresh) {
    mbest = 1;
  } else {
    for (int k = 2; k <= max_deg; k++) {
      sigma_new = 1.0 / (2.0 / sigma1 - sigma);
      s1 = sigma_new / e;
      s2 = 2 * s1 * c;
      s3 = sigma * sigma_new;
      for (int j = 2; j <= k; j++) tnew[j - 1] = s1 * (t1[j] + t1[j - 2]);
      tnew[1] = tnew[1] + s1 * t1[0];
      tnew[0] = s1 * t1[1];
      tnew[k] = s1 * t1[k - 1];
      for (int j = 0; j <= k; j++) tnew[j] = tnew[j] - s2 * t1[j] - s3 * t0[j];
      for (int j = 0; j <= k; j++) {
        t0[j] = t1[j];
        t1[j] = tnew[j];
      }
      gnew = 2 * (x - c) * s1 * g1 - s3 * g0;
      g0 = g1;
      g1 = gnew;
      mbest = k;
      if (g1 < thresh) break;
      sigma = sigma_new;
    }
  }
  memcpy(mu, t1, (mbest + 1) * sizeof(double));
  pol->deg = mbest;
  pol->bar = g1;
  pol->gam = gam;
  free(t0);
}

/* Newton / bisection balancing of the damped delta expansion so that
 * p(a) == p(b) (SRC/chebpoly.c:288-336).  Returns 1 if no root is bracketed. */
static int rootchb(int m, const double *v, const double *jac, double tha, double thb,
                   double *mu, double *thcOut) {
  const int MaxIterBalan = 30;
  double tolBal = fabs(tha - thb) * 1.e-13;
  double thc = 0.5 * (tha + thb);
  double fb = dif_eval(v, m, thb, jac);
  double fa = dif_eval(v, m, tha, jac);
  if ((fa > 0) || (fb < 0)) return 1;
  for (int it = 0; it <= MaxIterBalan; it++) {
    double fval = dif_eval(v, m, thc, jac);
    double d = 0.0;
    for (int j = 1; j <= m; j++) d += jac[j] * j * sin(j * (thc)) * v[j];
    double thN = thc + fval / d;
    if ((fabs(fval) < tolBal) || fabs(thc - thN) < DBL_EPSILON * fabs(thc)) break;
    if (fval > 0) {
      if ((thN < thb) || (thN > tha)) thN = 0.5 * (thc + tha);
      thb = thc;
      thc = thN;
    } else {
      if ((thN < thb) || (thN > tha)) thN = 0.5 * (thc + thb);
      tha = thc;
      thc = thN;
    }
  }
  for (int j = 0; j <= m; j++) mu[j] = cos(j * (thc)) * jac[j];
  *thcOut = thc;
  return 0;
}

/* SRC/chebpoly.c:363-490 */
int ora_find_pol(const double *intv, polparams *pol) {
  int max_deg = pol->max_deg, min_deg, damping = pol->damping;
  double tha = 0.0, thb = 0.0, thc = 0.0, t = 0.0, itv[2] = {0, 0}, vals[2] = {0, 0};
  double gam = 0.0, thresh;
  int m, nitv;
  if (check_intv(intv) < 0) return -1;
  int cap = max_deg;
  if (pol->deg > cap) cap = pol->deg;
  double *mu = calloc(cap + 2, sizeof(double));
  pol->mu = mu;
  double *v = calloc(cap + 2, sizeof(double));
  double *jac = calloc(cap + 2, sizeof(double));
  double IntTol = pol->intvtol;
  double aa = evsl_max(intv[0], intv[2]), bb = evsl_min(intv[1], intv[3]);
  double lmin = intv[2], lmax = intv[3];
  double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  pol->cc = cc;
  pol->dd = dd;
  aa = (aa - cc) / dd;
  bb = (bb - cc) / dd;
  aa = evsl_max(aa, -1.0);
  bb = evsl_min(bb, 1.0);
  thb = acos(bb);
  tha = acos(aa);
  if (aa - IntTol <= -1.0) {
    thc = tha;
    nitv = 1;
    aa = -1.0;
    gam = -1.0;
    pol->type = 1;
  } else if (bb + IntTol >= 1.0) {
    thc = thb;
    nitv = 1;
    bb = 1;
    gam = 1;
    pol->type = 2;
  } else {
    itv[0] = aa;
    itv[1] = bb;
    nitv = 2;
    pol->type = 0;
  }
  (void)thc;
  if (nitv == 1) {
    chext(pol, aa, bb);
  } else {
    min_deg = 2 + 0.5 / (bb - aa);
    min_deg = evsl_max(min_deg, 2);
    min_deg = evsl_min(min_deg, (max_deg + 1) / 2);
    thresh = pol->thresh_int;
    if (pol->deg > 0) {
      thresh = -1;
      max_deg = pol->deg;
    }
    for (int j = 0; j < min_deg; j++) v[j] = cos(j * thb) - cos(j * tha);
    for (m = min_deg; m < max_deg; m++) {
      ora_dampcf(m, damping, jac);
      v[m] = cos(m * thb) - cos(m * tha);
      if (rootchb(m, v, jac, tha, thb, mu, &thc)) continue;
      gam = cos(thc);
      ora_chebxPltd(m, mu, 1, &gam, &t);
      ora_chebxPltd(m, mu, nitv, itv, vals);
      if (vals[0] <= t * thresh && vals[1] <= t * thresh) {
        m++;
        break;
      }
    }
    int mbest = m - 1;
    for (int j = 0; j <= m; j++) mu[j] /= t;
    pol->bar = evsl_min(vals[0], vals[1]) / t;
    pol->gam = gam;
    pol->deg = mbest;
  }
  free(v);
  free(jac);
  return 0;
}

void ora_free_pol(polparams *pol) {
  free(pol->mu);
  pol->mu = NULL;
}

/* y = p((A - cc)/dd) v by the three-term recurrence; w is 3n workspace
 * (SRC/chebpoly.c:519-598, CPU branch :575-582) */
int ora_ChebAv(const polparams *pol, const double *v, double *y, double *w) {
  const int n = g_n, m = pol->deg;
  const double *mu = pol->mu, cc = pol->cc;
  const double t1 = 1.0 / pol->dd, t2 = 2.0 / pol->dd;
  double *vk = w, *vkp1 = w + n, *vkm1 = vkp1 + n;
  memcpy(vk, v, n * sizeof(double));
  for (int i = 0; i < n; i++) y[i] = mu[0] * v[i];
  for (int k = 1; k <= m; k++) {
    double t = (k == 1 ? t1 : t2), s = mu[k];
    ora_matvec(vk, vkp1);
    for (int i = 0; i < n; i++) {
      double r = t * (vkp1[i] - cc * vk[i]);
      if (k > 1) r -= vkm1[i];
      vkp1[i] = r;
      y[i] += s * r;
    }
    double *tmp = vkm1;
    vkm1 = vk;
    vk = vkp1;
    vkp1 = tmp;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* KPM DOS and slicing (SRC/spslice.c)                                 */
/* ------------------------------------------------------------------ */
/* SRC/spslice.c:29-148 (standard problem) */
int ora_kpmdos(int Mdeg, int damping, int nvec, const double *intv, double *mu,
               double *ecnt) {
  const int n = g_n;
  if (check_intv(intv) < 0) return -1;
  double *vkp1 = malloc(n * sizeof(double)), *v = malloc(n * sizeof(double));
  double *vkm1 = malloc(n * sizeof(double)), *vk = malloc(n * sizeof(double));
  double *jac = malloc((Mdeg + 1) * sizeof(double));
  double aa = evsl_max(intv[0], intv[2]), bb = evsl_min(intv[1], intv[3]);
  double ctr = (intv[3] + intv[2]) / 2.0, wid = (intv[3] - intv[2]) / 2.0;
  double t = evsl_max(-1.0 + DBL_EPSILON, (aa - ctr) / wid);
  double beta1 = acos(t);
  t = evsl_min(1.0 - DBL_EPSILON, (bb - ctr) / wid);
  double beta2 = acos(t);
  ora_dampcf(Mdeg, damping, jac);
  jac[0] = 1.0;
  memset(mu, 0, (Mdeg + 1) * sizeof(double));
  double tcnt = 0.0;
  for (int m = 0; m < nvec; m++) {
    ora_rand_double(n, v);
    t = 1.0 / nrm2(n, v);
    scal(n, t, v);
    memcpy(vk, v, n * sizeof(double));
    memset(vkm1, 0, n * sizeof(double));
    mu[0] += jac[0];
    tcnt -= jac[0] * (beta2 - beta1);
    for (int k = 0; k < Mdeg; k++) {
      ora_matvec(vk, vkp1);
      double sc = (k == 0 ? 1.0 : 2.0) / wid;
      for (int i = 0; i < n; i++) vkp1[i] = sc * (vkp1[i] - ctr * vk[i]) - vkm1[i];
      double *tmp = vkm1;
      vkm1 = vk;
      vk = vkp1;
      vkp1 = tmp;
      int k1 = k + 1;
      t = 2 * jac[k1] * dot(n, vk, v);
      mu[k1] += t;
      tcnt -= t * (sin(k1 * beta2) - sin(k1 * beta1)) / k1;
    }
  }
  t = 1.0 / (((double)nvec) * EVSL_PI);
  scal(Mdeg + 1, t, mu);
  tcnt *= t * ((double)n);
  *ecnt = tcnt;
  free(vkp1);
  free(v);
  free(vkm1);
  free(vk);
  free(jac);
  return 0;
}

/* yi[j] = integral from xi[0] to xi[j] of the KPM density, in closed form
 * (SRC/spslice.c:155-181) */
void ora_intChx(int Mdeg, const double *mu, int npts, const double *xi, double *yi) {
  if (npts <= 0) return;
  double *thetas = malloc(npts * sizeof(double));
  for (int j = 0; j < npts; j++) thetas[j] = acos(xi[j]);
  double theta0 = thetas[0];
  for (int j = 0; j < npts; j++) yi[j] = mu[0] * (theta0 - thetas[j]);
  for (int k = 1; k < Mdeg + 1; k++) {
    double val0 = sin(k * theta0) / k;
    for (int j = 0; j < npts; j++) yi[j] += mu[k] * (val0 - sin(k * thetas[j]) / k);
  }
  free(thetas);
}

/* equal-count slices on a linspace grid (SRC/spslice.c:205-297) */
int ora_spslicer(double *sli, const double *mu, int Mdeg, const double *intv,
                 int n_int, int npts) {
  int ls, ii, err = 0;
  if (check_intv(intv) < 0) return -1;
  double aa = evsl_max(intv[0], intv[2]), bb = evsl_min(intv[1], intv[3]);
  memset(sli, 0, (n_int + 1) * sizeof(double));
  if (n_int == 1) {
    sli[0] = intv[0];
    sli[1] = intv[1];
    return 0;
  }
  double ctr = (intv[3] + intv[2]) / 2, wid = (intv[3] - intv[2]) / 2;
  double aL = (aa - ctr) / wid, bL = (bb - ctr) / wid;
  aL = evsl_max(aL, -1.0);
  bL = evsl_min(bL, 1.0);
  npts = evsl_max(npts, 2 * n_int + 1);
  double *xi = malloc(npts * sizeof(double)), *yi = malloc(npts * sizeof(double));
  ora_linspace(aL, bL, npts, xi);
  ora_intChx(Mdeg, mu, npts, xi, yi);
  double target = yi[npts - 1] / (double)n_int;
  ls = 0;
  ii = 0;
  sli[ls] = intv[0];
  while (++ls < n_int) {
    while (ii < npts && yi[ii] < target) ii++;
    if (ii == npts) break;
    /* the nearer of the two grid points; the reference reads yi[-1] when
     * target <= yi[0] (a density whose integral is not positive), here the
     * first grid point has no lower neighbour and is kept */
    if (ORA_PLANT != 1 && ii > 0 && (target - yi[ii - 1]) < (yi[ii] - target)) ii--;
    sli[ls] = ctr + wid * xi[ii];
    if (ORA_PLANT == 2)
      target = yi[npts - 1] * (ls + 1) / (double)n_int;
    else
      target = yi[ii] + (yi[npts - 1] - yi[ii]) / (n_int - ls);
  }
  sli[n_int] = intv[1];
  if (ls != n_int) err = 1;
  for (ii = 1; ii <= n_int; ii++) {
    if (sli[ii] < sli[ii - 1]) {
      err += 2;
      break;
    }
  }
  free(xi);
  free(yi);
  return err;
}

/* ------------------------------------------------------------------ */
/* dense linear algebra                                                */
/* ------------------------------------------------------------------ */
/* classical Gram-Schmidt with the DGKS re-orthogonalization test, two GEMVs
 * per pass (SRC/misc_la.c:221-261, USE_DGEMV branch) */
void ora_CGS_DGKS(int n, int k, int i_max, const double *Q, double *v, double *nrmv,
                  double *w) {
  const double eta = 1.0 / sqrt(2.0);
  double old_nrm = nrm2(n, v), new_nrm = 0.0;
  for (int it = 0; it < i_max; it++) {
    for (int j = 0; j < k; j++) w[j] = dot(n, Q + (size_t)j * n, v);
    for (int j = 0; j < k; j++) axpy(n, -w[j], Q + (size_t)j * n, v);
    new_nrm = nrm2(n, v);
    if (new_nrm > eta * old_nrm) break;
    old_nrm = new_nrm;
  }
  if (nrmv) *nrmv = new_nrm;
}

/* Implicit QL with a Wilkinson-type shift on the symmetric tridiagonal
 * (d, e), e[n-1] used as scratch.  If z != NULL it holds an n x n column-major
 * matrix that is multiplied on the right by every rotation (pass the identity
 * for the eigenvectors of T, or a Householder basis Q for those of Q T Q^T).
 * Eigenvalues are left ascending with the columns of z permuted to match.
 * Replaces LAPACK dstev/dsteqr used by SymmTridEig (SRC/misc_la.c:33-73). */
static int tridiag_ql(int n, double *d, double *e, double *z, int ldz) {
  if (n <= 0) return 0;
  e[n - 1] = 0.0;
  for (int l = 0; l < n; l++) {
    int iter = 0, m;
    for (;;) {
      for (m = l; m < n - 1; m++) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) return -1;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      int early = 0;
      for (i = m - 1; i >= l; i--) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          early = 1;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (z) {
          double *zi = z + (size_t)i * ldz, *zi1 = z + (size_t)(i + 1) * ldz;
          for (int k = 0; k < n; k++) {
            double tz = zi1[k];
            zi1[k] = s * zi[k] + c * tz;
            zi[k] = c * zi[k] - s * tz;
          }
        }
      }
      if (early) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  /* selection sort, ascending */
  for (int i = 0; i < n - 1; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      double t = d[i];
      d[i] = d[k];
      d[k] = t;
      if (z) {
        double *zi = z + (size_t)i * ldz, *zk = z + (size_t)k * ldz;
        for (int r = 0; r < n; r++) {
          t = zi[r];
          zi[r] = zk[r];
          zk[r] = t;
        }
      }
    }
  }
  return 0;
}

/* eigenvalues (and, if evec != NULL, eigenvectors as columns of an n x n
 * column-major matrix) of the tridiagonal with diagonal diag[0..n-1] and
 * sub-diagonal sdiag[0..n-2] (SRC/misc_la.c:33-73) */
int ora_tridiag_eig(int n, const double *diag, const double *sdiag, double *eval,
                    double *evec) {
  double *e = malloc((n + 1) * sizeof(double));
  memcpy(eval, diag, n * sizeof(double));
  if (n > 1) memcpy(e, sdiag, (n - 1) * sizeof(double));
  if (evec) {
    memset(evec, 0, (size_t)n * n * sizeof(double));
    for (int i = 0; i < n; i++) evec[(size_t)i * n + i] = 1.0;
  }
  int info = tridiag_ql(n, eval, e, evec, n);
  free(e);
  return info;
}

/* Dense symmetric eigensolver: Householder reduction to tridiagonal form with
 * the reflections accumulated, then implicit QL.  Replaces LAPACK dsyev as
 * called by SymEigenSolver (SRC/misc_la.c:169-213). */
int ora_sym_eig(int n, const double *A, int lda, double *Q, int ldq, double *lam) {
  double *W = malloc((size_t)n * n * sizeof(double));
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) W[i + (size_t)j * n] = A[i + (size_t)j * lda];
  double *Z = malloc((size_t)n * n * sizeof(double));
  memset(Z, 0, (size_t)n * n * sizeof(double));
  for (int i = 0; i < n; i++) Z[i + (size_t)i * n] = 1.0;
  double *v = malloc(n * sizeof(double)), *p = malloc(n * sizeof(double));
  for (int k = 0; k + 2 < n; k++) {
    int len = n - k - 1;
    double *x = W + (k + 1) + (size_t)k * n;
    double xn = 0.0;
    for (int i = 0; i < len; i++) xn += x[i] * x[i];
    xn = sqrt(xn);
    if (xn == 0.0) continue;
    double alpha = (x[0] > 0 ? -xn : xn);
    for (int i = 0; i < len; i++) v[i] = x[i];
    v[0] -= alpha;
    double vn = 0.0;
    for (int i = 0; i < len; i++) vn += v[i] * v[i];
    vn = sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = 0; i < len; i++) v[i] /= vn;
    /* p = W22 v, K = v^T p, w = 2 (p - K v) */
    double K = 0.0;
    for (int i = 0; i < len; i++) {
      double s = 0.0;
      for (int j = 0; j < len; j++) s += W[(k + 1 + i) + (size_t)(k + 1 + j) * n] * v[j];
      p[i] = s;
      K += v[i] * s;
    }
    for (int i = 0; i < len; i++) p[i] = 2.0 * (p[i] - K * v[i]);
    for (int j = 0; j < len; j++)
      for (int i = 0; i < len; i++)
        W[(k + 1 + i) + (size_t)(k + 1 + j) * n] -= v[i] * p[j] + p[i] * v[j];
    W[(k + 1) + (size_t)k * n] = alpha;
    W[k + (size_t)(k + 1) * n] = alpha;
    for (int i = 1; i < len; i++) {
      W[(k + 1 + i) + (size_t)k * n] = 0.0;
      W[k + (size_t)(k + 1 + i) * n] = 0.0;
    }
    /* Z(:, k+1:) <- Z(:, k+1:) (I - 2 v v^T) */
    for (int r = 0; r < n; r++) {
      double s = 0.0;
      for (int i = 0; i < len; i++) s += Z[r + (size_t)(k + 1 + i) * n] * v[i];
      s *= 2.0;
      for (int i = 0; i < len; i++) Z[r + (size_t)(k + 1 + i) * n] -= s * v[i];
    }
  }
  double *e = malloc((n + 1) * sizeof(double));
  for (int i = 0; i < n; i++) lam[i] = W[i + (size_t)i * n];
  for (int i = 0; i + 1 < n; i++) e[i] = W[(i + 1) + (size_t)i * n];
  int info = tridiag_ql(n, lam, e, Z, n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) Q[i + (size_t)j * ldq] = Z[i + (size_t)j * n];
  free(W);
  free(Z);
  free(v);
  free(p);
  free(e);
  return info;
}

/* ------------------------------------------------------------------ */
/* filtered non-restarted Lanczos (SRC/cheblanNr.c:57-434)             */
/* ------------------------------------------------------------------ */
int ora_ChebLanNr(const double *intv, int maxit, double tol, const double *vinit,
                  const polparams *pol, int *nevOut, double **lamo, double **Wo,
                  double **reso) {
  const int n = g_n;
  const size_t nl = n;
  const int Ntest = 30, cycle = 20;
  double tr0 = 0.0, tr1;
  int k, kdim = 0, nconv = 0, count = 0;
  maxit = evsl_min(n, maxit);
  const double bar = pol->bar, gamB = pol->gam;
  if (check_intv(intv) < 0) {
    *nevOut = 0;
    *lamo = NULL;
    *Wo = NULL;
    *reso = NULL;
    return 0;
  }
  const double aa = intv[0], bb = intv[1];
  if (gamB > 1.0 || gamB < -1.0) return -1;
  double *V = malloc(nl * (maxit + 1) * sizeof(double));
  double *dT = malloc(maxit * sizeof(double)), *eT = malloc(maxit * sizeof(double));
  double *EvalT = malloc(maxit * sizeof(double));
  double *wk = malloc(3 * nl * sizeof(double));
  double *vrand = malloc(nl * sizeof(double));
  /* filter the start vector, normalize (FILTER_VINIT, :147-168) */
  ora_ChebAv(pol, vinit, V, wk);
  scal(n, 1.0 / nrm2(n, V), V);
  double wn = 0.0, beta = 0.0;
  int nwn = 0;
  for (k = 0; k < maxit; k++) {
    double *zold = k > 0 ? V + (size_t)(k - 1) * nl : NULL;
    double *v = V + (size_t)k * nl, *vnew = v + nl;
    ora_ChebAv(pol, v, vnew, wk);
    if (zold) axpy(n, -beta, zold, vnew);
    double alpha = dot(n, v, vnew);
    dT[k] = alpha;
    wn += fabs(alpha);
    axpy(n, -alpha, v, vnew);
    ora_CGS_DGKS(n, k + 1, NGS_MAX, V, vnew, &beta, wk);
    wn += 2.0 * beta;
    nwn += 3;
    if (beta * nwn < ORTH_TOL * wn) { /* lucky breakdown (:225-256) */
      ora_rand_double(n, vrand);
      ora_ChebAv(pol, vrand, vnew, wk);
      ora_CGS_DGKS(n, k + 1, NGS_MAX, V, vnew, &beta, wk);
      scal(n, 1.0 / beta, vnew);
      beta = 0.0;
    } else {
      scal(n, 1.0 / beta, vnew);
    }
    eT[k] = beta;
    if ((k < Ntest || (k - Ntest) % cycle != 0) && k != maxit - 1) continue;
    /* restricted-trace convergence test (:269-309) */
    kdim = k + 1;
    if (ora_tridiag_eig(kdim, dT, eT, EvalT, NULL)) return -3;
    count = kdim;
    tr1 = 0;
    nconv = 0;
    for (int i = 0; i < count; i++) {
      if (EvalT[i] + EPSM >= bar) {
        tr1 += EvalT[i];
        nconv++;
      }
    }
    if ((fabs(tr1 - tr0) < 2e-12) || (fabs(tr1) + fabs(tr0) < 1e-10)) break;
    tr0 = tr1;
  }
  g_last_steps = (k < maxit ? k + 1 : maxit);
  /* Ritz pairs and acceptance (:316-399) */
  double *EvecT = malloc((size_t)kdim * kdim * sizeof(double));
  if (ora_tridiag_eig(kdim, dT, eT, EvalT, EvecT)) return -3;
  int ncand = 0;
  for (int i = 0; i < count; i++)
    if (!(EvalT[i] < bar)) ncand++;
  double *Rvec = malloc(nl * (ncand > 0 ? ncand : 1) * sizeof(double));
  double *Lam = malloc((maxit > 0 ? maxit : 1) * sizeof(double));
  double *res = malloc((maxit > 0 ? maxit : 1) * sizeof(double));
  int nev = 0;
  for (int i = 0; i < count; i++) {
    double flami = EvalT[i];
    if (flami < bar) continue;
    double *y = EvecT + (size_t)i * kdim;
    double *u = Rvec + (size_t)nev * nl;
    gemv_n(n, kdim, V, nl, y, u);
    double t = nrm2(n, u);
    if (t == 0.0) return 2;
    scal(n, 1.0 / t, u);
    ora_matvec(u, wk);
    t = dot(n, wk, u);
    if (t < aa - EPSM || t > bb + EPSM) continue;
    axpy(n, -t, u, wk);
    double res0 = nrm2(n, wk);
    if (res0 < tol) {
      Lam[nev] = t;
      res[nev] = res0;
      nev++;
    }
  }
  *nevOut = nev;
  *lamo = Lam;
  *Wo = Rvec;
  *reso = res;
  free(V);
  free(dT);
  free(eT);
  free(EvalT);
  free(EvecT);
  free(wk);
  free(vrand);
  (void)nconv;
  return 0;
}

int ora_last_lanczos_steps(void) { return g_last_steps; }

/* ------------------------------------------------------------------ */
/* thick-restart filtered Lanczos (SRC/cheblanTr.c:59-707)             */
/* ------------------------------------------------------------------ */
int ora_ChebLanTr(int lanm, int nev, const double *intv, int maxit, double tol,
                  const double *vinit, const polparams *pol, int *nev2, double **vals,
                  double **W, double **resW) {
  const int n = g_n;
  const size_t nl = n;
  double tr, last_tr;
  lanm = evsl_min(lanm, n);
  int lanm1 = lanm + 1;
  size_t L1 = lanm1;
  if (lanm == n) maxit = evsl_min(maxit, n);
  const double nevInc = 0.2;
  int Ntest;
  const int cycle = 50;
  int i, ll, jl, last_jl;
  if (check_intv(intv) < 0) {
    *nev2 = 0;
    *vals = NULL;
    *W = NULL;
    *resW = NULL;
    return 0;
  }
  const double aa = intv[0], bb = intv[1];
  const double gamB = pol->gam, bar = pol->bar;
  if (gamB > 1.0 || gamB < -1.0) return -1;
  const int min_inner_step = 5;
  int it = 0;
  double *V = malloc(nl * L1 * sizeof(double));
  double *T = calloc(L1 * L1, sizeof(double));
  double *Y = malloc(nl * nev * sizeof(double));
  double *Lam = malloc(nev * sizeof(double)), *res = malloc(nev * sizeof(double));
  int lock = 0, trlen = 0, prtrlen = -1;
  double *Rval = malloc(lanm * sizeof(double)), *resi = malloc(lanm * sizeof(double));
  double *Rvec = malloc(nl * lanm * sizeof(double));
  double *EvecT = malloc(L1 * L1 * sizeof(double));
  double *s = malloc(lanm * sizeof(double));
  double *work = malloc(3 * nl * sizeof(double));
  double *vrand = malloc(nl * sizeof(double));
  double *hw = malloc(L1 * sizeof(double) + nl * sizeof(double));
  ora_ChebAv(pol, vinit, V, work);
  scal(n, 1.0 / nrm2(n, V), V);
  int k = 0;
  double beta = 0.0;
  while (it < maxit) {
    double wn = 0.0;
    int nwn = 0;
    beta = 0.0;
    k = trlen > 0 ? trlen + 1 : 0;
    if (k + min_inner_step > lanm1) return 3; /* reference exits (:217-222) */
    if (trlen > 0) { /* thick-restart special step (:224-319) */
      int k1 = k - 1;
      double *v = V + (size_t)k1 * nl, *vnew = v + nl;
      ora_ChebAv(pol, v, vnew, work);
      if (lock > 0) ora_CGS_DGKS(n, lock, NGS_MAX, Y, vnew, NULL, hw);
      for (i = 0; i < trlen; i++) {
        T[i * L1 + i] = Rval[i];
        wn += fabs(Rval[i]);
      }
      s[k1] = dot(n, v, vnew);
      for (int j = 0; j < k; j++) axpy(n, -s[j], V + (size_t)j * nl, vnew);
      for (i = 0; i < k1; i++) {
        T[trlen * L1 + i] = s[i];
        T[i * L1 + trlen] = s[i];
        wn += 2.0 * fabs(s[i]);
      }
      T[trlen * L1 + trlen] = s[k1];
      wn += fabs(s[k1]);
      beta = nrm2(n, vnew);
      wn += 2.0 * beta;
      nwn += 3 * k;
      if (beta * nwn < ORTH_TOL * wn) {
        ora_rand_double(n, vrand);
        ora_ChebAv(pol, vrand, vnew, work);
        ora_CGS_DGKS(n, lock, NGS_MAX, Y, vnew, NULL, hw);
        ora_CGS_DGKS(n, k, NGS_MAX, V, vnew, &beta, hw);
        scal(n, 1.0 / beta, vnew);
        beta = 0.0;
      } else {
        scal(n, 1.0 / beta, vnew);
      }
      T[k1 * L1 + k] = beta;
      T[k * L1 + k1] = beta;
    }
    Ntest = evsl_max(20, nev - lock + 10);
    last_jl = 0;
    last_tr = 0.0;
    double *zold = k > 0 ? V + (size_t)(k - 1) * nl : NULL;
    while (k < lanm && it < maxit) { /* Lanczos inner loop (:331-503) */
      k++;
      double *v = V + (size_t)(k - 1) * nl, *vnew = v + nl;
      ora_ChebAv(pol, v, vnew, work);
      it++;
      if (lock > 0) ora_CGS_DGKS(n, lock, NGS_MAX, Y, vnew, NULL, hw);
      if (zold) axpy(n, -beta, zold, vnew);
      double alpha = dot(n, v, vnew);
      T[(k - 1) * L1 + (k - 1)] = alpha;
      wn += fabs(alpha);
      axpy(n, -alpha, v, vnew);
      ora_CGS_DGKS(n, k, NGS_MAX, V, vnew, &beta, hw);
      wn += 2.0 * beta;
      nwn += 3;
      zold = v;
      if (beta * nwn < ORTH_TOL * wn) {
        ora_rand_double(n, vrand);
        ora_ChebAv(pol, vrand, vnew, work);
        ora_CGS_DGKS(n, lock, NGS_MAX, Y, vnew, NULL, hw);
        ora_CGS_DGKS(n, k, NGS_MAX, V, vnew, &beta, hw);
        scal(n, 1.0 / beta, vnew);
        beta = 0.0;
      } else {
        scal(n, 1.0 / beta, vnew);
      }
      T[k * L1 + (k - 1)] = beta;
      T[(k - 1) * L1 + k] = beta;
      int k1 = k - trlen - Ntest;
      if (((k1 >= 0) && (k1 % cycle == 0)) || (k == lanm) || it == maxit) {
        if (ora_sym_eig(k, T, lanm1, EvecT, lanm1, Rval)) return -3;
        if (k == lanm || it == maxit) break;
        jl = 0;
        tr = 0.0;
        for (i = 0; i < k; i++) {
          if (Rval[i] + EPSM >= bar) {
            jl++;
            tr += Rval[i];
          }
        }
        if (fabs(tr - last_tr) <= 2e-12 && jl == last_jl) break;
        last_tr = tr;
        last_jl = jl;
      }
    }
    /* pass 1: Ritz values of p(A) above bar, moved to the front (:512-527) */
    jl = 0;
    for (i = 0; i < k; i++) {
      if (Rval[i] + EPSM >= bar) {
        if (i != jl) {
          memcpy(EvecT + jl * L1, EvecT + i * L1, k * sizeof(double));
          Rval[jl] = Rval[i];
        }
        resi[jl] = fabs(beta * EvecT[i * L1 + (k - 1)]);
        jl++;
      }
    }
    /* Rvec(:,1:jl) = V(:,1:k) * EvecT(:,1:jl) (:535) */
    for (int j = 0; j < jl; j++) gemv_n(n, k, V, nl, EvecT + (size_t)j * L1, Rvec + j * nl);
    /* pass 2: Ritz values of A in [a,b]; lock or keep for restart (:541-632) */
    ll = 0;
    prtrlen = trlen;
    trlen = 0;
    for (i = 0; i < jl; i++) {
      double *y = Rvec + (size_t)i * nl, *w = work;
      double t = nrm2(n, y);
      if (t == 0.0) return 2;
      scal(n, 1.0 / t, y);
      ora_matvec(y, w);
      double t3 = dot(n, y, w);
      if (t3 >= aa - EPSM && t3 <= bb + EPSM) {
        ll++;
        axpy(n, -t3, y, w);
        double res0 = nrm2(n, w);
        if (res0 < tol) {
          if (lock >= nev) {
            nev += 1 + (int)(nev * nevInc);
            Y = realloc(Y, nl * nev * sizeof(double));
            Lam = realloc(Lam, nev * sizeof(double));
            res = realloc(res, nev * sizeof(double));
          }
          memcpy(Y + (size_t)lock * nl, y, nl * sizeof(double));
          Lam[lock] = t3;
          res[lock] = res0;
          lock++;
        } else {
          Rval[trlen] = Rval[i];
          memcpy(V + (size_t)trlen * nl, y, nl * sizeof(double));
          s[trlen] = beta * EvecT[i * L1 + (k - 1)];
          trlen++;
        }
      }
    }
    if ((prtrlen == 0) && (ll == 0)) break;
    if (it == maxit) break;
    memset(T, 0, L1 * L1 * sizeof(double));
    memcpy(V + (size_t)trlen * nl, V + (size_t)k * nl, nl * sizeof(double));
  }
  *nev2 = lock;
  *vals = Lam;
  *W = Y;
  *resW = res;
  free(V);
  free(T);
  free(Rval);
  free(resi);
  free(EvecT);
  free(Rvec);
  free(s);
  free(work);
  free(vrand);
  free(hw);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Lanczos spectrum bounds (SRC/lanbounds.c:21-144)                    */
/* ------------------------------------------------------------------ */
int ora_LanBounds(int msteps, const double *v, double *lmin, double *lmax) {
  const int n = g_n;
  const size_t nl = n;
  msteps = evsl_min(n, msteps);
  double *alp = malloc(msteps * sizeof(double)), *bet = malloc(msteps * sizeof(double));
  double *V = malloc((msteps + 1) * nl * sizeof(double));
  double t = 1.0 / nrm2(n, v);
  memcpy(V, v, nl * sizeof(double));
  scal(n, t, V);
  double wn = 0.0;
  for (int j = 0; j < msteps; j++) {
    double *vj = V + j * nl, *vn = V + (j + 1) * nl;
    ora_matvec(vj, vn);
    if (j) axpy(n, -bet[j - 1], V + (j - 1) * nl, vn);
    alp[j] = dot(n, vn, vj);
    wn += alp[j] * alp[j];
    axpy(n, -alp[j], vj, vn);
    for (int i = 0; i <= j; i++) { /* modified Gram-Schmidt (:78-84) */
      t = dot(n, vn, V + i * nl);
      axpy(n, -t, V + i * nl, vn);
    }
    bet[j] = dot(n, vn, vn);
    if (bet[j] * (j + 1) < ORTH_TOL * wn) {
      msteps = j + 1;
      break;
    }
    wn += 2.0 * bet[j];
    bet[j] = sqrt(bet[j]);
    scal(n, 1.0 / bet[j], vn);
  }
  double bottomBeta = bet[msteps - 1];
  size_t ml = msteps;
  double *S = malloc(ml * ml * sizeof(double)), *ritz = malloc(msteps * sizeof(double));
  int info = ora_tridiag_eig(msteps, alp, bet, ritz, S);
  if (ORA_PLANT == 4) {
    *lmin = ritz[0];
    *lmax = ritz[msteps - 1];
  } else if (ORA_PLANT == 5) {
    *lmin = ritz[0] - fabs(bottomBeta * S[0]);
    *lmax = ritz[msteps - 1] + fabs(bottomBeta * S[ml * ml - ml]);
  } else {
    *lmin = ritz[0] - fabs(bottomBeta * S[msteps - 1]);
    *lmax = ritz[msteps - 1] + fabs(bottomBeta * S[ml * ml - 1]);
  }
  free(alp);
  free(bet);
  free(V);
  free(S);
  free(ritz);
  return info ? -3 : 0;
}

/* ------------------------------------------------------------------ */
/* Thick-restart Lanczos spectrum bounds (SRC/lanTrbounds.c:35-421):   */
/* Lanczos on A with full reorthogonalization, restarted with the two  */
/* extreme Ritz pairs; simple (|beta s|) or Kato-Temple bounds.        */
/* ------------------------------------------------------------------ */
int ora_LanTrbounds(int lanm, int maxit, double tol, const double *vinit, int bndtype,
                    double *lammin, double *lammax) {
  const int n = g_n;
  const size_t nl = n;
  double lmin = 0.0, lmax = 0.0, t1, t2;
  lanm = evsl_min(lanm, n);
  const int lanm1 = lanm + 1;
  const size_t L1 = lanm1;
  if (lanm == n) maxit = evsl_min(maxit, n);
  const int min_inner_step = 5;
  int it = 0, trlen = 0, i;
  double *V = malloc(nl * L1 * sizeof(double));
  double *T = calloc(L1 * L1, sizeof(double));
  double *Rval = malloc(lanm * sizeof(double));
  double *Rvec = malloc(2 * nl * sizeof(double));
  double *EvecT = malloc(L1 * L1 * sizeof(double));
  double *hw = malloc(L1 * sizeof(double) + nl * sizeof(double));
  double s[3];
  memcpy(V, vinit, nl * sizeof(double));
  scal(n, 1.0 / nrm2(n, V), V);
  int ret = 0;
  while (it < maxit) {
    double wn = 0.0, beta = 0.0;
    int nwn = 0;
    int k = trlen > 0 ? trlen + 1 : 0;
    if (k + min_inner_step > lanm1) { /* the reference exits (:139-143) */
      ret = 3;
      break;
    }
    if (trlen > 0) { /* thick-restart special step (:145-215) */
      int k1 = k - 1;
      double *v = V + (size_t)k1 * nl, *vnew = v + nl;
      ora_matvec(v, vnew);
      for (i = 0; i < trlen; i++) {
        T[i * L1 + i] = Rval[i];
        wn += fabs(Rval[i]);
      }
      s[k1] = dot(n, v, vnew);
      for (int j = 0; j < k; j++) axpy(n, -s[j], V + (size_t)j * nl, vnew);
      for (i = 0; i < k1; i++) {
        T[trlen * L1 + i] = s[i];
        T[i * L1 + trlen] = s[i];
        wn += 2.0 * fabs(s[i]);
      }
      T[trlen * L1 + trlen] = s[k1];
      wn += fabs(s[k1]);
      beta = nrm2(n, vnew);
      wn += 2.0 * beta;
      nwn += 3 * k;
      if (beta * nwn < ORTH_TOL * wn) {
        ora_rand_double(n, vnew);
        ora_CGS_DGKS(n, k, NGS_MAX, V, vnew, &beta, hw);
        scal(n, 1.0 / beta, vnew);
        beta = 0.0;
      } else {
        scal(n, 1.0 / beta, vnew);
      }
      T[k1 * L1 + k] = beta;
      T[k * L1 + k1] = beta;
    }
    double *zold = k > 0 ? V + (size_t)(k - 1) * nl : NULL;
    while (k < lanm && it < maxit) { /* Lanczos inner loop (:222-300) */
      k++;
      double *v = V + (size_t)(k - 1) * nl, *vnew = v + nl;
      ora_matvec(v, vnew);
      it++;
      if (zold) axpy(n, -beta, zold, vnew);
      double alpha = dot(n, v, vnew);
      T[(k - 1) * L1 + (k - 1)] = alpha;
      wn += fabs(alpha);
      axpy(n, -alpha, v, vnew);
      ora_CGS_DGKS(n, k, NGS_MAX, V, vnew, &beta, hw);
      wn += 2.0 * beta;
      nwn += 3;
      zold = v;
      if (beta * nwn < ORTH_TOL * wn) {
        ora_rand_double(n, vnew);
        ora_CGS_DGKS(n, k, NGS_MAX, V, vnew, &beta, hw);
        scal(n, 1.0 / beta, vnew);
        beta = 0.0;
      } else {
        scal(n, 1.0 / beta, vnew);
      }
      T[k * L1 + (k - 1)] = beta;
      T[(k - 1) * L1 + k] = beta;
    }
    if (ora_sym_eig(k, T, lanm1, EvecT, lanm1, Rval)) {
      ret = -3;
      break;
    }
    /* spikes: beta times the last components of the extreme eigenvectors */
    s[0] = beta * EvecT[k - 1];
    s[1] = beta * EvecT[(size_t)(k - 1) * L1 + (k - 1)];
    if (bndtype <= 1) {
      t1 = fabs(s[0]);
      t2 = fabs(s[1]);
    } else { /* Kato-Temple */
      t1 = 2.0 * s[0] * s[0] / (Rval[1] - Rval[0]);
      t2 = 2.0 * s[1] * s[1] / (Rval[k - 1] - Rval[k - 2]);
    }
    lmin = Rval[0] - t1;
    lmax = Rval[k - 1] + t2;
    /* the two extreme Ritz vectors become the restart set (:333-357) */
    gemv_n(n, k, V, nl, EvecT, Rvec);
    gemv_n(n, k, V, nl, EvecT + (size_t)(k - 1) * L1, Rvec + nl);
    trlen = 2;
    memcpy(V, Rvec, 2 * nl * sizeof(double));
    Rval[1] = Rval[k - 1];
    if (t1 + t2 < tol * (fabs(lmin) + fabs(lmax))) break;
    memset(T, 0, L1 * L1 * sizeof(double));
    memcpy(V + (size_t)trlen * nl, V + (size_t)k * nl, nl * sizeof(double));
  }
  *lammin = lmin;
  *lammax = lmax;
  free(V);
  free(T);
  free(Rval);
  free(Rvec);
  free(EvecT);
  free(hw);
  return ret;
}

/* ------------------------------------------------------------------ */
/* Filtered subspace iteration (SRC/chebsi.c:39-336) and the block     */
/* orthonormalisation it uses (orth, SRC/misc_la.c:317-342).           */
/* ------------------------------------------------------------------ */
static void ora_orth(const double *V, int n, int k, double *Vo, double *hw) {
  const size_t nl = n;
  memcpy(Vo, V, nl * k * sizeof(double));
  double nrmv = nrm2(n, Vo);
  scal(n, 1.0 / nrmv, Vo);
  for (int i = 1; i < k; i++) {
    ora_CGS_DGKS(n, i, NGS_MAX, Vo, Vo + (size_t)i * nl, &nrmv, hw);
    scal(n, 1.0 / nrmv, Vo + (size_t)i * nl);
  }
}

/* T (m x p, ld m) = A(:, 0:m)^T B(:, 0:p), columns of length n */
static void gemm_tn_(int n, int m, int p, const double *A, const double *B, double *T) {
  for (int j = 0; j < p; j++)
    for (int i = 0; i < m; i++) T[(size_t)j * m + i] = dot(n, A + (size_t)i * n, B + (size_t)j * n);
}

int ora_ChebSI(int nev, const double *intv, int maxit, double tol, const double *vinit,
               const polparams *pol, int *nevo, double **lamo, double **Yo, double **reso) {
  const int n = g_n;
  const size_t nl = n;
  const double tolP = tol * 0.01;
  *nevo = 0;
  *lamo = *Yo = *reso = NULL;
  if (check_intv(intv) < 0) return 0;
  const double aa = intv[0], bb = intv[1];
  const int cvcheck = 5, nbuf = 2;
  if (pol->gam > 1.0 || pol->gam < -1.0) return -1;
  double *T = malloc((size_t)nev * nev * sizeof(double));
  double *evecT = malloc((size_t)nev * nev * sizeof(double));
  double *evalT = malloc(nev * sizeof(double));
  double *V = malloc(nl * nev * sizeof(double)), *PV = malloc(nl * nev * sizeof(double));
  double *R = malloc(nl * nev * sizeof(double)), *buf = malloc(nl * nev * sizeof(double));
  double *Lam = malloc(nev * sizeof(double)), *res = malloc(nev * sizeof(double));
  int *act_idx = malloc(nev * sizeof(int)), *lock_idx = malloc(nev * sizeof(int));
  double *work = malloc(3 * nl * sizeof(double));
  double *hw = malloc((nev + 1) * sizeof(double) + nl * sizeof(double));
  int nlock = 0, nact = nev, nlock_ab = 0, it = 0, find_more = 1;
  ora_orth(vinit, n, nev, V, hw);
  for (int j = 0; j < nev; j++) ora_ChebAv(pol, V + (size_t)j * nl, PV + (size_t)j * nl, work);
  while (it < maxit && find_more) {
    const size_t nnlock = nl * nlock, nnact = nl * nact;
    memcpy(V + nnlock, PV + nnlock, nnact * sizeof(double));
    if (nlock > 0) { /* V_act -= V_lock (V_lock^T V_act) */
      gemm_tn_(n, nlock, nact, V, V + nnlock, T);
      for (int j = 0; j < nact; j++)
        for (int i = 0; i < nlock; i++)
          axpy(n, -T[(size_t)j * nlock + i], V + (size_t)i * nl, V + nnlock + (size_t)j * nl);
    }
    ora_orth(V + nnlock, n, nact, buf, hw);
    memcpy(V + nnlock, buf, nnact * sizeof(double));
    for (int j = nlock; j < nlock + nact; j++)
      ora_ChebAv(pol, V + (size_t)j * nl, PV + (size_t)j * nl, work);
    if (nlock > 0) { /* PV_act -= V_lock (V_lock^T PV_act) */
      gemm_tn_(n, nlock, nact, V, PV + nnlock, T);
      for (int j = 0; j < nact; j++)
        for (int i = 0; i < nlock; i++)
          axpy(n, -T[(size_t)j * nlock + i], V + (size_t)i * nl, PV + nnlock + (size_t)j * nl);
    }
    if ((it + 1) % cvcheck == 0 || it + 1 == maxit) {
      gemm_tn_(n, nact, nact, V + nnlock, PV + nnlock, T);
      if (ora_sym_eig(nact, T, nact, evecT, nact, evalT + nlock)) return -3;
      for (int j = 0; j < nact; j++) gemv_n(n, nact, V + nnlock, nl, evecT + (size_t)j * nact, buf + nnlock + (size_t)j * nl);
      memcpy(V + nnlock, buf + nnlock, nnact * sizeof(double));
      for (int j = 0; j < nact; j++) gemv_n(n, nact, PV + nnlock, nl, evecT + (size_t)j * nact, buf + nnlock + (size_t)j * nl);
      memcpy(PV + nnlock, buf + nnlock, nnact * sizeof(double));
      for (int i = nlock; i < nev; i++) {
        memcpy(R + (size_t)i * nl, PV + (size_t)i * nl, nl * sizeof(double));
        axpy(n, -evalT[i], V + (size_t)i * nl, R + (size_t)i * nl);
      }
      int nlock_new = 0, nlock_ab_new = 0;
      nact = 0;
      for (int i = nlock; i < nev; i++) {
        double resP = nrm2(n, R + (size_t)i * nl);
        int locked = 0;
        if (resP < tolP) {
          ora_matvec(V + (size_t)i * nl, buf);
          double rq = dot(n, V + (size_t)i * nl, buf);
          memcpy(R + (size_t)i * nl, buf, nl * sizeof(double));
          axpy(n, -rq, V + (size_t)i * nl, R + (size_t)i * nl);
          double resA = nrm2(n, R + (size_t)i * nl);
          if (resA < tol) {
            lock_idx[nlock_new] = i;
            res[nlock + nlock_new] = resA;
            Lam[nlock + nlock_new] = rq;
            nlock_new++;
            if (rq >= aa - EPSM && rq <= bb + EPSM) nlock_ab_new++;
            locked = 1;
          }
        }
        if (!locked) act_idx[nact++] = i;
      }
      for (int i = 0; i < nlock_new; i++) {
        memcpy(buf + nnlock + (size_t)i * nl, V + (size_t)lock_idx[i] * nl, nl * sizeof(double));
        memcpy(R + nnlock + (size_t)i * nl, PV + (size_t)lock_idx[i] * nl, nl * sizeof(double));
      }
      for (int i = 0; i < nact; i++) {
        memcpy(buf + nnlock + (size_t)(nlock_new + i) * nl, V + (size_t)act_idx[i] * nl, nl * sizeof(double));
        memcpy(R + nnlock + (size_t)(nlock_new + i) * nl, PV + (size_t)act_idx[i] * nl, nl * sizeof(double));
      }
      memcpy(V + nnlock, buf + nnlock, nnact * sizeof(double));
      memcpy(PV + nnlock, R + nnlock, nnact * sizeof(double));
      nlock += nlock_new;
      nlock_ab += nlock_ab_new;
      if ((nlock - nlock_ab) >= nbuf || nlock == nev) find_more = 0;
    }
    it++;
  }
  double *lam_o = malloc(nev * sizeof(double)), *res_o = malloc(nev * sizeof(double));
  double *Y_o = malloc(nl * nev * sizeof(double));
  int idx = 0;
  for (int i = 0; i < nlock; i++) {
    const double t = Lam[i];
    if (t >= aa - EPSM && t <= bb + EPSM) {
      lam_o[idx] = t;
      res_o[idx] = res[i];
      memcpy(Y_o + (size_t)idx * nl, V + (size_t)i * nl, nl * sizeof(double));
      idx++;
    }
  }
  *nevo = idx;
  *lamo = lam_o;
  *Yo = Y_o;
  *reso = res_o;
  free(T); free(evecT); free(evalT); free(V); free(PV); free(R); free(buf);
  free(Lam); free(res); free(act_idx); free(lock_idx); free(work); free(hw);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Lanczos DOS (SRC/landos.c:37-196), Simpson (SRC/simpson.c:25-50),   */
/* slicer on the DOS curve (SRC/spslice2.c:25-49)                      */
/* ------------------------------------------------------------------ */
void ora_simpson(const double *xi, double *yi, int npts) {
  double tc = 0.0, ti, tm, ysav;
  int i = 1;
  ysav = yi[0];
  yi[0] = 0;
  while (i < npts - 1) {
    ti = (xi[i] - xi[i - 1]) * (ysav + yi[i]) / 2.0;
    tm = tc + ti;
    ti = (xi[i + 1] - xi[i - 1]) * (ysav + 4 * yi[i] + yi[i + 1]) / 6.0;
    tc += ti;
    yi[i] = tm;
    ysav = yi[i + 1];
    yi[i + 1] = tc;
    i += 2;
  }
  if (i == npts - 1) {
    ti = (xi[i] - xi[i - 1]) * (yi[i] + ysav) / 2.0;
    yi[i] = tc + ti;
  }
}

void ora_spslicer2(const double *xi, double *yi, int n_int, int npts, double *sli) {
  int ls = 0;
  ora_simpson(xi, yi, npts);
  double t = yi[0];
  double want = (yi[npts - 1] - yi[0]) / (double)n_int;
  sli[ls] = xi[0];
  for (int k = 1; k < npts; k++) {
    if (yi[k] - t >= want) {
      ls = ls + 1;
      sli[ls] = xi[k];
      t = ORA_PLANT == 3 ? t + want : yi[k];
    }
  }
  sli[n_int] = xi[npts - 1];
}

int ora_LanDos(int nvec, int msteps, int npts, double *xdos, double *ydos, double *neig,
               const double *intv) {
  const int n = g_n;
  const size_t nl = n;
  double *v = malloc(nl * sizeof(double)), *y = calloc(npts, sizeof(double));
  int *ind = malloc(npts * sizeof(int));
  const size_t ml0 = msteps; /* allocation-time msteps, used for S indexing */
  double *S = malloc(ml0 * ml0 * sizeof(double));
  double *gamma2 = malloc(msteps * sizeof(double)), *ritz = malloc(msteps * sizeof(double));
  const double lm = intv[2], lM = intv[3];
  const double tolBdwn = 1.e-13 * (fabs(lM) + fabs(lm));
  const double aa = evsl_max(intv[0], intv[2]), bb = evsl_min(intv[1], intv[3]);
  const double kappa = 1.25;
  const int M = ORA_PLANT == 8 ? msteps : evsl_min(msteps, 30);
  const double H = (lM - lm) / (M - 1);
  const double sigma = H / sqrt(8 * log(kappa));
  double sigma2 = 2 * sigma * sigma;
  const double tol = 1e-08;
  double width = sigma * sqrt(-2.0 * log(tol));
  ora_linspace(aa, bb, npts, xdos);
  double *alp = malloc(msteps * sizeof(double)), *bet = malloc(msteps * sizeof(double));
  double *V = malloc((msteps + 1) * nl * sizeof(double));
  for (int m = 0; m < nvec; m++) {
    ora_randn_double(n, v);
    double t = dot(n, v, v);
    t = 1.0 / sqrt(t);
    scal(n, t, v);
    memcpy(V, v, nl * sizeof(double));
    double wn = 0.0;
    for (int j = 0; j < msteps; j++) {
      double *vj = V + j * nl, *vn = V + (j + 1) * nl;
      ora_matvec(vj, vn);
      if (j) axpy(n, -bet[j - 1], V + (j - 1) * nl, vn);
      alp[j] = dot(n, vn, vj);
      wn += alp[j] * alp[j];
      axpy(n, -alp[j], vj, vn);
      for (int i = 0; i <= j; i++) {
        t = dot(n, vn, V + i * nl);
        axpy(n, -t, V + i * nl, vn);
      }
      bet[j] = dot(n, vn, vn);
      if (bet[j] * (j + 1) < ORTH_TOL * wn) {
        msteps = j + 1;
        break;
      }
      if (bet[j] > tolBdwn) {
        wn += 2.0 * bet[j];
        bet[j] = sqrt(bet[j]);
        scal(n, 1.0 / bet[j], vn);
      } else {
        /* the reference draws the replacement into v, not V(:,j+1) (:132) */
        ora_randn_double(n, v);
        for (int i = 0; i <= j; i++) {
          t = dot(n, vn, V + i * nl);
          axpy(n, -t, V + i * nl, vn);
        }
        bet[j] = dot(n, vn, vn);
        wn += 2.0 * bet[j];
        bet[j] = sqrt(bet[j]);
        scal(n, 1.0 / bet[j], vn);
        bet[j] = 0;
      }
    }
    if (ora_tridiag_eig(msteps, alp, bet, ritz, S)) return -3;
    for (int i = 0; i < msteps; i++) {
      const size_t at = ORA_PLANT == 6 ? i * ml0 + msteps - 1 : i * ml0;
      gamma2[i] = S[at] * S[at];
    }
    for (int i = 0; i < msteps; i++) {
      const double tt = ritz[i];
      int numPlaced = 0;
      for (int j = 0; j < npts; j++)
        if (fabs(xdos[j] - tt) < width) ind[numPlaced++] = j;
      for (int j = 0; j < numPlaced; j++)
        y[ind[j]] += gamma2[i] * exp(-((xdos[ind[j]] - tt) * (xdos[ind[j]] - tt)) / sigma2);
    }
  }
  double scaling = 1.0 / ((ORA_PLANT == 7 ? 1 : nvec) * sqrt(sigma2 * EVSL_PI));
  scal(npts, scaling, y);
  memcpy(ydos, y, npts * sizeof(double));
  ora_simpson(xdos, y, npts);
  *neig = y[npts - 1] * n;
  free(gamma2);
  free(S);
  free(ritz);
  free(alp);
  free(bet);
  free(V);
  free(v);
  free(y);
  free(ind);
  return 0;
}

/* ------------------------------------------------------------------ */
void ora_free(void *p) { free(p); }

/* CLOCK_MONOTONIC, as SRC/timing.c:60-74 */
double ora_timer(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}
