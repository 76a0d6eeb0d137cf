// kpm.cpp -- KPM density of states on the device and the spectrum slicers.
//
// kpmdos follows SRC/spslice.c:29-148 of EVSL 1.1.1: Chebyshev moments
// mu_k = (2 - delta_k0) jac_k / (nvec pi) * sum_m v_m^T T_k(A~) v_m and the
// eigenvalue count from their exact integral.  The nvec random vectors are
// independent, so they run as column blocks of up to 32: one fused SpMM
// launch per degree advances the three-term recurrence of every column and
// emits the per-column dot products v_m^T T_k v_m, reading A once per degree
// for the whole block.  spslicer/intChx (SRC/spslice.c:155-297) and
// spslicer2 (SRC/spslice2.c:25-49) are host scalar code as in the reference.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"

namespace evsl {

// mu_k = (2 - delta_k0) jac_k raw_k / (nvec pi) with raw_0 = nvec, and the
// eigenvalue count n * integral of the KPM density over [a, b] ∩ [lmin, lmax]
// (SRC/spslice.c:57-72, 120-146)
void kpm_moments(int Mdeg, int damping, int nvec, const double* intv, int n, const double* raw,
                 double* mu, double* ecnt) {
  const double aa = std::max(intv[0], intv[2]), bb = std::min(intv[1], intv[3]);
  const double ctr = (intv[3] + intv[2]) / 2.0, wid = (intv[3] - intv[2]) / 2.0;
  const double beta1 = std::acos(std::max(-1.0 + DBL_EPSILON, (aa - ctr) / wid));
  const double beta2 = std::acos(std::min(1.0 - DBL_EPSILON, (bb - ctr) / wid));
  std::vector<double> jac(Mdeg + 1);
  dampcf(Mdeg, damping, jac.data());
  jac[0] = 1.0;
  double tcnt = -(double)nvec * jac[0] * (beta2 - beta1);
  mu[0] = (double)nvec * jac[0];
  for (int k = 1; k <= Mdeg; ++k) {
    const double t = 2 * jac[k] * raw[k];
    mu[k] = t;
    tcnt -= t * (std::sin(k * beta2) - std::sin(k * beta1)) / k;
  }
  const double t = 1.0 / (((double)nvec) * EVSL_PI);
  for (int k = 0; k <= Mdeg; ++k) mu[k] *= t;
  *ecnt = tcnt * t * (double)n;
}

namespace {

int kpm_impl(int Mdeg, int damping, int nvec, const double* intv, int kind,
             unsigned long long seed, double* mu, double* ecnt) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (c.op == 0) return EVSL_ERR_NOMAT;
  if (intv[0] >= intv[1] || intv[0] >= intv[3] || intv[1] <= intv[2]) return -1;
  if (Mdeg < 0 || nvec <= 0) return EVSL_ERR_ARG;
  const int n = c.n;
  const double ctr = (intv[3] + intv[2]) / 2.0, wid = (intv[3] - intv[2]) / 2.0;
  std::vector<double> raw(Mdeg + 1, 0.0);  // raw[k] = sum_m v_m . T_k v_m
  const int Bmax = 32;
  DevBuf X0, X, dots;
  EVSL_HIP(X0.alloc((size_t)n * Bmax * sizeof(double)));
  EVSL_HIP(X.alloc((size_t)3 * n * Bmax * sizeof(double)));
  EVSL_HIP(dots.alloc((size_t)(Mdeg + 1) * Bmax * sizeof(double)));
  std::vector<double> hbuf, hcol, hdots;
  const evsl_dev::Csr A = c.A.view();
  int done = 0;
  while (done < nvec) {
    int B = Bmax;
    while (B > nvec - done) B >>= 1;
    const size_t nb = (size_t)n * B;
    double* x0 = X0.as<double>();
    if (kind == EVSL_VEC_RAND_DOUBLE) {
      // the libc stream, vector after vector, as the reference draws it
      hbuf.resize(nb);
      hcol.resize(n);
      for (int col = 0; col < B; ++col) {
        rand_double(n, hcol.data());
        for (int r = 0; r < n; ++r) hbuf[(size_t)r * B + col] = hcol[r];
      }
      EVSL_HIP(hipMemcpy(x0, hbuf.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    } else {
      EVSL_HIP(evsl_dev::rm_random(n, B, kind == EVSL_VEC_RADEMACHER ? 0 : 1, seed, done, B, x0,
                                   c.stream));
    }
    double* s2 = red_scalar(8);
    EVSL_HIP(evsl_dev::rm_col_nrm2(n, B, x0, red_part(), s2, c.stream));
    EVSL_HIP(evsl_dev::rm_col_scale_rsqrt(n, B, x0, s2, c.stream));
    double* buf[3] = {X.as<double>(), X.as<double>() + nb, X.as<double>() + 2 * nb};
    if (c.op == 1) {
      const int nblk_kpm = evsl_dev::cheb_spmm_nblk(A, B, false, evsl_dev::kKpm);
      if (nblk_kpm < 0) return EVSL_ERR_ARG;
      for (int k = 0; k < Mdeg; ++k) {
        const double sc = (k == 0 ? 1.0 : 2.0) / wid;
        void* xkp1 = buf[k % 3];
        const void* xk = (k == 0) ? x0 : buf[(k - 1) % 3];
        const void* xkm1 = (k == 0) ? nullptr : (k == 1 ? x0 : buf[(k - 2) % 3]);
        EVSL_HIP(evsl_dev::cheb_spmm(A, B, false, evsl_dev::kKpm, xk, xkm1, xkp1, nullptr, x0,
                                     nullptr, nullptr, sc, ctr, false, 0, red_part(), c.stream));
        EVSL_HIP(evsl_dev::sum_block_partials(red_part(), nblk_kpm, B,
                                              dots.as<double>() + (size_t)(k + 1) * B,
                                              red_scratch(), c.stream));
        c.astats.filter_bytes += c.A.mat_bytes(8, B) + 4.0 * (n + 1) +
                                 (double)(k == 0 ? 3 : 4) * nb * sizeof(double);
      }
      c.astats.n_filter_steps += Mdeg;
      c.astats.n_filter_cols += (size_t)Mdeg * B;
    } else {
      // user operator: one column at a time through the callback
      DevBuf col;
      EVSL_HIP(col.alloc((size_t)4 * n * sizeof(double)));
      double* v = col.as<double>();
      double* cb[3] = {v + n, v + 2 * (size_t)n, v + 3 * (size_t)n};
      std::vector<double> vh(nb), dh((size_t)(Mdeg + 1) * B, 0.0);
      EVSL_HIP(hipMemcpy(vh.data(), x0, nb * sizeof(double), hipMemcpyDeviceToHost));
      hcol.resize(n);
      for (int cc = 0; cc < B; ++cc) {
        for (int r = 0; r < n; ++r) hcol[r] = vh[(size_t)r * B + cc];
        EVSL_HIP(hipMemcpy(v, hcol.data(), n * sizeof(double), hipMemcpyHostToDevice));
        for (int k = 0; k < Mdeg; ++k) {
          const double sc = (k == 0 ? 1.0 : 2.0) / wid;
          double* xkp1 = cb[k % 3];
          const double* xk = (k == 0) ? v : cb[(k - 1) % 3];
          const double* xkm1 = (k == 0) ? nullptr : (k == 1 ? v : cb[(k - 2) % 3]);
          if (int e = apply_A(xk, xkp1)) return e;
          EVSL_HIP(evsl_dev::scal(n, sc, xkp1, c.stream));
          EVSL_HIP(evsl_dev::axpy_dot(n, -sc * ctr, xk, xkp1, nullptr, nullptr, c.stream));
          if (xkm1) EVSL_HIP(evsl_dev::axpy_dot(n, -1.0, xkm1, xkp1, nullptr, nullptr, c.stream));
          EVSL_HIP(evsl_dev::dot_partial(n, xkp1, v, red_part(), c.stream));
          EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(0),
                                        c.stream));
          double d;
          if (int e = read_scalars(red_scalar(0), 1, &d)) return e;
          dh[(size_t)(k + 1) * B + cc] = d;
        }
      }
      EVSL_HIP(hipMemcpy(dots.p, dh.data(), dh.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    hdots.resize((size_t)(Mdeg + 1) * B);
    EVSL_HIP(hipMemcpy(hdots.data(), dots.p, hdots.size() * sizeof(double),
                       hipMemcpyDeviceToHost));
    for (int k = 1; k <= Mdeg; ++k)
      for (int col = 0; col < B; ++col) raw[k] += hdots[(size_t)k * B + col];
    done += B;
  }
  kpm_moments(Mdeg, damping, nvec, intv, n, raw.data(), mu, ecnt);
  return 0;
}

// yi[j] = integral_{xi[0]}^{xi[j]} of the KPM density (SRC/spslice.c:155-181)
void int_chx(int Mdeg, const double* mu, int npts, const double* xi, double* yi) {
  std::vector<double> th(npts);
  for (int j = 0; j < npts; j++) th[j] = std::acos(xi[j]);
  const double th0 = th[0];
  for (int j = 0; j < npts; j++) yi[j] = mu[0] * (th0 - th[j]);
  for (int k = 1; k <= Mdeg; k++) {
    const double v0 = std::sin(k * th0) / k;
    for (int j = 0; j < npts; j++) yi[j] += mu[k] * (v0 - std::sin(k * th[j]) / k);
  }
}

}  // namespace
}  // namespace evsl

using namespace evsl;

extern "C" {

int kpmdos(int Mdeg, int damping, int nvec, double* intv, double* mu, double* ecnt) {
  return kpm_impl(Mdeg, damping, nvec, intv, EVSL_VEC_RAND_DOUBLE, 0ULL, mu, ecnt);
}

int kpmdos_ex(int Mdeg, int damping, int nvec, double* intv, int vec_kind,
              unsigned long long seed, double* mu, double* ecnt) {
  if (vec_kind < EVSL_VEC_RAND_DOUBLE || vec_kind > EVSL_VEC_GAUSSIAN) return EVSL_ERR_ARG;
  return kpm_impl(Mdeg, damping, nvec, intv, vec_kind, seed, mu, ecnt);
}

// equal-count slices on a linspace grid (SRC/spslice.c:205-297)
int spslicer(double* sli, double* mu, int Mdeg, double* intv, int n_int, int npts) {
  if (intv[0] >= intv[1] || intv[0] >= intv[3] || intv[1] <= intv[2]) return -1;
  const double aa = std::max(intv[0], intv[2]), bb = std::min(intv[1], intv[3]);
  std::memset(sli, 0, (size_t)(n_int + 1) * sizeof(double));
  if (n_int == 1) {
    sli[0] = intv[0];
    sli[1] = intv[1];
    return 0;
  }
  const double ctr = (intv[3] + intv[2]) / 2, wid = (intv[3] - intv[2]) / 2;
  const double aL = std::max((aa - ctr) / wid, -1.0), bL = std::min((bb - ctr) / wid, 1.0);
  npts = std::max(npts, 2 * n_int + 1);
  std::vector<double> xi(npts), yi(npts);
  linspace(aL, bL, npts, xi.data());
  int_chx(Mdeg, mu, npts, xi.data(), yi.data());
  double target = yi[npts - 1] / (double)n_int;
  int ls = 0, ii = 0, err = 0;
  sli[0] = intv[0];
  while (++ls < n_int) {
    while (ii < npts && yi[ii] < target) ii++;
    if (ii == npts) break;
    // the nearer of the two grid points; the reference reads yi[-1] when
    // target <= yi[0] (a density whose integral is not positive): the first
    // grid point has no lower neighbour and is kept
    if (ii > 0 && (target - yi[ii - 1]) < (yi[ii] - target)) ii--;
    sli[ls] = ctr + wid * xi[ii];
    target = yi[ii] + (yi[npts - 1] - yi[ii]) / (n_int - ls);
  }
  sli[n_int] = intv[1];
  if (ls != n_int) err = 1;
  for (ii = 1; ii <= n_int; ii++)
    if (sli[ii] < sli[ii - 1]) {
      err += 2;
      break;
    }
  return err;
}

// slices from a Lanczos DOS curve: Simpson integral, then equal shares
// (SRC/spslice2.c:25-49, SRC/simpson.c:25-50)
void spslicer2(double* xi, double* yi, int n_int, int npts, double* sli) {
  double tc = 0.0, ti, tm, ysav = yi[0];
  int i = 1;
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
  if (i == npts - 1) yi[i] = tc + (xi[i] - xi[i - 1]) * (yi[i] + ysav) / 2.0;
  double t = yi[0];
  const double want = (yi[npts - 1] - yi[0]) / (double)n_int;
  int ls = 0;
  sli[ls] = xi[0];
  for (int k = 1; k < npts; k++) {
    // (ls < n_int: a further boundary would fall past sli[n_int], which the
    // reference writes out of bounds when rounding lets one more share fit)
    if (yi[k] - t >= want && ls < n_int) {
      ls = ls + 1;
      sli[ls] = xi[k];
      t = yi[k];
    }
  }
  sli[n_int] = xi[npts - 1];
}

// The exact smoothed DOS of known eigenvalues (SRC/exDOS.c:17-69): Gaussians
// of width sigma = H / sqrt(8 ln 1.25), H = (lmax - lmin) / (min(n, 30) - 1),
// cut where they fall below 1e-8, on npts points of [a, b], scaled by
// 1 / (n sqrt(2 pi sigma^2)).  The reference walks each Gaussian's points by
// repeated addition of the grid step (xi += h) from its first point; the same
// walk here keeps its values.  Returns 1 when the last eigenvalue's walk left
// the point counter at 0 (in particular when no eigenvalue lies in [a, b]),
// else 0.  Host-only.
int exDOS(double* vals, int n, int npts, double* x, double* y, double* intv) {
  const double lm = intv[2], lM = intv[3], a = intv[0], b = intv[1];
  const int M = std::min(n, 30);
  const double H = (lM - lm) / (M - 1);
  const double sigma = H / std::sqrt(8 * std::log(1.25));
  const double sigma2 = 2 * sigma * sigma;
  const double width = sigma * std::sqrt(-2.0 * std::log(1e-08));
  linspace(a, b, npts, x);
  const double h = x[1] - x[0];
  std::memset(y, 0, (size_t)npts * sizeof(double));
  int j = 0;
  for (int i = 0; i < n; i++) {
    const double t = vals[i];
    if (t < a || t > b) continue;
    j = std::max((int)((t - a - width) / h), 0);
    const double stop = std::min(t + width, b);
    // j < npts: the reference relies on the accumulated xi passing b first
    for (double xi = a + j * h; xi <= stop && j < npts; xi += h) {
      y[j] += std::exp(-(xi - t) * (xi - t) / sigma2);
      j++;
    }
  }
  if (j == 0) return 1;
  const double scaling = 1.0 / (n * std::sqrt(sigma2 * EVSL_PI));
  for (int k = 0; k < npts; k++) y[k] *= scaling;
  return 0;
}

// Time nsteps fused filter launches (block width B, prec) on the current
// matrix with HIP events; reports the average launch time and the
// algorithmic bytes of one launch.  mode 1: ChebAv step, 2: KPM step.
int evsl_amd_bench_filter(int B, int prec, int mode, int nsteps, double cc, double dd,
                          double* ms_per_step, double* bytes_per_step) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (c.op != 1) return EVSL_ERR_NOMAT;
  if (nsteps <= 0 || !(dd > 0.0) || evsl_dev::width_slot(B) < 0) return EVSL_ERR_ARG;
  if (mode == 2 && (B & (B - 1))) return EVSL_ERR_ARG;
  const bool f32 = (prec == EVSL_PREC_MIXED) && mode != 2;
  FilterOp op;
  if (int e = make_filter_op(cc, dd, f32, &op)) return e;
  const int n = c.n;
  const size_t vs = f32 ? 4 : 8;
  const size_t nb = (size_t)n * B;
  // ring of three blocks and the residual block of the vector size, y FP64
  DevBuf X, coef;
  const size_t xbytes = 4 * nb * vs + nb * sizeof(double);
  EVSL_HIP(X.alloc(xbytes));
  EVSL_HIP(coef.alloc(256 * sizeof(double)));
  char* base = X.as<char>();
  EVSL_HIP(hipMemsetAsync(base, 0, xbytes, c.stream));
  EVSL_HIP(evsl_dev::fill(256, 0.01, coef.as<double>(), c.stream));
  // coefficient rows with room for the look-back of a 3-term y flush
  const double* scol = coef.as<double>() + 64;
  const double* mu0col = coef.as<double>() + 160;
  // buf[3] = y (FP64), buf[4] = v0 / R
  void* buf[5];
  for (int i = 0; i < 3; ++i) buf[i] = base + i * nb * vs;
  buf[4] = base + 3 * nb * vs;
  buf[3] = base + 4 * nb * vs;
  const evsl_dev::Mode md =
      mode == 2 ? evsl_dev::kKpm : (mode == 3 ? evsl_dev::kRes : evsl_dev::kCheb);
  // steady state of a long series: y is flushed (3 terms) every third launch
  nsteps = (nsteps + 2) / 3 * 3;
  // the layout tuner settles first (its measuring launches run candidate
  // layouts), so the timed launches run the layout the solvers will use
  if (md != evsl_dev::kKpm && op.A.tune) {
    for (int it = 0; it < 4096 && !evsl_dev::spmm_tuner_decided(op.A.tune, B, f32, md); ++it) {
      EVSL_HIP(evsl_dev::cheb_spmm(op.A, B, f32, md, buf[it % 3], buf[(it + 1) % 3],
                                   buf[(it + 2) % 3], buf[3], buf[4], scol, mu0col, op.t2, op.cc,
                                   false, it % 3 == 2 ? 3 : 0, red_part(), c.stream));
      if (it % 64 == 63) EVSL_HIP(hipStreamSynchronize(c.stream));
    }
  }
  hipEvent_t e0, e1;
  EVSL_HIP(hipEventCreate(&e0));
  EVSL_HIP(hipEventCreate(&e1));
  int flushes = 0;
  for (int it = 0; it < nsteps + 3; ++it) {
    if (it == 3) EVSL_HIP(hipEventRecord(e0, c.stream));
    const int yf = (mode == 2 || it % 3 != 2) ? 0 : 3;
    if (it >= 3 && yf) ++flushes;
    EVSL_HIP(evsl_dev::cheb_spmm(op.A, B, f32, md, buf[it % 3], buf[(it + 1) % 3],
                                 buf[(it + 2) % 3], buf[3], buf[4], scol, mu0col, op.t2,
                                 mode == 2 ? cc : op.cc, false, yf, red_part(), c.stream));
  }
  EVSL_HIP(hipEventRecord(e1, c.stream));
  EVSL_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  EVSL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_step = ms / nsteps;
  // per launch: KPM xk (gathered once), xkm1, v0 in, xkp1 out; ChebAv xk,
  // xkm1 in, xkp1 out, plus the FP64 y in and out on a flush; the residual
  // recurrence adds the residual block R
  double vbytes;
  if (mode == 2)
    vbytes = 4.0 * vs;
  else
    vbytes = 3.0 * vs + 16.0 * flushes / nsteps + (mode == 3 ? vs : 0.0);
  *bytes_per_step = c.A.mat_bytes((int)vs, B) + 4.0 * (n + 1) + vbytes * nb;
  return 0;
}

int evsl_amd_bench_dense(int kind, int n, int k, int m, int reps, double* ms_per_rep,
                         double* bytes, double* flops) {
  if (int e = ensure_started()) return e;
  if (kind < 0 || kind > 3 || n <= 0 || k <= 0 || m <= 0 || reps <= 0 || !ms_per_rep ||
      !bytes || !flops)
    return EVSL_ERR_ARG;
  Context& c = ctx();
  const size_t nn = (size_t)n;
  DevBuf A, B, C, part, h, A32;
  // kind 0: A = V (n x k), B = w (n); kind 1: A = V, B = Z (k x m), C = n x m;
  // kind 2: A = U (n x m), C = m x m; kind 3: kind 0 with A32 = (float)V,
  // columns ld32 apart
  const size_t acols = kind == 2 ? (size_t)m : (size_t)k;
  const size_t ld32 = (nn + 63) / 64 * 64;
  EVSL_HIP(A.alloc(nn * acols * sizeof(double)));
  EVSL_HIP(evsl_dev::rm_random(nn * acols, 1, 1, 17, 0, 1, A.as<double>(), c.stream));
  if (kind == 3) {
    EVSL_HIP(A32.alloc(ld32 * acols * sizeof(float)));
    for (size_t j = 0; j < acols; ++j)
      EVSL_HIP(evsl_dev::f64_to_f32(nn, A.as<double>() + j * nn, A32.as<float>() + j * ld32,
                                    c.stream));
  }
  if (kind == 0 || kind == 3) {
    EVSL_HIP(B.alloc(nn * sizeof(double)));
    EVSL_HIP(evsl_dev::rm_random(nn, 1, 1, 18, 0, 1, B.as<double>(), c.stream));
    // V^T w partials (chunks x k), and one norm partial per 256 rows
    const size_t np = std::max((size_t)evsl_dev::gemv_t_chunks(nn) * (k + 1),
                               (nn + evsl_dev::kWG - 1) / evsl_dev::kWG);
    EVSL_HIP(part.alloc(np * sizeof(double)));
    EVSL_HIP(h.alloc((size_t)(k + 1) * sizeof(double)));
  } else if (kind == 1) {
    EVSL_HIP(B.alloc((size_t)k * m * sizeof(double)));
    EVSL_HIP(evsl_dev::rm_random((size_t)k * m, 1, 1, 19, 0, 1, B.as<double>(), c.stream));
    EVSL_HIP(C.alloc(nn * m * sizeof(double)));
  } else {
    EVSL_HIP(C.alloc((size_t)m * m * sizeof(double)));
    EVSL_HIP(part.alloc(evsl_dev::gemm_tn_part_elems(nn, m, m) * sizeof(double)));
  }
  auto once = [&]() -> hipError_t {
    if (kind == 3) {
      hipError_t e = evsl_dev::gemv_t_f32(nn, k, A32.as<float>(), ld32, B.as<double>(),
                                          part.as<double>(), h.as<double>(), c.stream);
      if (e != hipSuccess) return e;
      e = evsl_dev::scal(k, 1e-30, h.as<double>(), c.stream);
      if (e != hipSuccess) return e;
      return evsl_dev::gemv_n_sub_nrm_f32(nn, k, A32.as<float>(), ld32, h.as<double>(),
                                          B.as<double>(), part.as<double>(), c.stream);
    }
    if (kind == 0) {
      hipError_t e = evsl_dev::gemv_t(nn, k, A.as<double>(), nn, B.as<double>(),
                                      part.as<double>(), h.as<double>(), c.stream);
      if (e != hipSuccess) return e;
      // the pass subtracts V h from w: scale h so that w stays bounded
      e = evsl_dev::scal(k, 1e-30, h.as<double>(), c.stream);
      if (e != hipSuccess) return e;
      return evsl_dev::gemv_n_sub_nrm(nn, k, A.as<double>(), nn, h.as<double>(), B.as<double>(),
                                      part.as<double>(), c.stream);
    }
    if (kind == 1)
      return evsl_dev::gemm_nn(nn, m, k, A.as<double>(), nn, B.as<double>(), k, C.as<double>(),
                               nn, c.stream);
    return evsl_dev::gemm_tn(nn, m, m, A.as<double>(), nn, A.as<double>(), nn, C.as<double>(), m,
                             part.as<double>(), c.stream);
  };
  EVSL_HIP(once());  // warm-up
  hipEvent_t e0, e1;
  EVSL_HIP(hipEventCreate(&e0));
  EVSL_HIP(hipEventCreate(&e1));
  EVSL_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) EVSL_HIP(once());
  EVSL_HIP(hipEventRecord(e1, c.stream));
  EVSL_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  EVSL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_rep = ms / reps;
  const double dn = (double)n, dk = (double)k, dm = (double)m;
  if (kind == 0 || kind == 3) {
    // V twice (8 or 4 bytes an entry), w in, out, in
    *bytes = (kind == 0 ? 16.0 : 8.0) * dn * dk + 24.0 * dn;
    *flops = 4.0 * dn * dk;
  } else if (kind == 1) {
    *bytes = 8.0 * (dn * dk + dn * dm + dk * dm);
    *flops = 2.0 * dn * dk * dm;
  } else {
    *bytes = 8.0 * (dn * dm + dm * dm);
    *flops = 2.0 * dn * dm * dm;
  }
  return 0;
}

int evsl_amd_spmm_index_bytes(void) {
  const Context& c = ctx();
  return c.op == 1 ? c.A.idx_bytes() : 0;
}

int evsl_amd_spmm_index_mode(int B, int prec, int mode) {
  const Context& c = ctx();
  if (c.op != 1 || B < 1) return -1;
  const int vs = (prec == EVSL_PREC_MIXED && mode != 2) ? 4 : 8;
  return (c.A.pc && B * vs <= evsl_dev::kPatMaxRowBytes) ? (c.A.pshort ? 2 : 3)
                                                         : (c.A.jd ? 1 : 0);
}

int evsl_amd_set_pattern_codes(int on) {
  ctx().pattern_codes = on != 0;
  return 0;
}

int evsl_amd_pattern_codes(void) {
  const Context& c = ctx();
  return c.op == 1 && c.A.pc ? c.A.npat : 0;
}

int evsl_amd_spmm_layout(int B, int prec, int mode, int* decided) {
  const Context& c = ctx();
  const evsl_dev::Mode md =
      mode == 2 ? evsl_dev::kKpm : (mode == 3 ? evsl_dev::kRes : evsl_dev::kCheb);
  const bool f32 = prec == EVSL_PREC_MIXED && mode != 2;
  if (decided) *decided = 1;
  if (mode < 1 || mode > 3) return -1;
  if (c.op != 1 || !c.A.tune || !c.spmm_tuning) return evsl_dev::spmm_table_layout(B, f32, md);
  if (decided) *decided = evsl_dev::spmm_tuner_decided(c.A.tune, B, f32, md);
  return evsl_dev::spmm_tuner_layout(c.A.tune, B, f32, md);
}

int evsl_amd_set_spmm_layout(int B, int prec, int mode, int layout) {
  const Context& c = ctx();
  if (c.op != 1 || !c.A.tune) {
    set_error_msg("evsl_amd_set_spmm_layout: no device matrix (SetAMatrix first)");
    return EVSL_ERR_NOMAT;
  }
  if (mode != 1 && mode != 3) return EVSL_ERR_ARG;
  const evsl_dev::Mode md = mode == 3 ? evsl_dev::kRes : evsl_dev::kCheb;
  if (evsl_dev::spmm_tuner_set(c.A.tune, B, prec == EVSL_PREC_MIXED, md, layout) != 0) {
    set_error_msg("evsl_amd_set_spmm_layout: B must be 4, 8, 16 or 32 and the layout one of "
                  "its candidates, 0 or -1");
    return EVSL_ERR_ARG;
  }
  return 0;
}

int evsl_amd_set_spmm_tuning(int on) {
  ctx().spmm_tuning = on != 0;
  return 0;
}

int evsl_amd_spmm_vec(int B, int prec, int mode) {
  const int L = evsl_amd_spmm_layout(B, prec, mode, nullptr);
  return L > 0 ? L / 10 : -1;
}

int evsl_amd_spmm_streaming(int B, int prec, int mode) {
  const Context& c = ctx();
  if (c.op != 1) return 0;
  const evsl_dev::Mode md =
      mode == 2 ? evsl_dev::kKpm : (mode == 3 ? evsl_dev::kRes : evsl_dev::kCheb);
  return evsl_dev::cheb_spmm_streaming(c.A.view(), B, prec == EVSL_PREC_MIXED && mode != 2, md)
             ? 1
             : 0;
}

int evsl_amd_bench_copy(size_t bytes, int reps, double* ms_per_copy) {
  if (int e = ensure_started()) return e;
  if (reps <= 0 || bytes < 16 || !ms_per_copy) return EVSL_ERR_ARG;
  Context& c = ctx();
  bytes &= ~(size_t)15;
  DevBuf src, dst;
  EVSL_HIP(src.alloc(bytes));
  EVSL_HIP(dst.alloc(bytes));
  EVSL_HIP(evsl_dev::fill(bytes / 8, 1.0, src.as<double>(), c.stream));
  hipEvent_t e0, e1;
  EVSL_HIP(hipEventCreate(&e0));
  EVSL_HIP(hipEventCreate(&e1));
  for (int it = 0; it < reps + 2; ++it) {
    if (it == 2) EVSL_HIP(hipEventRecord(e0, c.stream));
    EVSL_HIP(evsl_dev::stream_copy(bytes, src.p, dst.p, c.stream));
  }
  EVSL_HIP(hipEventRecord(e1, c.stream));
  EVSL_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  EVSL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_copy = ms / reps;
  return 0;
}

}  // extern "C"
