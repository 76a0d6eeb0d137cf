// lanczos.cpp -- filtered Lanczos on the device: ChebAv, CGS/DGKS,
// ChebLanNr, LanBounds and LanDos of the drop-in interface.
//
// Follows SRC/chebpoly.c:519-598 (ChebAv), SRC/misc_la.c:221-261 (CGS_DGKS),
// SRC/cheblanNr.c:57-434, SRC/lanbounds.c:21-144 and SRC/landos.c:37-196 of
// EVSL 1.1.1.  The Lanczos basis V stays in HBM (column-major, n x (maxit+1));
// all vector work runs in HIP kernels; the host keeps the scalar recurrence
// (alpha, beta), the tridiagonal eigenvalues and the control flow.  One
// device->host read per Lanczos step carries alpha and the DGKS norms.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"

namespace evsl {

namespace {

constexpr double kOrthTol = 1e-14;  // INC/internal_header.h:21
constexpr int kNgsMax = 2;          // INC/internal_header.h:24
const double kEpsM = EVSL_DBL_EPS_MULT * DBL_EPSILON;


}  // namespace

int check_intv(const double* intv) {
  if (intv[0] >= intv[1]) return -1;
  if (intv[0] >= intv[3] || intv[1] <= intv[2]) return -2;
  return 0;
}

int as_device_vector(const double* p, size_t n, DevBuf& tmp, const double** out) {
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e == hipSuccess && (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged)) {
    *out = p;
    return 0;
  }
  (void)hipGetLastError();  // clear the error of an unregistered host pointer
  EVSL_HIP(tmp.alloc(n * sizeof(double)));
  EVSL_HIP(hipMemcpy(tmp.p, p, n * sizeof(double), hipMemcpyHostToDevice));
  *out = tmp.as<double>();
  return 0;
}

namespace {

// HIP-event phase timing accumulated into the extension statistics
struct EventClock {
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool ok = false;
  EventClock() {
    ok = true;
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) ok = false;
  }
  ~EventClock() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  void mark(int i) {
    if (ok) (void)hipEventRecord(ev[i], ctx().stream);
  }
  double secs(int a, int b) {
    float ms = 0.f;
    if (!ok || hipEventSynchronize(ev[b]) != hipSuccess ||
        hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess)
      return 0.0;
    return ms * 1e-3;
  }
};

// matrix bytes of one launch (values of sa bytes) plus vec_bytes per row
// and column of the B-column block
double filter_step_bytes(const Context& c, int B, int sa, double vec_bytes) {
  return c.A.mat_bytes((int)sa, B) + 4.0 * (c.n + 1) + vec_bytes * B * c.n;
}

}  // namespace

// ---------------------------------------------------------------------
// y = p((A - cc)/dd) v  -- one fused SpMM launch per degree
// ---------------------------------------------------------------------
int cheb_av(const polparams* pol, const double* v, double* y, ChebWork& w, int prec) {
  Context& c = ctx();
  const int n = c.n, m = pol->deg;
  const double cc = pol->cc, t1 = 1.0 / pol->dd, t2 = 2.0 / pol->dd;
  (void)t1;
  (void)t2;
  if (w.deg_loaded != m || w.mu_src != pol->mu) {
    EVSL_HIP(w.mu_d.alloc((size_t)(m + 1) * sizeof(double)));
    EVSL_HIP(hipMemcpy(w.mu_d.p, pol->mu, (size_t)(m + 1) * sizeof(double),
                       hipMemcpyHostToDevice));
    w.deg_loaded = m;
    w.mu_src = pol->mu;
  }
  const double* mu_d = w.mu_d.as<double>();
  if (m == 0) {
    EVSL_HIP(evsl_dev::scal_copy(n, pol->mu[0], v, y, c.stream));
    return 0;
  }
  if (c.op == 2 && prec != EVSL_PREC_FP64) prec = EVSL_PREC_FP64;  // user operator: FP64 only
  if (c.op == 1) {
    const bool f32 = (prec == EVSL_PREC_MIXED);
    const size_t vs = f32 ? sizeof(float) : sizeof(double);
    FilterOp op;
    if (int e = make_filter_op(cc, pol->dd, f32, &op)) return e;
    // the kernel accumulates y in FP64 in both precisions, straight into
    // the caller's vector
    DevBuf& store = f32 ? w.b32 : w.bufs;
    const int nvec = f32 ? 4 : 3;
    EVSL_HIP(store.alloc((size_t)nvec * n * vs));
    char* base = store.as<char>();
    void* buf[3] = {base, base + (size_t)n * vs, base + 2 * (size_t)n * vs};
    const void* vin = v;
    void* yout = y;
    if (f32) {
      void* v32 = base + 3 * (size_t)n * vs;
      EVSL_HIP(evsl_dev::f64_to_f32(n, v, (float*)v32, c.stream));
      vin = v32;
    }
    for (int k = 1; k <= m; ++k) {
      void* xkp1 = buf[(k - 1) % 3];
      const void* xk = (k == 1) ? vin : buf[(k - 2) % 3];
      const void* xkm1 = (k == 1) ? nullptr : (k == 2 ? vin : buf[(k - 3) % 3]);
      const int yf = evsl_dev::y_flush(k, m);
      EVSL_HIP(evsl_dev::cheb_spmm(op.A, 1, f32, evsl_dev::kCheb, xk, xkm1, xkp1, yout, nullptr,
                                   mu_d + k, mu_d, k == 1 ? op.t1 : op.t2, op.cc, k == 1, yf,
                                   nullptr, c.stream));
      c.astats.filter_bytes +=
          filter_step_bytes(c, 1, (int)vs, evsl_dev::cheb_vec_bytes(yf, k == 1, (int)vs));
    }
    c.astats.n_filter_steps += m;
    c.astats.n_filter_cols += m;
    return 0;
  }
  // user operator: y = mu0 v; per degree vkp1 = A vk then a fused update
  EVSL_HIP(w.bufs.alloc((size_t)4 * n * sizeof(double)));
  double* b = w.bufs.as<double>();
  double* buf[3] = {b, b + n, b + 2 * (size_t)n};
  double* tmp = b + 3 * (size_t)n;
  EVSL_HIP(evsl_dev::scal_copy(n, pol->mu[0], v, y, c.stream));
  for (int k = 1; k <= m; ++k) {
    double* xkp1 = buf[(k - 1) % 3];
    const double* xk = (k == 1) ? v : buf[(k - 2) % 3];
    const double* xkm1 = (k == 1) ? nullptr : (k == 2 ? v : buf[(k - 3) % 3]);
    if (int e = apply_A(xk, tmp)) return e;
    // xkp1 = t*(tmp - cc*xk) - xkm1 ; y += mu_k xkp1
    const double t = (k == 1 ? t1 : t2);
    EVSL_HIP(evsl_dev::scal_copy(n, t, tmp, xkp1, c.stream));
    EVSL_HIP(evsl_dev::axpy_dot(n, -t * cc, xk, xkp1, nullptr, nullptr, c.stream));
    if (xkm1) EVSL_HIP(evsl_dev::axpy_dot(n, -1.0, xkm1, xkp1, nullptr, nullptr, c.stream));
    EVSL_HIP(evsl_dev::axpy_dot(n, pol->mu[k], xkp1, y, nullptr, nullptr, c.stream));
  }
  c.astats.n_filter_steps += m;
  c.astats.n_filter_cols += m;
  return 0;
}

// CGS with the DGKS test, two GEMV passes at most (SRC/misc_la.c:221-261).
// On entry *old2 = ||v||^2 (already known); returns ||v|| in *nrm.
int cgs_dgks(int n, int k, const double* Q, double* v, double old2, double* nrm, LanWork& L) {
  Context& c = ctx();
  const double eta = 1.0 / std::sqrt(2.0);
  double old_nrm = std::sqrt(old2), new_nrm = 0.0;
  const int nb = (n + evsl_dev::kWG - 1) / evsl_dev::kWG;
  for (int it = 0; it < kNgsMax; ++it) {
    EVSL_HIP(evsl_dev::gemv_t(n, k, Q, n, v, L.gt_part.as<double>(), L.h.as<double>(), c.stream));
    EVSL_HIP(evsl_dev::gemv_n_sub_nrm(n, k, Q, n, L.h.as<double>(), v, red_part(), c.stream));
    EVSL_HIP(evsl_dev::sum_scalar(red_part(), nb, red_scalar(2), c.stream));
    double s2;
    if (int e = read_scalars(red_scalar(2), 1, &s2)) return e;
    new_nrm = std::sqrt(s2);
    c.astats.n_reorth_pass++;
    c.astats.reorth_bytes += 16.0 * n * k + 24.0 * n;
    if (new_nrm > eta * old_nrm) break;
    old_nrm = new_nrm;
  }
  *nrm = new_nrm;
  return 0;
}

int nrm2_dev(int n, const double* x, double* out) {
  Context& c = ctx();
  EVSL_HIP(evsl_dev::dot_partial(n, x, x, red_part(), c.stream));
  EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(0), c.stream));
  double s;
  if (int e = read_scalars(red_scalar(0), 1, &s)) return e;
  *out = std::sqrt(s);
  return 0;
}


// Ritz extraction shared by the filtered Lanczos solvers: given the
// tridiagonal T_kdim (dT, eT) and the basis V, compute Ritz pairs of A for
// the eigenvalues of T at or above bar, keep those in [a, b] with residual
// below tol (SRC/cheblanNr.c:316-399).
static int ritz_extract(int n, int kdim, const double* dT, const double* eT, const double* V,
                        double bar, double aa, double bb, double tol, int* nevOut,
                        double** lamo, double** Wo, double** reso, int maxit_alloc) {
  Context& c = ctx();
  std::vector<double> eval(kdim), e(kdim + 1, 0.0);
  std::memcpy(eval.data(), dT, kdim * sizeof(double));
  if (kdim > 1) std::memcpy(e.data(), eT, (kdim - 1) * sizeof(double));
  RotationLog rec;
  if (int err = tridiag_ql(kdim, eval.data(), e.data(), &rec)) return err;
  std::vector<int> sel;
  for (int i = 0; i < kdim; ++i)
    if (!(eval[i] < bar)) sel.push_back(i);
  const int m = (int)sel.size();
  double* Lam = (double*)std::malloc((size_t)std::max(maxit_alloc, 1) * sizeof(double));
  double* res = (double*)std::malloc((size_t)std::max(maxit_alloc, 1) * sizeof(double));
  *lamo = Lam;
  *reso = res;
  *Wo = nullptr;
  *nevOut = 0;
  if (m == 0) return 0;
  DevBuf Zsel, U, W, part, s2, tv, r2;
  EVSL_HIP(Zsel.alloc((size_t)kdim * m * sizeof(double)));
  if (int err = tridiag_rotations_to_device(kdim, rec, sel, Zsel.as<double>())) return err;
  EVSL_HIP(U.alloc((size_t)n * m * sizeof(double)));
  EVSL_HIP(W.alloc((size_t)n * m * sizeof(double)));
  const int nch = (n + 8191) / 8192;
  EVSL_HIP(part.alloc((size_t)nch * m * sizeof(double)));
  EVSL_HIP(s2.alloc(m * sizeof(double)));
  EVSL_HIP(tv.alloc(m * sizeof(double)));
  EVSL_HIP(r2.alloc(m * sizeof(double)));
  double* Ud = U.as<double>();
  double* Wd = W.as<double>();
  // U = V(:, 0:kdim) * Zsel
  EVSL_HIP(evsl_dev::gemm_nn(n, m, kdim, V, n, Zsel.as<double>(), kdim, Ud, n, c.stream));
  // normalise the columns
  EVSL_HIP(evsl_dev::col_dot(n, m, Ud, n, Ud, n, part.as<double>(), s2.as<double>(), c.stream));
  EVSL_HIP(evsl_dev::col_scale_rsqrt(n, m, Ud, n, s2.as<double>(), c.stream));
  // W = A U, Rayleigh quotients, residual norms
  for (int j = 0; j < m; ++j)
    if (int err = apply_A(Ud + (size_t)j * n, Wd + (size_t)j * n)) return err;
  EVSL_HIP(evsl_dev::col_dot(n, m, Wd, n, Ud, n, part.as<double>(), tv.as<double>(), c.stream));
  EVSL_HIP(evsl_dev::col_res2(n, m, Wd, n, Ud, n, tv.as<double>(), part.as<double>(),
                              r2.as<double>(), c.stream));
  std::vector<double> hs2(m), ht(m), hr2(m);
  EVSL_HIP(hipMemcpy(hs2.data(), s2.p, m * sizeof(double), hipMemcpyDeviceToHost));
  EVSL_HIP(hipMemcpy(ht.data(), tv.p, m * sizeof(double), hipMemcpyDeviceToHost));
  EVSL_HIP(hipMemcpy(hr2.data(), r2.p, m * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int> keep;
  int nev = 0;
  for (int j = 0; j < m; ++j) {
    if (hs2[j] == 0.0) return 2;  // zero Ritz vector (SRC/cheblanNr.c:362-364)
    const double t = ht[j];
    if (t < aa - kEpsM || t > bb + kEpsM) continue;
    const double r = std::sqrt(std::max(hr2[j], 0.0));
    if (r < tol) {
      Lam[nev] = t;
      res[nev] = r;
      keep.push_back(j);
      ++nev;
    }
  }
  *nevOut = nev;
  if (nev > 0) {
    double* Y = (double*)_evsl_Malloc_device((size_t)n * nev * sizeof(double));
    if (!Y) return EVSL_ERR_ALLOC;
    DevBuf idx;
    EVSL_HIP(idx.alloc(nev * sizeof(int)));
    EVSL_HIP(hipMemcpy(idx.p, keep.data(), nev * sizeof(int), hipMemcpyHostToDevice));
    EVSL_HIP(evsl_dev::gather_cols(n, nev, Ud, n, idx.as<int>(), Y, n, c.stream));
    EVSL_HIP(hipStreamSynchronize(c.stream));
    *Wo = Y;
  }
  return 0;
}

}  // namespace evsl

using namespace evsl;

extern "C" {

// SRC/cheblanNr.c:57-434
int ChebLanNr(double* intv, int maxit, double tol, double* vinit, polparams* pol, int* nevOut,
              double** lamo, double** Wo, double** reso, FILE* fstats) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (c.op == 0) return EVSL_ERR_NOMAT;
  // MIXED filter precision on a CSR matrix: the one-slice case of the
  // batched solver, whose FP64 Rayleigh-Ritz refinement brings the
  // FP32-filtered Ritz pairs to tol
  if (c.filter_prec == EVSL_PREC_MIXED && c.op == 1)
    return ChebLanNr_batch(1, intv, maxit, tol, vinit, pol, nevOut, lamo, Wo, reso, fstats);
  const double t_all = evsl_timer();
  const int n = c.n;
  maxit = std::min(n, maxit);
  const double bar = pol->bar, gamB = pol->gam;
  if (check_intv(intv) < 0) {
    *nevOut = 0;
    *lamo = nullptr;
    *Wo = nullptr;
    *reso = nullptr;
    return 0;
  }
  const double aa = intv[0], bb = intv[1];
  if (fstats)
    std::fprintf(fstats, " ** Cheb Poly of deg = %d, gam = %.15e, bar: %.15e\n", pol->deg,
                 gamB, bar);
  if (gamB > 1.0 || gamB < -1.0) {
    std::fprintf(stdout, "gamB error %.15e\n", gamB);
    return -1;
  }
  const int prec = c.filter_prec;
  LanWork L;
  const size_t nl = n;
  EVSL_HIP(L.V.alloc(nl * (maxit + 1) * sizeof(double)));
  L.nch = evsl_dev::gemv_t_chunks(n);
  EVSL_HIP(L.gt_part.alloc((size_t)L.nch * (maxit + 1) * sizeof(double)));
  EVSL_HIP(L.h.alloc((size_t)(maxit + 1) * sizeof(double)));
  EVSL_HIP(L.vrand.alloc(nl * sizeof(double)));
  double* V = L.V.as<double>();
  std::vector<double> dT(maxit), eT(maxit), EvalT(maxit);
  DevBuf vtmp;
  const double* v0;
  if (int e = as_device_vector(vinit, n, vtmp, &v0)) return e;
  EventClock clk;
  // filter the start vector and normalise it (FILTER_VINIT)
  clk.mark(0);
  if (int e = cheb_av(pol, v0, V, L.cw, prec)) return e;
  clk.mark(1);
  double nv;
  if (int e = nrm2_dev(n, V, &nv)) return e;
  c.astats.t_filter += clk.secs(0, 1);
  EVSL_HIP(evsl_dev::scal(n, 1.0 / nv, V, c.stream));
  double wn = 0.0, beta = 0.0, tr0 = 0.0;
  int nwn = 0, kdim = 0, k;
  const int Ntest = 30, cycle = 20;
  for (k = 0; k < maxit; ++k) {
    double* v = V + (size_t)k * nl;
    double* vnew = v + nl;
    const double* zold = k > 0 ? v - nl : nullptr;
    clk.mark(0);
    if (int e = cheb_av(pol, v, vnew, L.cw, prec)) return e;
    clk.mark(1);
    // vnew -= beta*zold; alpha = v.vnew; vnew -= alpha*v; ||vnew||^2
    if (zold)
      EVSL_HIP(evsl_dev::axpy_dot(n, -beta, zold, vnew, v, red_part(), c.stream));
    else
      EVSL_HIP(evsl_dev::dot_partial(n, vnew, v, red_part(), c.stream));
    EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(0), c.stream));
    EVSL_HIP(evsl_dev::axpy_dev_nrm(n, -1.0, red_scalar(0), v, vnew, red_part(), c.stream));
    EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(1), c.stream));
    double sc[2];
    if (int e = read_scalars(red_scalar(0), 2, sc)) return e;
    const double alpha = sc[0];
    dT[k] = alpha;
    wn += std::fabs(alpha);
    if (int e = cgs_dgks(n, k + 1, V, vnew, sc[1], &beta, L)) return e;
    clk.mark(2);
    wn += 2.0 * beta;
    nwn += 3;
    if (beta * nwn < kOrthTol * wn) {  // lucky breakdown (SRC/cheblanNr.c:225-256)
      if (fstats) std::fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", k, beta);
      rand_double_device(n, L.vrand.as<double>());
      if (int e = cheb_av(pol, L.vrand.as<double>(), vnew, L.cw, prec)) return e;
      double nn;
      if (int e = nrm2_dev(n, vnew, &nn)) return e;
      if (int e = cgs_dgks(n, k + 1, V, vnew, nn * nn, &beta, L)) return e;
      EVSL_HIP(evsl_dev::scal(n, 1.0 / beta, vnew, c.stream));
      beta = 0.0;
    } else {
      EVSL_HIP(evsl_dev::scal(n, 1.0 / beta, vnew, c.stream));
    }
    eT[k] = beta;
    c.astats.t_filter += clk.secs(0, 1);
    c.astats.t_reorth += clk.secs(1, 2);
    c.astats.n_lanczos++;
    if ((k < Ntest || (k - Ntest) % cycle != 0) && k != maxit - 1) continue;
    // restricted-trace convergence test (SRC/cheblanNr.c:269-309)
    kdim = k + 1;
    const double te = evsl_timer();
    std::vector<double> e(kdim + 1, 0.0);
    std::memcpy(EvalT.data(), dT.data(), kdim * sizeof(double));
    std::memcpy(e.data(), eT.data(), (kdim - 1) * sizeof(double));
    if (tridiag_ql(kdim, EvalT.data(), e.data(), nullptr)) return EVSL_ERR_EIG;
    c.astats.t_eig += evsl_timer() - te;
    double tr1 = 0.0;
    int nconv = 0;
    for (int i = 0; i < kdim; ++i)
      if (EvalT[i] + kEpsM >= bar) {
        tr1 += EvalT[i];
        ++nconv;
      }
    if (fstats) std::fprintf(fstats, "k %4d: nconv %4d  tr1 %21.15e\n", k, nconv, tr1);
    if (std::fabs(tr1 - tr0) < 2e-12 || std::fabs(tr1) + std::fabs(tr0) < 1e-10) break;
    tr0 = tr1;
  }
  if (k >= maxit && fstats)
    std::fprintf(fstats, "The max number of iterations [%d] has been reached. The eigenvalues "
                         "computed may not have converged\n", maxit);
  const double tr = evsl_timer();
  int err = ritz_extract(n, kdim, dT.data(), eT.data(), V, bar, aa, bb, tol, nevOut, lamo, Wo,
                         reso, maxit);
  const double tnow = evsl_timer();
  c.astats.t_ritz += tnow - tr;
  evslstat.t_ritz = tnow - tr;
  evslstat.t_iter = tnow - t_all;
  return err;
}

// Lanczos with full (modified Gram-Schmidt) re-orthogonalization giving
// spectrum bounds (SRC/lanbounds.c:21-144).  v may be a host or device vector.
int LanBounds(int msteps, double* vin, double* lmin, double* lmax) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (c.op == 0) return EVSL_ERR_NOMAT;
  const int n = c.n;
  const size_t nl = n;
  msteps = std::min(n, msteps);
  std::vector<double> alp(msteps), bet(msteps);
  DevBuf Vb, vtmp;
  EVSL_HIP(Vb.alloc((msteps + 1) * nl * sizeof(double)));
  double* V = Vb.as<double>();
  const double* v;
  if (int e = as_device_vector(vin, n, vtmp, &v)) return e;
  double nv;
  if (int e = nrm2_dev(n, v, &nv)) return e;
  EVSL_HIP(evsl_dev::scal_copy(n, 1.0 / nv, v, V, c.stream));
  double wn = 0.0;
  double* part = red_part();
  for (int j = 0; j < msteps; ++j) {
    double* vj = V + j * nl;
    double* vn = vj + nl;
    if (int e = apply_A(vj, vn)) return e;
    if (j)
      EVSL_HIP(evsl_dev::axpy_dot(n, -bet[j - 1], vj - nl, vn, vj, part, c.stream));
    else
      EVSL_HIP(evsl_dev::dot_partial(n, vn, vj, part, c.stream));
    EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(0), c.stream));
    // vn -= alp vj, then modified Gram-Schmidt against V(:,0..j): each step
    // subtracts the projection computed by the previous launch
    EVSL_HIP(evsl_dev::axpy_dev_dot(n, -1.0, red_scalar(0), vj, vn, V, part, c.stream));
    for (int i = 0; i <= j; ++i) {
      EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(1), c.stream));
      const double* next = (i < j) ? V + (i + 1) * nl : vn;
      EVSL_HIP(evsl_dev::axpy_dev_dot(n, -1.0, red_scalar(1), V + i * nl, vn, next, part,
                                      c.stream));
    }
    EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(2), c.stream));
    double a_b[3];
    if (int e = read_scalars(red_scalar(0), 3, a_b)) return e;
    alp[j] = a_b[0];
    bet[j] = a_b[2];
    wn += alp[j] * alp[j];
    if (bet[j] * (j + 1) < kOrthTol * wn) {
      std::fprintf(stdout, "lanbounds: lucky break, j=%d, beta=%e, break\n", j, bet[j]);
      msteps = j + 1;
      break;
    }
    wn += 2.0 * bet[j];
    bet[j] = std::sqrt(bet[j]);
    EVSL_HIP(evsl_dev::scal(n, 1.0 / bet[j], vn, c.stream));
  }
  const double bottomBeta = bet[msteps - 1];
  std::vector<double> S((size_t)msteps * msteps), ritz(msteps);
  if (tridiag_eig_host(msteps, alp.data(), bet.data(), ritz.data(), S.data()))
    return EVSL_ERR_EIG;
  *lmin = ritz[0] - std::fabs(bottomBeta * S[msteps - 1]);
  *lmax = ritz[msteps - 1] + std::fabs(bottomBeta * S[(size_t)msteps * msteps - 1]);
  return 0;
}

// Lanczos DOS: nvec independent Lanczos runs with full re-orthogonalization,
// Gauss quadrature nodes/weights smoothed by Gaussians, Simpson integral
// (SRC/landos.c:37-196).  Random vectors are Box-Muller draws from libc
// rand() as in the reference.
int LanDos(const int nvec, int msteps, int npts, double* xdos, double* ydos, double* neig,
           const double* const intv) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (c.op == 0) return EVSL_ERR_NOMAT;
  const int n = c.n;
  const size_t nl = n;
  const size_t ml0 = msteps;
  std::vector<double> y(npts, 0.0), S(ml0 * ml0), gamma2(msteps), ritz(msteps);
  std::vector<int> ind(npts);
  const double lm = intv[2], lM = intv[3];
  const double tolBdwn = 1.e-13 * (std::fabs(lM) + std::fabs(lm));
  const double aa = std::max(intv[0], intv[2]), bb = std::min(intv[1], intv[3]);
  const double kappa = 1.25;
  const int M = std::min(msteps, 30);
  const double H = (lM - lm) / (M - 1);
  const double sigma = H / std::sqrt(8 * std::log(kappa));
  const double sigma2 = 2 * sigma * sigma;
  const double width = sigma * std::sqrt(-2.0 * std::log(1e-08));
  linspace(aa, bb, npts, xdos);
  std::vector<double> alp(msteps), bet(msteps), hv(nl);
  DevBuf Vb, vb;
  EVSL_HIP(Vb.alloc((msteps + 1) * nl * sizeof(double)));
  EVSL_HIP(vb.alloc(nl * sizeof(double)));
  double* V = Vb.as<double>();
  double* part = red_part();
  for (int m = 0; m < nvec; ++m) {
    randn_double(n, hv.data());
    EVSL_HIP(hipMemcpy(vb.p, hv.data(), nl * sizeof(double), hipMemcpyHostToDevice));
    double nv;
    if (int e = nrm2_dev(n, vb.as<double>(), &nv)) return e;
    EVSL_HIP(evsl_dev::scal_copy(n, 1.0 / nv, vb.as<double>(), V, c.stream));
    double wn = 0.0;
    for (int j = 0; j < msteps; ++j) {
      double* vj = V + j * nl;
      double* vn = vj + nl;
      if (int e = apply_A(vj, vn)) return e;
      if (j)
        EVSL_HIP(evsl_dev::axpy_dot(n, -bet[j - 1], vj - nl, vn, vj, part, c.stream));
      else
        EVSL_HIP(evsl_dev::dot_partial(n, vn, vj, part, c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(0), c.stream));
      EVSL_HIP(evsl_dev::axpy_dev_dot(n, -1.0, red_scalar(0), vj, vn, V, part, c.stream));
      for (int i = 0; i <= j; ++i) {
        EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(1), c.stream));
        const double* next = (i < j) ? V + (i + 1) * nl : vn;
        EVSL_HIP(evsl_dev::axpy_dev_dot(n, -1.0, red_scalar(1), V + i * nl, vn, next, part,
                                        c.stream));
      }
      EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(2), c.stream));
      double a_b[3];
      if (int e = read_scalars(red_scalar(0), 3, a_b)) return e;
      alp[j] = a_b[0];
      bet[j] = a_b[2];
      wn += alp[j] * alp[j];
      if (bet[j] * (j + 1) < kOrthTol * wn) {
        std::fprintf(stdout, "lanbounds: lucky break, j=%d, beta=%e, break\n", j, bet[j]);
        msteps = j + 1;
        break;
      }
      if (bet[j] > tolBdwn) {
        wn += 2.0 * bet[j];
        bet[j] = std::sqrt(bet[j]);
        EVSL_HIP(evsl_dev::scal(n, 1.0 / bet[j], vn, c.stream));
      } else {
        // the reference draws the replacement into v, which is then unused
        // (SRC/landos.c:130-144); the stream of rand() calls is kept the same
        randn_double(n, hv.data());
        EVSL_HIP(evsl_dev::dot_partial(n, vn, V, part, c.stream));
        for (int i = 0; i <= j; ++i) {
          EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(1), c.stream));
          const double* next = (i < j) ? V + (i + 1) * nl : vn;
          EVSL_HIP(evsl_dev::axpy_dev_dot(n, -1.0, red_scalar(1), V + i * nl, vn, next, part,
                                          c.stream));
        }
        EVSL_HIP(evsl_dev::sum_scalar(part, evsl_dev::kRedBlocks, red_scalar(2), c.stream));
        double b2;
        if (int e = read_scalars(red_scalar(2), 1, &b2)) return e;
        wn += 2.0 * b2;
        const double bj = std::sqrt(b2);
        EVSL_HIP(evsl_dev::scal(n, 1.0 / bj, vn, c.stream));
        bet[j] = 0;
      }
    }
    if (tridiag_eig_host(msteps, alp.data(), bet.data(), ritz.data(), S.data()))
      return EVSL_ERR_EIG;
    for (int i = 0; i < msteps; ++i) gamma2[i] = S[i * ml0] * S[i * ml0];
    for (int i = 0; i < msteps; ++i) {
      const double t = ritz[i];
      int np = 0;
      for (int j = 0; j < npts; ++j)
        if (std::fabs(xdos[j] - t) < width) ind[np++] = j;
      for (int j = 0; j < np; ++j)
        y[ind[j]] += gamma2[i] * std::exp(-((xdos[ind[j]] - t) * (xdos[ind[j]] - t)) / sigma2);
    }
  }
  const double scaling = 1.0 / (nvec * std::sqrt(sigma2 * EVSL_PI));
  for (int j = 0; j < npts; ++j) y[j] *= scaling;
  std::memcpy(ydos, y.data(), npts * sizeof(double));
  // in-place Simpson integration (SRC/simpson.c:25-50)
  {
    double tc = 0.0, ti, tm, ysav = y[0];
    int i = 1;
    y[0] = 0;
    while (i < npts - 1) {
      ti = (xdos[i] - xdos[i - 1]) * (ysav + y[i]) / 2.0;
      tm = tc + ti;
      ti = (xdos[i + 1] - xdos[i - 1]) * (ysav + 4 * y[i] + y[i + 1]) / 6.0;
      tc += ti;
      y[i] = tm;
      ysav = y[i + 1];
      y[i + 1] = tc;
      i += 2;
    }
    if (i == npts - 1) y[i] = tc + (xdos[i] - xdos[i - 1]) * (y[i] + ysav) / 2.0;
  }
  *neig = y[npts - 1] * n;
  return 0;
}

// Thick-restart filtered Lanczos (SRC/cheblanTr.c:59-707).  The basis V
// (n x (lanm+1)), the locked block Y and the Ritz block live in HBM; the
// arrow-head matrix T and its dense eigensolve stay on the host.  The filter
// always runs in FP64: restarts lock Ritz pairs on their residual with A, so
// an FP32-filtered basis would never lock pairs below its rounding level.
int ChebLanTr(int lanm, int nev, double* intv, int maxit, double tol, double* vinit,
              polparams* pol, int* nev2, double** vals, double** W, double** resW,
              FILE* fstats) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (c.op == 0) return EVSL_ERR_NOMAT;
  const double t_all = evsl_timer();
  const int n = c.n;
  const size_t nl = n;
  lanm = std::min(lanm, n);
  const int lanm1 = lanm + 1;
  const size_t L1 = lanm1;
  if (lanm == n) maxit = std::min(maxit, n);
  const double nevInc = 0.2;
  const int cycle = 50, min_inner_step = 5;
  if (check_intv(intv) < 0) {
    *nev2 = 0;
    *vals = nullptr;
    *W = nullptr;
    *resW = nullptr;
    return 0;
  }
  const double aa = intv[0], bb = intv[1];
  const double gamB = pol->gam, bar = pol->bar;
  if (gamB > 1.0 || gamB < -1.0) {
    if (fstats) std::fprintf(fstats, "gamB error %.15e\n", gamB);
    return -1;
  }
  if (lanm < 1) {
    set_error_msg("ChebLanTr: lanm must be positive");
    return EVSL_ERR_ARG;
  }
  if (fstats) std::fprintf(fstats, " Cheb-LanTR, dim %d  cycle %d  \n", lanm, cycle);
  const int prec = EVSL_PREC_FP64;
  int ycap = std::max(nev, 1);
  LanWork L;
  L.nch = evsl_dev::gemv_t_chunks(n);
  int gcap = 0;  // columns the Gram-Schmidt scratch holds
  auto ensure_gs = [&](int cols) -> int {
    if (cols <= gcap) return 0;
    gcap = cols;
    EVSL_HIP(L.gt_part.alloc((size_t)L.nch * (gcap + 1) * sizeof(double)));
    EVSL_HIP(L.h.alloc((size_t)(gcap + 1) * sizeof(double)));
    return 0;
  };
  if (int e = ensure_gs(std::max(lanm1, ycap))) return e;
  EVSL_HIP(L.V.alloc(nl * L1 * sizeof(double)));
  EVSL_HIP(L.vrand.alloc(nl * sizeof(double)));
  DevBuf Yd, Rvec, Wblk, Ed, part, s2d, tvd, r2d;
  EVSL_HIP(Yd.alloc(nl * ycap * sizeof(double)));
  EVSL_HIP(Rvec.alloc(nl * lanm * sizeof(double)));
  constexpr int kChunk = 64;  // Ritz columns per A-application block
  EVSL_HIP(Wblk.alloc(nl * std::min(kChunk, lanm) * sizeof(double)));
  const int nch8 = (n + 8191) / 8192;
  EVSL_HIP(part.alloc((size_t)nch8 * lanm * sizeof(double)));
  EVSL_HIP(s2d.alloc(lanm * sizeof(double)));
  EVSL_HIP(tvd.alloc(lanm * sizeof(double)));
  EVSL_HIP(r2d.alloc(lanm * sizeof(double)));
  double* V = L.V.as<double>();
  std::vector<double> T(L1 * L1, 0.0), EvecT(L1 * L1), Rval(lanm), resi(lanm), s(lanm);
  std::vector<double> Lam(ycap), res(ycap);
  DevBuf vtmp;
  const double* v0;
  if (int e = as_device_vector(vinit, n, vtmp, &v0)) return e;
  EventClock clk;
  int lock = 0, trlen = 0, prtrlen = -1, it = 0;
  // filter the start vector and normalise it (FILTER_VINIT)
  if (int e = cheb_av(pol, v0, V, L.cw, prec)) return e;
  double nv;
  if (int e = nrm2_dev(n, V, &nv)) return e;
  EVSL_HIP(evsl_dev::scal(n, 1.0 / nv, V, c.stream));
  // deflation against the locked block: v -= Y Y^T v (two CGS passes at most)
  auto deflate = [&](double* vnew) -> int {
    if (lock == 0) return 0;
    double nn, dummy;
    if (int e = nrm2_dev(n, vnew, &nn)) return e;
    return cgs_dgks(n, lock, Yd.as<double>(), vnew, nn * nn, &dummy, L);
  };
  // lucky breakdown: a filtered random vector orthogonalised against Y and V(:,0:k)
  auto breakdown = [&](int k, double* vnew, double* beta) -> int {
    rand_double_device(n, L.vrand.as<double>());
    if (int e = cheb_av(pol, L.vrand.as<double>(), vnew, L.cw, prec)) return e;
    if (int e = deflate(vnew)) return e;
    double nn;
    if (int e = nrm2_dev(n, vnew, &nn)) return e;
    if (int e = cgs_dgks(n, k, V, vnew, nn * nn, beta, L)) return e;
    EVSL_HIP(evsl_dev::scal(n, 1.0 / *beta, vnew, c.stream));
    *beta = 0.0;
    return 0;
  };
  while (it < maxit) {
    double wn = 0.0, beta = 0.0;
    int nwn = 0;
    int k = trlen > 0 ? trlen + 1 : 0;
    if (k + min_inner_step > lanm1) {  // the reference exits here
      if (fstats) std::fprintf(fstats, "Krylov dim too small for this problem. Try a larger dim\n");
      set_error_msg("ChebLanTr: Krylov dimension too small for the restart set");
      return EVSL_ERR_ARG;
    }
    if (trlen > 0) {  // thick-restart special step (SRC/cheblanTr.c:224-319)
      const int k1 = k - 1;
      double* v = V + (size_t)k1 * nl;
      double* vnew = v + nl;
      clk.mark(0);
      if (int e = cheb_av(pol, v, vnew, L.cw, prec)) return e;
      clk.mark(1);
      c.astats.t_filter += clk.secs(0, 1);
      if (int e = deflate(vnew)) return e;
      for (int i = 0; i < trlen; ++i) {
        T[i * L1 + i] = Rval[i];
        wn += std::fabs(Rval[i]);
      }
      EVSL_HIP(evsl_dev::dot_partial(n, v, vnew, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(0), c.stream));
      if (int e = read_scalars(red_scalar(0), 1, &s[k1])) return e;
      // vnew -= V(:, 0:k) s(0:k), with ||vnew||^2
      EVSL_HIP(hipMemcpy(L.h.p, s.data(), (size_t)k * sizeof(double), hipMemcpyHostToDevice));
      const int nb = (n + evsl_dev::kWG - 1) / evsl_dev::kWG;
      EVSL_HIP(evsl_dev::gemv_n_sub_nrm(n, k, V, n, L.h.as<double>(), vnew, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), nb, red_scalar(1), c.stream));
      for (int i = 0; i < k1; ++i) {
        T[trlen * L1 + i] = s[i];
        T[i * L1 + trlen] = s[i];
        wn += 2.0 * std::fabs(s[i]);
      }
      T[trlen * L1 + trlen] = s[k1];
      wn += std::fabs(s[k1]);
      double b2;
      if (int e = read_scalars(red_scalar(1), 1, &b2)) return e;
      beta = std::sqrt(b2);
      wn += 2.0 * beta;
      nwn += 3 * k;
      if (beta * nwn < kOrthTol * wn) {
        if (fstats) std::fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", k, beta);
        if (int e = breakdown(k, vnew, &beta)) return e;
      } else {
        EVSL_HIP(evsl_dev::scal(n, 1.0 / beta, vnew, c.stream));
      }
      T[k1 * L1 + k] = beta;
      T[k * L1 + k1] = beta;
    }
    const int Ntest = std::max(20, nev - lock + 10);
    int last_jl = 0;
    double last_tr = 0.0;
    const double* zold = k > 0 ? V + (size_t)(k - 1) * nl : nullptr;
    while (k < lanm && it < maxit) {  // Lanczos inner loop (SRC/cheblanTr.c:331-503)
      ++k;
      double* v = V + (size_t)(k - 1) * nl;
      double* vnew = v + nl;
      clk.mark(0);
      if (int e = cheb_av(pol, v, vnew, L.cw, prec)) return e;
      clk.mark(1);
      ++it;
      if (int e = deflate(vnew)) return e;
      // vnew -= beta*zold; alpha = v.vnew; vnew -= alpha*v; ||vnew||^2
      if (zold)
        EVSL_HIP(evsl_dev::axpy_dot(n, -beta, zold, vnew, v, red_part(), c.stream));
      else
        EVSL_HIP(evsl_dev::dot_partial(n, vnew, v, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(0), c.stream));
      EVSL_HIP(evsl_dev::axpy_dev_nrm(n, -1.0, red_scalar(0), v, vnew, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(1), c.stream));
      double sc[2];
      if (int e = read_scalars(red_scalar(0), 2, sc)) return e;
      const double alpha = sc[0];
      T[(k - 1) * L1 + (k - 1)] = alpha;
      wn += std::fabs(alpha);
      if (int e = cgs_dgks(n, k, V, vnew, sc[1], &beta, L)) return e;
      clk.mark(2);
      wn += 2.0 * beta;
      nwn += 3;
      zold = v;
      if (beta * nwn < kOrthTol * wn) {
        if (fstats) std::fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", it, beta);
        if (int e = breakdown(k, vnew, &beta)) return e;
      } else {
        EVSL_HIP(evsl_dev::scal(n, 1.0 / beta, vnew, c.stream));
      }
      T[k * L1 + (k - 1)] = beta;
      T[(k - 1) * L1 + k] = beta;
      c.astats.t_filter += clk.secs(0, 1);
      c.astats.t_reorth += clk.secs(1, 2);
      c.astats.n_lanczos++;
      const int k1 = k - trlen - Ntest;
      if (((k1 >= 0) && (k1 % cycle == 0)) || (k == lanm) || it == maxit) {
        const double te = evsl_timer();
        if (int e = sym_eig_host(k, T.data(), lanm1, EvecT.data(), lanm1, Rval.data())) return e;
        c.astats.t_eig += evsl_timer() - te;
        if (k == lanm || it == maxit) break;
        int jl = 0;
        double tr = 0.0;
        for (int i = 0; i < k; ++i)
          if (Rval[i] + kEpsM >= bar) {
            ++jl;
            tr += Rval[i];
          }
        if (std::fabs(tr - last_tr) <= 2e-12 && jl == last_jl) {
          if (fstats)
            std::fprintf(fstats,
                         "break: [it %d, k %d]: last_tr %.15e, tr %.15e, last_jl %d  jl %d\n", it,
                         k, last_tr, tr, last_jl, jl);
          break;
        }
        if (fstats)
          std::fprintf(fstats, "[it %d, k %d] testing: last_tr %.15e, tr %.15e, last_jl %d  jl %d\n",
                       it, k, last_tr, tr, last_jl, jl);
        last_tr = tr;
        last_jl = jl;
      }
    }
    const double tr0 = evsl_timer();
    // pass 1: Ritz values of p(A) at or above bar, moved to the front
    int jl = 0;
    for (int i = 0; i < k; ++i) {
      if (Rval[i] + kEpsM >= bar) {
        if (i != jl) {
          std::memcpy(&EvecT[jl * L1], &EvecT[i * L1], (size_t)k * sizeof(double));
          Rval[jl] = Rval[i];
        }
        resi[jl] = std::fabs(beta * EvecT[i * L1 + (k - 1)]);
        ++jl;
      }
    }
    // Rvec(:, 0:jl) = V(:, 0:k) EvecT(0:k, 0:jl)
    double* R = Rvec.as<double>();
    std::vector<double> hs2(jl), ht(jl), hr2(jl);
    if (jl > 0) {
      std::vector<double> Eh((size_t)k * jl);
      for (int j = 0; j < jl; ++j)
        std::memcpy(&Eh[(size_t)j * k], &EvecT[j * L1], (size_t)k * sizeof(double));
      EVSL_HIP(Ed.alloc(Eh.size() * sizeof(double)));
      EVSL_HIP(hipMemcpy(Ed.p, Eh.data(), Eh.size() * sizeof(double), hipMemcpyHostToDevice));
      EVSL_HIP(evsl_dev::gemm_nn(n, jl, k, V, n, Ed.as<double>(), k, R, n, c.stream));
      EVSL_HIP(evsl_dev::col_dot(n, jl, R, n, R, n, part.as<double>(), s2d.as<double>(), c.stream));
      EVSL_HIP(hipMemcpy(hs2.data(), s2d.p, jl * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < jl; ++i)
        if (hs2[i] == 0.0) return 2;  // zero Ritz vector (SRC/cheblanTr.c:562-564)
      EVSL_HIP(evsl_dev::col_scale_rsqrt(n, jl, R, n, s2d.as<double>(), c.stream));
      // Rayleigh quotients and residuals, kChunk columns at a time
      double* Wb = Wblk.as<double>();
      for (int j0 = 0; j0 < jl; j0 += kChunk) {
        const int m = std::min(kChunk, jl - j0);
        for (int j = 0; j < m; ++j)
          if (int e = apply_A(R + (size_t)(j0 + j) * nl, Wb + (size_t)j * nl)) return e;
        EVSL_HIP(evsl_dev::col_dot(n, m, Wb, n, R + (size_t)j0 * nl, n, part.as<double>(),
                                   tvd.as<double>() + j0, c.stream));
        EVSL_HIP(evsl_dev::col_res2(n, m, Wb, n, R + (size_t)j0 * nl, n, tvd.as<double>() + j0,
                                    part.as<double>(), r2d.as<double>() + j0, c.stream));
      }
      EVSL_HIP(hipMemcpy(ht.data(), tvd.p, jl * sizeof(double), hipMemcpyDeviceToHost));
      EVSL_HIP(hipMemcpy(hr2.data(), r2d.p, jl * sizeof(double), hipMemcpyDeviceToHost));
    }
    // pass 2: Ritz values of A in [a, b] are locked or kept for the restart
    int ll = 0;
    prtrlen = trlen;
    trlen = 0;
    for (int i = 0; i < jl; ++i) {
      const double t3 = ht[i];
      if (!(t3 >= aa - kEpsM && t3 <= bb + kEpsM)) continue;
      ++ll;
      const double res0 = std::sqrt(std::max(hr2[i], 0.0));
      const double* y = R + (size_t)i * nl;
      if (res0 < tol) {
        if (lock >= nev) {  // grow the locked block by 1 + 20%
          nev += 1 + (int)(nev * nevInc);
          if (fstats) std::fprintf(fstats, "-- More eigval found: realloc space for %d evs\n", nev);
        }
        if (lock >= ycap) {
          const int ncap = std::max(nev, lock + 1);
          DevBuf Y2;
          EVSL_HIP(Y2.alloc(nl * ncap * sizeof(double)));
          EVSL_HIP(hipMemcpyAsync(Y2.p, Yd.p, nl * lock * sizeof(double),
                                  hipMemcpyDeviceToDevice, c.stream));
          std::swap(Yd.p, Y2.p);
          std::swap(Yd.bytes, Y2.bytes);
          ycap = ncap;
          Lam.resize(ycap);
          res.resize(ycap);
          if (int e = ensure_gs(std::max(lanm1, ycap))) return e;
        }
        EVSL_HIP(hipMemcpyAsync(Yd.as<double>() + (size_t)lock * nl, y, nl * sizeof(double),
                                hipMemcpyDeviceToDevice, c.stream));
        Lam[lock] = t3;
        res[lock] = res0;
        ++lock;
      } else {
        Rval[trlen] = Rval[i];
        EVSL_HIP(hipMemcpyAsync(V + (size_t)trlen * nl, y, nl * sizeof(double),
                                hipMemcpyDeviceToDevice, c.stream));
        s[trlen] = beta * EvecT[i * L1 + (k - 1)];
        ++trlen;
      }
    }
    c.astats.t_ritz += evsl_timer() - tr0;
    if (fstats)
      std::fprintf(fstats, "it %4d:  k %3d, jl %3d, ll %3d, lock %3d, trlen %3d\n", it, k, jl, ll,
                   lock, trlen);
    if (prtrlen == 0 && ll == 0) {
      if (fstats)
        std::fprintf(fstats, "--------------------------------------------------\n"
                             " --> No ev.s left to be computed\n");
      break;
    }
    if (it == maxit) {
      if (fstats)
        std::fprintf(fstats, "--------------------------------------------------\n"
                             " --> Max its reached without convergence\n");
      break;
    }
    std::fill(T.begin(), T.end(), 0.0);
    if (trlen != k)
      EVSL_HIP(hipMemcpyAsync(V + (size_t)trlen * nl, V + (size_t)k * nl, nl * sizeof(double),
                              hipMemcpyDeviceToDevice, c.stream));
  }
  if (fstats)
    std::fprintf(fstats, "     Number of evals found = %d\n"
                         "--------------------------------------------------\n", lock);
  EVSL_HIP(hipStreamSynchronize(c.stream));
  // outputs: host Lam/res (evsl_Malloc), device Y (evsl_Malloc_device)
  double* lam_o = (double*)std::malloc((size_t)std::max(lock, 1) * sizeof(double));
  double* res_o = (double*)std::malloc((size_t)std::max(lock, 1) * sizeof(double));
  double* Y_o = (double*)_evsl_Malloc_device(nl * std::max(lock, 1) * sizeof(double));
  if (!lam_o || !res_o || !Y_o) {
    std::free(lam_o);
    std::free(res_o);
    if (Y_o) _evsl_Free_device(Y_o);
    return EVSL_ERR_ALLOC;
  }
  if (lock > 0) {
    std::memcpy(lam_o, Lam.data(), lock * sizeof(double));
    std::memcpy(res_o, res.data(), lock * sizeof(double));
    EVSL_HIP(hipMemcpy(Y_o, Yd.p, nl * lock * sizeof(double), hipMemcpyDeviceToDevice));
  }
  *nev2 = lock;
  *vals = lam_o;
  *W = Y_o;
  *resW = res_o;
  evslstat.t_iter = evsl_timer() - t_all;
  return 0;
}


// Thick-restart Lanczos spectrum bounds (SRC/lanTrbounds.c:35-421): Lanczos
// on A with full reorthogonalization, restarted from the two extreme Ritz
// pairs until the bound corrections fall below tol * (|lmin| + |lmax|).
// bndtype <= 1: simple bounds |beta s|; otherwise Kato-Temple.
int LanTrbounds(int lanm, int maxit, double tol, double* vinit, int bndtype, double* lammin,
                double* lammax, FILE* fstats) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (c.op == 0) return EVSL_ERR_NOMAT;
  const int n = c.n;
  const size_t nl = n;
  lanm = std::min(lanm, n);
  const int lanm1 = lanm + 1;
  const size_t L1 = lanm1;
  if (lanm == n) maxit = std::min(maxit, n);
  const int min_inner_step = 5;
  if (fstats) std::fprintf(fstats, " LanTR for bounds: dim %d, maxits %d\n", lanm, maxit);
  LanWork L;
  L.nch = evsl_dev::gemv_t_chunks(n);
  EVSL_HIP(L.gt_part.alloc((size_t)L.nch * (lanm1 + 1) * sizeof(double)));
  EVSL_HIP(L.h.alloc((size_t)(lanm1 + 1) * sizeof(double)));
  EVSL_HIP(L.V.alloc(nl * L1 * sizeof(double)));
  DevBuf Rvec, Ed;
  EVSL_HIP(Rvec.alloc(2 * nl * sizeof(double)));
  double* V = L.V.as<double>();
  std::vector<double> T(L1 * L1, 0.0), EvecT(L1 * L1), Rval(std::max(lanm, 2));
  double s[3] = {0.0, 0.0, 0.0};
  DevBuf vtmp;
  const double* v0;
  if (int e = as_device_vector(vinit, n, vtmp, &v0)) return e;
  EVSL_HIP(hipMemcpyAsync(V, v0, nl * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  double nv;
  if (int e = nrm2_dev(n, V, &nv)) return e;
  EVSL_HIP(evsl_dev::scal(n, 1.0 / nv, V, c.stream));
  const int nb = (n + evsl_dev::kWG - 1) / evsl_dev::kWG;
  // lucky breakdown: a random vector orthogonalised against V(:, 0:k)
  auto breakdown = [&](int k, double* vnew, double* beta) -> int {
    rand_double_device(n, vnew);
    double nn;
    if (int e = nrm2_dev(n, vnew, &nn)) return e;
    if (int e = cgs_dgks(n, k, V, vnew, nn * nn, beta, L)) return e;
    EVSL_HIP(evsl_dev::scal(n, 1.0 / *beta, vnew, c.stream));
    *beta = 0.0;
    return 0;
  };
  double lmin = 0.0, lmax = 0.0;
  int it = 0, trlen = 0;
  while (it < maxit) {
    double wn = 0.0, beta = 0.0;
    int nwn = 0;
    int k = trlen > 0 ? trlen + 1 : 0;
    if (k + min_inner_step > lanm1) {  // the reference exits here
      set_error_msg("LanTrbounds: Krylov dimension too small");
      return EVSL_ERR_ARG;
    }
    if (trlen > 0) {  // thick-restart special step (SRC/lanTrbounds.c:145-215)
      const int k1 = k - 1;
      double* v = V + (size_t)k1 * nl;
      double* vnew = v + nl;
      if (int e = apply_A(v, vnew)) return e;
      for (int i = 0; i < trlen; ++i) {
        T[i * L1 + i] = Rval[i];
        wn += std::fabs(Rval[i]);
      }
      EVSL_HIP(evsl_dev::dot_partial(n, v, vnew, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(0), c.stream));
      if (int e = read_scalars(red_scalar(0), 1, &s[k1])) return e;
      EVSL_HIP(hipMemcpy(L.h.p, s, (size_t)k * sizeof(double), hipMemcpyHostToDevice));
      EVSL_HIP(evsl_dev::gemv_n_sub_nrm(n, k, V, n, L.h.as<double>(), vnew, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), nb, red_scalar(1), c.stream));
      for (int i = 0; i < k1; ++i) {
        T[trlen * L1 + i] = s[i];
        T[i * L1 + trlen] = s[i];
        wn += 2.0 * std::fabs(s[i]);
      }
      T[trlen * L1 + trlen] = s[k1];
      wn += std::fabs(s[k1]);
      double b2;
      if (int e = read_scalars(red_scalar(1), 1, &b2)) return e;
      beta = std::sqrt(b2);
      wn += 2.0 * beta;
      nwn += 3 * k;
      if (beta * nwn < kOrthTol * wn) {
        if (int e = breakdown(k, vnew, &beta)) return e;
      } else {
        EVSL_HIP(evsl_dev::scal(n, 1.0 / beta, vnew, c.stream));
      }
      T[k1 * L1 + k] = beta;
      T[k * L1 + k1] = beta;
    }
    const double* zold = k > 0 ? V + (size_t)(k - 1) * nl : nullptr;
    while (k < lanm && it < maxit) {  // Lanczos inner loop (SRC/lanTrbounds.c:222-300)
      ++k;
      double* v = V + (size_t)(k - 1) * nl;
      double* vnew = v + nl;
      if (int e = apply_A(v, vnew)) return e;
      ++it;
      if (zold)
        EVSL_HIP(evsl_dev::axpy_dot(n, -beta, zold, vnew, v, red_part(), c.stream));
      else
        EVSL_HIP(evsl_dev::dot_partial(n, vnew, v, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(0), c.stream));
      EVSL_HIP(evsl_dev::axpy_dev_nrm(n, -1.0, red_scalar(0), v, vnew, red_part(), c.stream));
      EVSL_HIP(evsl_dev::sum_scalar(red_part(), evsl_dev::kRedBlocks, red_scalar(1), c.stream));
      double sc[2];
      if (int e = read_scalars(red_scalar(0), 2, sc)) return e;
      T[(k - 1) * L1 + (k - 1)] = sc[0];
      wn += std::fabs(sc[0]);
      if (int e = cgs_dgks(n, k, V, vnew, sc[1], &beta, L)) return e;
      wn += 2.0 * beta;
      nwn += 3;
      zold = v;
      if (beta * nwn < kOrthTol * wn) {
        if (fstats) std::fprintf(fstats, "it %4d: Lucky breakdown, beta = %.15e\n", it, beta);
        if (int e = breakdown(k, vnew, &beta)) return e;
      } else {
        EVSL_HIP(evsl_dev::scal(n, 1.0 / beta, vnew, c.stream));
      }
      T[k * L1 + (k - 1)] = beta;
      T[(k - 1) * L1 + k] = beta;
      c.astats.n_lanczos++;
    }
    if (k < 2) {
      set_error_msg("LanTrbounds: fewer than two Lanczos steps");
      return EVSL_ERR_ARG;
    }
    if (int e = sym_eig_host(k, T.data(), lanm1, EvecT.data(), lanm1, Rval.data())) return e;
    // spikes: beta times the last components of the extreme eigenvectors
    s[0] = beta * EvecT[k - 1];
    s[1] = beta * EvecT[(size_t)(k - 1) * L1 + (k - 1)];
    double t1, t2;
    if (bndtype <= 1) {
      t1 = std::fabs(s[0]);
      t2 = std::fabs(s[1]);
    } else {  // Kato-Temple
      t1 = 2.0 * s[0] * s[0] / (Rval[1] - Rval[0]);
      t2 = 2.0 * s[1] * s[1] / (Rval[k - 1] - Rval[k - 2]);
    }
    lmin = Rval[0] - t1;
    lmax = Rval[k - 1] + t2;
    // the two extreme Ritz vectors become the restart set
    std::vector<double> Eh((size_t)2 * k);
    std::memcpy(Eh.data(), &EvecT[0], (size_t)k * sizeof(double));
    std::memcpy(Eh.data() + k, &EvecT[(size_t)(k - 1) * L1], (size_t)k * sizeof(double));
    EVSL_HIP(Ed.alloc(Eh.size() * sizeof(double)));
    EVSL_HIP(hipMemcpy(Ed.p, Eh.data(), Eh.size() * sizeof(double), hipMemcpyHostToDevice));
    EVSL_HIP(evsl_dev::gemm_nn(n, 2, k, V, n, Ed.as<double>(), k, Rvec.as<double>(), n,
                               c.stream));
    trlen = 2;
    EVSL_HIP(hipMemcpyAsync(V, Rvec.p, 2 * nl * sizeof(double), hipMemcpyDeviceToDevice,
                            c.stream));
    Rval[1] = Rval[k - 1];
    if (fstats)
      std::fprintf(fstats, "it %4d, k %3d: ritz %.15e %.15e, t1,t2 %e %e, res %.15e %.15e\n", it,
                   k, Rval[0], Rval[1], t1, t2, std::fabs(s[0]), std::fabs(s[1]));
    if (t1 + t2 < tol * (std::fabs(lmin) + std::fabs(lmax))) break;
    std::fill(T.begin(), T.end(), 0.0);
    EVSL_HIP(hipMemcpyAsync(V + (size_t)trlen * nl, V + (size_t)k * nl, nl * sizeof(double),
                            hipMemcpyDeviceToDevice, c.stream));
  }
  EVSL_HIP(hipStreamSynchronize(c.stream));
  *lammin = lmin;
  *lammax = lmax;
  return 0;
}

}  // extern "C"
