// tridiag.cpp -- symmetric tridiagonal eigensolver for the Lanczos matrix T.
//
// Replaces LAPACK dstev (SymmTridEig, SRC/misc_la.c:33-73): implicit QL with
// a Wilkinson-type shift (the EISPACK tql1/tql2 iteration).  The O(k^2)
// scalar iteration stays on the host; when eigenvectors are wanted the plane
// rotations are recorded and replayed on the device (thread per row of Z),
// which moves the O(k^3) part of dsteqr onto the GPU next to the Lanczos
// basis that consumes it.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.h"

namespace evsl {

int tridiag_ql(int n, double* d, double* e, RotationLog* rec) {
  if (rec) rec->clear();
  if (n <= 0) return 0;
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) return EVSL_ERR_EIG;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      if (rec) {
        rec->top.push_back(m - 1);
        rec->off.push_back((long long)(rec->cs.size() / 2));
        rec->cnt.push_back(0);
      }
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (rec) {
          rec->cs.push_back(c);
          rec->cs.push_back(s);
          rec->cnt.back()++;
        }
      }
      if (rec && rec->cnt.back() == 0) {
        rec->top.pop_back();
        rec->off.pop_back();
        rec->cnt.pop_back();
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // ascending order; perm[j] = original position of the j-th smallest
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> ds(n);
  for (int j = 0; j < n; ++j) ds[j] = d[perm[j]];
  std::memcpy(d, ds.data(), n * sizeof(double));
  if (rec) rec->perm = perm;
  return 0;
}

int tridiag_eig_host(int n, const double* diag, const double* sdiag, double* eval, double* Z) {
  std::vector<double> e(n + 1, 0.0);
  std::memcpy(eval, diag, n * sizeof(double));
  if (n > 1) std::memcpy(e.data(), sdiag, (n - 1) * sizeof(double));
  if (!Z) return tridiag_ql(n, eval, e.data(), nullptr);
  RotationLog rec;
  int info = tridiag_ql(n, eval, e.data(), &rec);
  if (info) return info;
  // replay the rotations on the identity (host, row by row)
  std::vector<double> W((size_t)n * n, 0.0);  // row-major: W[r*n + col]
  for (int r = 0; r < n; ++r) W[(size_t)r * n + r] = 1.0;
  for (size_t sw = 0; sw < rec.top.size(); ++sw) {
    const int ihi = rec.top[sw], c = rec.cnt[sw];
    const double* rot = rec.cs.data() + 2 * rec.off[sw];
    for (int r = 0; r < n; ++r) {
      double* z = W.data() + (size_t)r * n;
      double carry = z[ihi + 1];
      for (int j = 0; j < c; ++j) {
        const int i = ihi - j;
        const double zi = z[i];
        z[i + 1] = rot[2 * j + 1] * zi + rot[2 * j] * carry;
        carry = rot[2 * j] * zi - rot[2 * j + 1] * carry;
      }
      z[ihi - c + 1] = carry;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int r = 0; r < n; ++r) Z[(size_t)j * n + r] = W[(size_t)r * n + rec.perm[j]];
  return 0;
}

int sym_eig_host(int n, const double* A, int lda, double* Q, int ldq, double* lam) {
  // Householder reduction W = H^T A H (H accumulated in Z), then QL on the
  // tridiagonal with rotations replayed on Z.
  std::vector<double> W((size_t)n * n), Z((size_t)n * n, 0.0), v(n), p(n), zv(n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) W[i + (size_t)j * n] = A[i + (size_t)j * lda];
  for (int i = 0; i < n; ++i) Z[i + (size_t)i * n] = 1.0;
  for (int k = 0; k + 2 < n; ++k) {
    const int len = n - k - 1;
    double* x = &W[(k + 1) + (size_t)k * n];
    double xn = 0.0;
    for (int i = 0; i < len; ++i) xn += x[i] * x[i];
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const double alpha = x[0] > 0 ? -xn : xn;
    for (int i = 0; i < len; ++i) v[i] = x[i];
    v[0] -= alpha;
    double vn = 0.0;
    for (int i = 0; i < len; ++i) vn += v[i] * v[i];
    vn = std::sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = 0; i < len; ++i) v[i] /= vn;
    // p = W v by columns (contiguous), K = v.p
    for (int i = 0; i < len; ++i) p[i] = 0.0;
    for (int j = 0; j < len; ++j) {
      const double* wc = &W[(k + 1) + (size_t)(k + 1 + j) * n];
      const double vj = v[j];
      for (int i = 0; i < len; ++i) p[i] += wc[i] * vj;
    }
    double K = 0.0;
    for (int i = 0; i < len; ++i) K += v[i] * p[i];
    for (int i = 0; i < len; ++i) p[i] = 2.0 * (p[i] - K * v[i]);
#pragma omp parallel for schedule(static) if (len > 192)
    for (int j = 0; j < len; ++j) {
      double* wc = &W[(k + 1) + (size_t)(k + 1 + j) * n];
      const double vj = v[j], pj = p[j];
      for (int i = 0; i < len; ++i) wc[i] -= v[i] * pj + p[i] * vj;
    }
    W[(k + 1) + (size_t)k * n] = alpha;
    W[k + (size_t)(k + 1) * n] = alpha;
    for (int i = 1; i < len; ++i) {
      W[(k + 1 + i) + (size_t)k * n] = 0.0;
      W[k + (size_t)(k + 1 + i) * n] = 0.0;
    }
    // Z(:, k+1:) -= 2 (Z(:, k+1:) v) v^T, column by column
    for (int r = 0; r < n; ++r) zv[r] = 0.0;
    for (int i = 0; i < len; ++i) {
      const double* zc = &Z[(size_t)(k + 1 + i) * n];
      const double vi = v[i];
      for (int r = 0; r < n; ++r) zv[r] += zc[r] * vi;
    }
#pragma omp parallel for schedule(static) if (len > 192)
    for (int i = 0; i < len; ++i) {
      double* zc = &Z[(size_t)(k + 1 + i) * n];
      const double vi2 = 2.0 * v[i];
      for (int r = 0; r < n; ++r) zc[r] -= zv[r] * vi2;
    }
  }
  std::vector<double> e(n + 1, 0.0);
  for (int i = 0; i < n; ++i) lam[i] = W[i + (size_t)i * n];
  for (int i = 0; i + 1 < n; ++i) e[i] = W[(i + 1) + (size_t)i * n];
  RotationLog rec;
  int info = tridiag_ql(n, lam, e.data(), &rec);
  if (info) return info;
  // replay rotations on the rows of Z (row-major copy for locality)
  std::vector<double> R((size_t)n * n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) R[(size_t)r * n + c] = Z[r + (size_t)c * n];
  // rows are independent: each applies every sweep in order
  const size_t nsw = rec.top.size();
#pragma omp parallel for schedule(static) if (n > 128)
  for (int r = 0; r < n; ++r) {
    double* z = R.data() + (size_t)r * n;
    for (size_t sw = 0; sw < nsw; ++sw) {
      const int ihi = rec.top[sw], cnt = rec.cnt[sw];
      const double* rot = rec.cs.data() + 2 * rec.off[sw];
      double carry = z[ihi + 1];
      for (int j = 0; j < cnt; ++j) {
        const int i = ihi - j;
        const double zi = z[i];
        z[i + 1] = rot[2 * j + 1] * zi + rot[2 * j] * carry;
        carry = rot[2 * j] * zi - rot[2 * j + 1] * carry;
      }
      z[ihi - cnt + 1] = carry;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int r = 0; r < n; ++r) Q[r + (size_t)j * ldq] = R[(size_t)r * n + rec.perm[j]];
  return 0;
}

int tridiag_rotations_to_device(int kdim, const RotationLog& rec, const std::vector<int>& sel,
                                double* Zsel_d) {
  Context& c = ctx();
  const int nsw = (int)rec.top.size();
  const size_t kk = (size_t)kdim * kdim;
  DevBuf Z, top, cnt, off, cs, idx;
  EVSL_HIP(Z.alloc(kk * sizeof(double)));
  EVSL_HIP(evsl_dev::set_identity(kdim, Z.as<double>(), c.stream));
  if (nsw > 0) {
    EVSL_HIP(top.alloc(nsw * sizeof(int)));
    EVSL_HIP(cnt.alloc(nsw * sizeof(int)));
    EVSL_HIP(off.alloc(nsw * sizeof(long long)));
    EVSL_HIP(cs.alloc(std::max<size_t>(rec.cs.size(), 2) * sizeof(double)));
    EVSL_HIP(hipMemcpy(top.p, rec.top.data(), nsw * sizeof(int), hipMemcpyHostToDevice));
    EVSL_HIP(hipMemcpy(cnt.p, rec.cnt.data(), nsw * sizeof(int), hipMemcpyHostToDevice));
    EVSL_HIP(hipMemcpy(off.p, rec.off.data(), nsw * sizeof(long long), hipMemcpyHostToDevice));
    EVSL_HIP(hipMemcpy(cs.p, rec.cs.data(), rec.cs.size() * sizeof(double),
                       hipMemcpyHostToDevice));
    EVSL_HIP(evsl_dev::apply_rotations(kdim, Z.as<double>(), nsw, top.as<int>(), cnt.as<int>(),
                                       off.as<long long>(), cs.as<double>(), c.stream));
  }
  const int m = (int)sel.size();
  if (m > 0) {
    std::vector<int> cols(m);
    for (int j = 0; j < m; ++j) cols[j] = rec.perm[sel[j]];
    EVSL_HIP(idx.alloc(m * sizeof(int)));
    EVSL_HIP(hipMemcpy(idx.p, cols.data(), m * sizeof(int), hipMemcpyHostToDevice));
    EVSL_HIP(evsl_dev::gather_cols(kdim, m, Z.as<double>(), kdim, idx.as<int>(), Zsel_d, kdim,
                                   c.stream));
  }
  EVSL_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

}  // namespace evsl
