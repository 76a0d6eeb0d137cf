// kpm_batch.cpp -- KPM densities of many matrices in one batched run.
//
// Each request is kpmdos (SRC/spslice.c:29-148) on its own matrix and
// interval: nvec seeded device random vectors in column blocks of up to 32,
// a three-term Chebyshev recurrence per block, and the moments
// mu_k ~ sum_m v_m^T T_k(A~) v_m.  The requests advance together: one grouped
// launch per degree runs the row blocks of every matrix (owner table ->
// matrix entry), so small matrices fill the device beside large ones, and
// one grouped reduction per degree sums each matrix's block partials in the
// order the one-matrix path uses.  In FP64 a request's moments equal
// kpmdos_ex on its matrix with the same seed bit for bit; EVSL_PREC_MIXED
// runs the recurrence in FP32 on (A - ctr I) * 2^-s with the moment dots
// accumulated in FP64.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"

namespace evsl {
namespace {

// one distinct matrix of the batch, resident for the call
struct BatchMatrix {
  const csrMat* host = nullptr;
  int n = 0, nnz = 0;
  bool full_diag = false;
  DevBuf ia, ja, a64;
  std::map<int, DevBuf> table;  // row-block table by lanes per row
  std::map<int, int> nblk;
  evsl_dev::Csr view() const {
    evsl_dev::Csr v;
    v.n = n;
    v.nnz = nnz;
    v.ia = ia.as<int>();
    v.ja = ja.as<int>();
    v.a64 = a64.as<double>();
    return v;
  }
  int ensure_table(int lanes) {
    if (table.count(lanes)) return 0;
    std::vector<int> rb = make_row_blocks(host->ia, n, evsl_dev::rows_cap_for(lanes));
    std::vector<int> rbp(2 * rb.size());
    for (size_t i = 0; i < rb.size(); ++i) {
      rbp[2 * i] = rb[i];
      rbp[2 * i + 1] = host->ia[rb[i]];
    }
    DevBuf& d = table[lanes];
    EVSL_HIP(d.alloc(rbp.size() * sizeof(int)));
    EVSL_HIP(hipMemcpy(d.p, rbp.data(), rbp.size() * sizeof(int), hipMemcpyHostToDevice));
    nblk[lanes] = (int)rb.size() - 1;
    return 0;
  }
};

// one request's device state
struct BatchRequest {
  BatchMatrix* mat = nullptr;
  const double* intv = nullptr;
  unsigned long long seed = 0;
  double ctr = 0, wid = 0, t0 = 0, t = 0, cc = 0;
  DevBuf a32;  // FP32 values (shifted and scaled when the diagonal is stored)
  DevBuf x0, ring, part, dots, scratch;
  std::vector<double> raw;
};

int upload(DevBuf& d, const void* src, size_t bytes) {
  EVSL_HIP(d.alloc(std::max<size_t>(bytes, 8)));
  if (bytes) EVSL_HIP(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace
}  // namespace evsl

using namespace evsl;

extern "C" {

int kpmdos_batch(int nreq, csrMat* const* mats, const double* intvs, int Mdeg, int damping,
                 int nvec, int vec_kind, const unsigned long long* seeds, int prec, double* mu,
                 double* ecnt) {
  if (int e = ensure_started()) return e;
  Context& c = ctx();
  if (nreq <= 0 || !mats || !intvs || !mu || !ecnt || Mdeg < 0 || nvec <= 0 ||
      (prec != EVSL_PREC_FP64 && prec != EVSL_PREC_MIXED)) {
    set_error_msg("kpmdos_batch: invalid arguments");
    return EVSL_ERR_ARG;
  }
  if (vec_kind != EVSL_VEC_RADEMACHER && vec_kind != EVSL_VEC_GAUSSIAN) {
    set_error_msg("kpmdos_batch: vec_kind must be EVSL_VEC_RADEMACHER or EVSL_VEC_GAUSSIAN");
    return EVSL_ERR_UNSUPP;
  }
  for (int r = 0; r < nreq; ++r) {
    const csrMat* A = mats[r];
    if (!A || A->nrows <= 0 || A->nrows != A->ncols || !A->ia || (A->ia[A->nrows] > 0 && (!A->ja || !A->a))) {
      set_error_msg("kpmdos_batch: request " + std::to_string(r) + " has no square CSR matrix");
      return EVSL_ERR_ARG;
    }
    const double* iv = intvs + 4 * r;
    if (iv[0] >= iv[1] || iv[0] >= iv[3] || iv[1] <= iv[2]) {
      set_error_msg("kpmdos_batch: request " + std::to_string(r) + " has an invalid interval");
      return -1;
    }
  }
  const bool f32 = prec == EVSL_PREC_MIXED;
  const size_t vs = f32 ? 4 : 8;
  // distinct matrices on the device
  std::vector<std::unique_ptr<BatchMatrix>> mat;
  std::map<const csrMat*, BatchMatrix*> by_ptr;
  std::vector<BatchRequest> req(nreq);
  for (int r = 0; r < nreq; ++r) {
    const csrMat* A = mats[r];
    auto it = by_ptr.find(A);
    if (it == by_ptr.end()) {
      auto m = std::make_unique<BatchMatrix>();
      m->host = A;
      m->n = A->nrows;
      m->nnz = A->ia[A->nrows];
      if (int e = upload(m->ia, A->ia, (size_t)(m->n + 1) * sizeof(int))) return e;
      if (int e = upload(m->ja, A->ja, (size_t)m->nnz * sizeof(int))) return e;
      if (int e = upload(m->a64, A->a, (size_t)m->nnz * sizeof(double))) return e;
      bool full = true;
      for (int i = 0; i < m->n && full; ++i) {
        bool has = false;
        for (int k = A->ia[i]; k < A->ia[i + 1]; ++k)
          if (A->ja[k] == i) {
            has = true;
            break;
          }
        full = has;
      }
      m->full_diag = full;
      it = by_ptr.emplace(A, m.get()).first;
      mat.push_back(std::move(m));
    }
    BatchRequest& q = req[r];
    q.mat = it->second;
    q.intv = intvs + 4 * r;
    q.seed = seeds ? seeds[r] : (unsigned long long)r;
    q.ctr = (q.intv[3] + q.intv[2]) / 2.0;
    q.wid = (q.intv[3] - q.intv[2]) / 2.0;
    q.t0 = 1.0 / q.wid;
    q.t = 2.0 / q.wid;
    q.cc = q.ctr;
    q.raw.assign(Mdeg + 1, 0.0);
    if (f32 && q.mat->nnz > 0) {
      EVSL_HIP(q.a32.alloc((size_t)q.mat->nnz * sizeof(float)));
      if (q.mat->full_diag) {
        // as the mixed filter: round (A - ctr I) * 2^-s, not A, to FP32
        const double scale = std::ldexp(1.0, -(int)std::lround(std::log2(q.wid)));
        EVSL_HIP(evsl_dev::shift_scale_f32(q.mat->view(), q.ctr, scale, q.a32.as<float>(),
                                           c.stream));
        q.t0 /= scale;
        q.t /= scale;
        q.cc = 0.0;
      } else {
        EVSL_HIP(evsl_dev::f64_to_f32(q.mat->nnz, q.mat->a64.as<double>(), q.a32.as<float>(),
                                      c.stream));
      }
    }
  }
  int maxn = 0;
  for (auto& m : mat) maxn = std::max(maxn, m->n);
  DevBuf x64, nrm_part, nrm2;
  EVSL_HIP(x64.alloc((size_t)maxn * 32 * sizeof(double)));
  EVSL_HIP(nrm_part.alloc((size_t)1024 * 32 * sizeof(double)));
  EVSL_HIP(nrm2.alloc(32 * sizeof(double)));
  DevBuf d_tab, d_owner, d_sums, d_owner1, d_first1, d_owner2;
  std::vector<double> hdots;
  int done = 0;
  while (done < nvec) {
    int B = 32;
    while (B > nvec - done) B >>= 1;
    const int V = evsl_dev::cheb_spmm_vec(B, f32, evsl_dev::kKpm);
    const int R = evsl_dev::cheb_spmm_rows(B, f32, evsl_dev::kKpm);
    if (V <= 0 || R <= 0) return EVSL_ERR_ARG;
    const int lanes = B / V / R;
    std::vector<evsl_dev::KpmGroupEntry> tab(nreq);
    std::vector<evsl_dev::KpmSumEntry> sums(nreq);
    std::vector<int> owner, owner1, first1(nreq), owner2;
    int total = 0;
    for (int r = 0; r < nreq; ++r) {
      BatchRequest& q = req[r];
      BatchMatrix& m = *q.mat;
      if (int e = m.ensure_table(lanes)) return e;
      const int nblk = m.nblk[lanes];
      const size_t nb = (size_t)m.n * B;
      EVSL_HIP(q.x0.alloc(nb * vs));
      EVSL_HIP(q.ring.alloc(3 * nb * vs));
      EVSL_HIP(q.part.alloc((size_t)std::max(nblk, 1) * B * sizeof(double)));
      EVSL_HIP(q.dots.alloc((size_t)(Mdeg + 1) * B * sizeof(double)));
      const int G = evsl_dev::sum_levels_blocks(nblk, B);
      if (G > 1) EVSL_HIP(q.scratch.alloc((size_t)G * B * sizeof(double)));
      // start block: the kpmdos_ex vectors of this request's seed, unit columns
      double* xd = f32 ? x64.as<double>() : q.x0.as<double>();
      EVSL_HIP(evsl_dev::rm_random(m.n, B, vec_kind == EVSL_VEC_RADEMACHER ? 0 : 1, q.seed, done,
                                   B, xd, c.stream));
      EVSL_HIP(evsl_dev::rm_col_nrm2(m.n, B, xd, nrm_part.as<double>(), nrm2.as<double>(),
                                     c.stream));
      EVSL_HIP(evsl_dev::rm_col_scale_rsqrt(m.n, B, xd, nrm2.as<double>(), c.stream));
      if (f32) EVSL_HIP(evsl_dev::f64_to_f32(nb, xd, q.x0.as<float>(), c.stream));
      evsl_dev::KpmGroupEntry& e = tab[r];
      e.ia = m.ia.as<int>();
      e.ja = m.ja.as<int>();
      e.a = f32 ? (const void*)q.a32.p : (const void*)m.a64.p;
      e.rbp = m.table[lanes].as<int>();
      e.x0 = q.x0.p;
      e.ring = q.ring.p;
      e.part = q.part.as<double>();
      e.t0 = q.t0;
      e.t = q.t;
      e.cc = q.cc;
      e.n = m.n;
      e.blk0 = total;
      e.k0 = 0;
      e.pad = 0;
      owner.insert(owner.end(), nblk, r);
      total += nblk;
      evsl_dev::KpmSumEntry& s = sums[r];
      s.part = q.part.as<double>();
      s.scratch = q.scratch.as<double>();
      s.dots = q.dots.as<double>();
      s.nblk = nblk;
      s.G = G;
      s.k0 = 0;
      s.pad = 0;
      first1[r] = (int)owner1.size();
      owner1.insert(owner1.end(), G, r);
      if (G > 1) owner2.push_back(r);
    }
    if (int e = upload(d_tab, tab.data(), tab.size() * sizeof(tab[0]))) return e;
    if (int e = upload(d_owner, owner.data(), owner.size() * sizeof(int))) return e;
    if (int e = upload(d_sums, sums.data(), sums.size() * sizeof(sums[0]))) return e;
    if (int e = upload(d_owner1, owner1.data(), owner1.size() * sizeof(int))) return e;
    if (int e = upload(d_first1, first1.data(), first1.size() * sizeof(int))) return e;
    if (int e = upload(d_owner2, owner2.data(), owner2.size() * sizeof(int))) return e;
    double step_bytes = 0.0;
    for (int r = 0; r < nreq; ++r)
      step_bytes += (double)req[r].mat->nnz * (vs + 4) + 4.0 * (req[r].mat->n + 1) +
                    4.0 * req[r].mat->n * B * vs;
    for (int k = 0; k < Mdeg; ++k) {
      EVSL_HIP(evsl_dev::kpm_group_step(
          f32, B, d_tab.as<evsl_dev::KpmGroupEntry>(), d_owner.as<int>(), total, k,
          d_sums.as<evsl_dev::KpmSumEntry>(), d_owner1.as<int>(), d_first1.as<int>(),
          (int)owner1.size(), d_owner2.as<int>(), (int)owner2.size(), c.stream));
      c.astats.filter_bytes += step_bytes;
    }
    c.astats.n_filter_steps += Mdeg;
    c.astats.n_filter_cols += (size_t)Mdeg * B * nreq;
    hdots.resize((size_t)(Mdeg + 1) * B);
    for (int r = 0; r < nreq; ++r) {
      EVSL_HIP(hipMemcpy(hdots.data(), req[r].dots.p, hdots.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
      for (int k = 1; k <= Mdeg; ++k)
        for (int col = 0; col < B; ++col) req[r].raw[k] += hdots[(size_t)k * B + col];
    }
    done += B;
  }
  for (int r = 0; r < nreq; ++r)
    kpm_moments(Mdeg, damping, nvec, req[r].intv, req[r].mat->n, req[r].raw.data(),
                mu + (size_t)r * (Mdeg + 1), ecnt + r);
  return 0;
}

}  // extern "C"
