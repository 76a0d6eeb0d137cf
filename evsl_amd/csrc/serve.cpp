// serve.cpp -- a request loop over device-resident matrices.
//
// Matrices are registered once and stay in HBM (CSR, row-block tables,
// FP32 copies).  Requests arrive at any time: slice requests (ChebLanNr on
// an interval of a registered matrix) and DOS requests (kpmdos with seeded
// device vectors).  Each evsl_amd_server_step:
//   1. admits queued requests in arrival order while their projected memory
//      (the Lanczos basis n x (maxit+1) x 8 B and, for a mixed slice that
//      filters in FP32, its FP32 copy, or the KPM vector blocks) fits
//      the budget;
//   2. advances, per matrix, every running slice by one filtered Lanczos step
//      in one blocked pass (LanczosGroup: a slice admitted now joins at this
//      step; each slice counts its own steps, so its result is the same alone
//      and in the mix);
//   3. advances every running DOS request by up to kDosQuantum Chebyshev
//      degrees in grouped launches over all of them (one launch per degree);
// and stamps each request's completion time.  The loop is driven by the
// caller (one step per call), so arrivals between steps are admitted at the
// next step boundary.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"

namespace evsl {
namespace {

constexpr int kDosQuantum = 64;  // KPM degrees per DOS request and server step

struct SrvMatrix {
  DevMatrix M;
  double lmin = 0, lmax = 0;
  std::unique_ptr<LanczosGroup> grp;
  ~SrvMatrix() { M.release(); }
};

// makes a server matrix the context operator while it lives (the caller's
// own operator is set aside and restored)
struct Binding {
  DevMatrix& M;
  int n, op;
  explicit Binding(DevMatrix& m) : M(m) {
    Context& c = ctx();
    std::swap(c.A, M);
    n = c.n;
    op = c.op;
    c.n = c.A.n;
    c.op = 1;
  }
  ~Binding() {
    Context& c = ctx();
    std::swap(c.A, M);
    c.n = n;
    c.op = op;
  }
};

struct SrvRequest {
  int id = 0, kind = 0, mat = 0;  // kind 0 slice, 1 DOS
  int state = 0;                  // 0 queued, 1 running, 2 done
  double t_submit = 0, t_start = 0, t_done = 0;
  size_t mem = 0;
  int err = 0;
  // slice
  double intv[4] = {0, 0, 0, 0};
  int maxit = 0;
  double tol = 0;
  polparams pol{};
  DevBuf v0;
  SliceResult res;
  // DOS
  int Mdeg = 0, damping = 0, nvec = 0, vec_kind = 0;
  unsigned long long seed = 0;
  int done_vec = 0, B = 0, k = 0, nblk = 0, G = 1;
  double t0 = 0, t = 0, cc = 0;
  const float* a32 = nullptr;
  DevBuf x0, ring, part, dots, scratch;
  std::vector<double> raw, mu;
  double ecnt = 0;
  ~SrvRequest() {
    if (pol.mu) free_pol(&pol);
    std::free(res.lam);
    std::free(res.res);
    if (res.W) _evsl_Free_device(res.W);
  }
};

}  // namespace
}  // namespace evsl

using namespace evsl;

struct evslAmdServer {
  size_t budget = 0, in_use = 0, peak = 0;
  bool mixed = false;
  std::vector<std::unique_ptr<SrvMatrix>> mats;
  std::vector<std::unique_ptr<SrvRequest>> reqs;
  std::deque<int> queue;
  long long kg = 0;  // grouped KPM launches so far (DOS entries start at k0 = kg - k)
  DevBuf d_tab, d_owner, d_sums, d_owner1, d_first1, d_owner2, x64, nrm_part, nrm2;
  GroupedWork gw;  // the slices' grouped filter passes
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~evslAmdServer() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

int upload(DevBuf& d, const void* src, size_t bytes) {
  EVSL_HIP(d.alloc(std::max<size_t>(bytes, 8)));
  if (bytes) EVSL_HIP(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

size_t dos_mem(const SrvMatrix& m, bool f32) {
  const size_t vs = f32 ? 4 : 8;
  return (size_t)m.M.n * 32 * (4 * vs + 8) + (size_t)m.M.n * 8;
}

// starts the next column block of a DOS request: seeded start vectors
// (kpmdos_ex's, unit columns), ring and partial buffers
int dos_round(evslAmdServer* s, SrvRequest& q) {
  Context& c = ctx();
  SrvMatrix& m = *s->mats[q.mat];
  const bool f32 = s->mixed;
  const size_t vs = f32 ? 4 : 8;
  int B = 32;
  while (B > q.nvec - q.done_vec) B >>= 1;
  q.B = B;
  q.k = 0;
  const int V = evsl_dev::cheb_spmm_vec(B, f32, evsl_dev::kKpm);
  const int R = evsl_dev::cheb_spmm_rows(B, f32, evsl_dev::kKpm);
  const int slot = evsl_dev::width_slot(B / V / R);
  if (slot < 0) return EVSL_ERR_ARG;
  q.nblk = m.M.nblk[slot];
  q.G = evsl_dev::sum_levels_blocks(q.nblk, B);
  const int n = m.M.n;
  const size_t nb = (size_t)n * B;
  EVSL_HIP(q.x0.alloc(nb * vs));
  EVSL_HIP(q.ring.alloc(3 * nb * vs));
  EVSL_HIP(q.part.alloc((size_t)std::max(q.nblk, 1) * B * sizeof(double)));
  EVSL_HIP(q.dots.alloc((size_t)(q.Mdeg + 1) * B * sizeof(double)));
  if (q.G > 1) EVSL_HIP(q.scratch.alloc((size_t)q.G * B * sizeof(double)));
  EVSL_HIP(s->x64.alloc(nb * sizeof(double)));
  EVSL_HIP(s->nrm_part.alloc((size_t)1024 * 32 * sizeof(double)));
  EVSL_HIP(s->nrm2.alloc(32 * sizeof(double)));
  double* xd = f32 ? s->x64.as<double>() : q.x0.as<double>();
  EVSL_HIP(evsl_dev::rm_random(n, B, q.vec_kind == EVSL_VEC_RADEMACHER ? 0 : 1, q.seed,
                               q.done_vec, B, xd, c.stream));
  EVSL_HIP(evsl_dev::rm_col_nrm2(n, B, xd, s->nrm_part.as<double>(), s->nrm2.as<double>(),
                                 c.stream));
  EVSL_HIP(evsl_dev::rm_col_scale_rsqrt(n, B, xd, s->nrm2.as<double>(), c.stream));
  if (f32) EVSL_HIP(evsl_dev::f64_to_f32(nb, xd, q.x0.as<float>(), c.stream));
  return 0;
}

int admit(evslAmdServer* s) {
  while (!s->queue.empty()) {
    SrvRequest& q = *s->reqs[s->queue.front()];
    if (s->in_use + q.mem > s->budget && s->in_use > 0) break;  // arrival order
    s->queue.pop_front();
    s->in_use += q.mem;
    s->peak = std::max(s->peak, s->in_use);
    q.state = 1;
    q.t_start = evsl_timer();
    SrvMatrix& m = *s->mats[q.mat];
    if (q.kind == 0) {
      if (!m.grp) m.grp = std::make_unique<LanczosGroup>(s->mixed, nullptr);
      Binding bind(m.M);
      if (int e = m.grp->add(q.intv, q.maxit, q.tol, q.v0.as<double>(), &q.pol, &q.res))
        return e;
    } else {
      q.raw.assign(q.Mdeg + 1, 0.0);
      q.t0 = 1.0 / ((m.lmax - m.lmin) / 2.0);
      q.t = 2.0 * q.t0;
      q.cc = (m.lmax + m.lmin) / 2.0;
      if (s->mixed) {
        // the filter's shifted FP32 copy: the same (cc, scale) as the slices'
        Binding bind(m.M);
        const double dd = (m.lmax - m.lmin) / 2.0;
        const double scale = std::ldexp(1.0, -(int)std::lround(std::log2(dd)));
        bool ok = false;
        if (int e = ensure_a32_shifted(q.cc, scale, &ok)) return e;
        if (ok) {
          q.a32 = ctx().A.a32s;
          q.t0 /= scale;
          q.t /= scale;
          q.cc = 0.0;
        } else {
          if (int e = ensure_a32()) return e;
          q.a32 = ctx().A.a32;
        }
      }
      if (int e = dos_round(s, q)) return e;
    }
  }
  return 0;
}

// A matrix whose filter launch moves more than this is filtered on its own
// (the launch fills the device by itself; its own filter keeps its tuned
// lane layout and drops columns as their degrees end); smaller ones, whose
// launches are latency-bound, share grouped launches.
constexpr double kGroupLaunchBytes = 32e6;

// One filter pass of the slices of every listed matrix: the start vectors
// of slices added since the last step (start), or the current Lanczos
// vectors of the active slices -- the small matrices' columns in grouped
// launches, each large one with its own filter; then, for a start pass, the
// new slices are normalised and join.

int grouped_pass(evslAmdServer* s, const std::vector<int>& live, bool start) {
  Context& c = ctx();
  if (!s->ev[0])
    for (auto& e : s->ev) EVSL_HIP(hipEventCreate(&e));
  EVSL_HIP(hipEventRecord(s->ev[0], c.stream));
  bool any = false;
  std::vector<GroupedCols> parts;
  for (int mi : live) {
    SrvMatrix& m = *s->mats[mi];
    Binding bind(m.M);
    std::vector<LanczosGroup::FilterCol> cols;
    if (int e = start ? m.grp->start_cols(cols) : m.grp->step_cols(cols)) return e;
    for (int pr = 0; pr < 2; ++pr) {
      GroupedCols gc;
      for (const auto& col : cols)
        if (col.f32 == (pr == 1)) gc.cols.push_back(col);
      if (gc.cols.empty()) continue;
      any = true;
      // (the operator's view carries this matrix's sizes; inside the
      // binding m.M holds the context's previous matrix)
      const polparams* p0 = gc.cols[0].pol;
      if (int e = make_filter_op(p0->cc, p0->dd, pr == 1, &gc.op)) return e;
      const double vs = pr == 1 ? 4.0 : 8.0;
      const double launch = (double)gc.op.A.nnz * (vs + (gc.op.A.jd ? 2 : 4)) +
                            (double)gc.op.A.n * gc.cols.size() * (3.0 * vs + 16.0 / 3.0);
      if (launch > kGroupLaunchBytes) {
        if (int e = m.grp->filter_cols(gc.cols)) return e;
      } else {
        parts.push_back(std::move(gc));
      }
    }
  }
  if (any) {
    if (!parts.empty())
      if (int e = grouped_filter(parts, s->gw)) return e;
    EVSL_HIP(hipEventRecord(s->ev[1], c.stream));
    float ms = 0.f;
    EVSL_HIP(hipEventSynchronize(s->ev[1]));
    EVSL_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    c.astats.t_filter += ms * 1e-3;
  }
  if (start)
    for (int mi : live) {
      Binding bind(s->mats[mi]->M);
      if (int e = s->mats[mi]->grp->start_done()) return e;
    }
  return 0;
}

void finish(evslAmdServer* s, SrvRequest& q) {
  q.state = 2;
  q.t_done = evsl_timer();
  s->in_use -= std::min(s->in_use, q.mem);
  q.v0.release();
  q.x0.release();
  q.ring.release();
  q.part.release();
  q.dots.release();
  q.scratch.release();
}

int dos_steps(evslAmdServer* s) {
  const bool f32 = s->mixed;
  int budget = kDosQuantum;
  while (budget > 0) {
    // running DOS requests of one block width (a launch has one width)
    std::vector<int> ids;
    int B = 0;
    for (auto& r : s->reqs)
      if (r->kind == 1 && r->state == 1 && r->k < r->Mdeg && (B == 0 || r->B == B)) {
        B = r->B;
        ids.push_back(r->id);
      }
    if (ids.empty()) break;
    const int nr = (int)ids.size();
    std::vector<evsl_dev::KpmGroupEntry> tab(nr);
    std::vector<evsl_dev::KpmSumEntry> sums(nr);
    std::vector<int> owner, owner1, first1(nr), owner2;
    int total = 0;
    const int V = evsl_dev::cheb_spmm_vec(B, f32, evsl_dev::kKpm);
    const int R = evsl_dev::cheb_spmm_rows(B, f32, evsl_dev::kKpm);
    const int slot = evsl_dev::width_slot(B / V / R);
    for (int i = 0; i < nr; ++i) {
      SrvRequest& q = *s->reqs[ids[i]];
      SrvMatrix& m = *s->mats[q.mat];
      evsl_dev::KpmGroupEntry& e = tab[i];
      e.ia = m.M.ia;
      e.ja = m.M.ja;
      e.a = f32 ? (const void*)q.a32 : (const void*)m.M.a64;
      e.rbp = m.M.rb[slot];
      e.x0 = q.x0.p;
      e.ring = q.ring.p;
      e.part = q.part.as<double>();
      e.t0 = q.t0;
      e.t = q.t;
      e.cc = q.cc;
      e.n = m.M.n;
      e.blk0 = total;
      e.k0 = (int)(s->kg - q.k);
      e.pad = 0;
      owner.insert(owner.end(), q.nblk, i);
      total += q.nblk;
      evsl_dev::KpmSumEntry& se = sums[i];
      se.part = q.part.as<double>();
      se.scratch = q.scratch.as<double>();
      se.dots = q.dots.as<double>();
      se.nblk = q.nblk;
      se.G = q.G;
      se.k0 = e.k0;
      se.pad = 0;
      first1[i] = (int)owner1.size();
      owner1.insert(owner1.end(), q.G, i);
      if (q.G > 1) owner2.push_back(i);
    }
    if (int e = upload(s->d_tab, tab.data(), tab.size() * sizeof(tab[0]))) return e;
    if (int e = upload(s->d_owner, owner.data(), owner.size() * sizeof(int))) return e;
    if (int e = upload(s->d_sums, sums.data(), sums.size() * sizeof(sums[0]))) return e;
    if (int e = upload(s->d_owner1, owner1.data(), owner1.size() * sizeof(int))) return e;
    if (int e = upload(s->d_first1, first1.data(), first1.size() * sizeof(int))) return e;
    if (int e = upload(s->d_owner2, owner2.data(), owner2.size() * sizeof(int))) return e;
    // launches while the membership stays the same
    int run = budget;
    for (int id : ids) run = std::min(run, s->reqs[id]->Mdeg - s->reqs[id]->k);
    for (int j = 0; j < run; ++j) {
      EVSL_HIP(evsl_dev::kpm_group_step(
          f32, B, s->d_tab.as<evsl_dev::KpmGroupEntry>(), s->d_owner.as<int>(), total,
          (int)s->kg, s->d_sums.as<evsl_dev::KpmSumEntry>(), s->d_owner1.as<int>(),
          s->d_first1.as<int>(), (int)owner1.size(), s->d_owner2.as<int>(), (int)owner2.size(),
          ctx().stream));
      ++s->kg;
    }
    for (int id : ids) s->reqs[id]->k += run;
    budget -= run;
    // requests whose column block is through: moments, next block or done
    for (int id : ids) {
      SrvRequest& q = *s->reqs[id];
      if (q.k < q.Mdeg) continue;
      std::vector<double> h((size_t)(q.Mdeg + 1) * q.B);
      EVSL_HIP(hipMemcpy(h.data(), q.dots.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int k = 1; k <= q.Mdeg; ++k)
        for (int col = 0; col < q.B; ++col) q.raw[k] += h[(size_t)k * q.B + col];
      q.done_vec += q.B;
      if (q.done_vec < q.nvec) {
        if (int e = dos_round(s, q)) return e;
      } else {
        q.mu.assign(q.Mdeg + 1, 0.0);
        kpm_moments(q.Mdeg, q.damping, q.nvec, q.intv, s->mats[q.mat]->M.n, q.raw.data(),
                    q.mu.data(), &q.ecnt);
        finish(s, q);
      }
    }
  }
  return 0;
}

SrvRequest* get(evslAmdServer* s, int id) {
  if (!s || id < 0 || id >= (int)s->reqs.size()) return nullptr;
  return s->reqs[id].get();
}

}  // namespace

extern "C" {

evslAmdServer* evsl_amd_server_create(size_t mem_budget, int prec) {
  if (ensure_started()) return nullptr;
  if (prec != EVSL_PREC_FP64 && prec != EVSL_PREC_MIXED) {
    set_error_msg("evsl_amd_server_create: prec must be EVSL_PREC_FP64 or EVSL_PREC_MIXED");
    return nullptr;
  }
  auto* s = new evslAmdServer;
  s->budget = mem_budget;
  s->mixed = prec == EVSL_PREC_MIXED;
  return s;
}

void evsl_amd_server_destroy(evslAmdServer* s) {
  if (!s) return;
  (void)hipDeviceSynchronize();
  delete s;
}

int evsl_amd_server_add_matrix(evslAmdServer* s, csrMat* A, double lmin, double lmax) {
  if (int e = ensure_started()) return e;
  if (!s || !A || A->nrows != A->ncols || A->nrows <= 0 || !(lmin < lmax)) {
    set_error_msg("evsl_amd_server_add_matrix: invalid arguments");
    return EVSL_ERR_ARG;
  }
  const int n = A->nrows, nnz = A->ia[n];
  int *ia = nullptr, *ja = nullptr;
  double* a = nullptr;
  EVSL_HIP(hipMalloc((void**)&ia, (n + 1) * sizeof(int)));
  EVSL_HIP(hipMalloc((void**)&ja, (size_t)std::max(nnz, 1) * sizeof(int)));
  EVSL_HIP(hipMalloc((void**)&a, (size_t)std::max(nnz, 1) * sizeof(double)));
  EVSL_HIP(hipMemcpy(ia, A->ia, (n + 1) * sizeof(int), hipMemcpyHostToDevice));
  EVSL_HIP(hipMemcpy(ja, A->ja, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
  EVSL_HIP(hipMemcpy(a, A->a, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice));
  auto m = std::make_unique<SrvMatrix>();
  size_t need = 0;
  if (int e = build_device_matrix(n, nnz, ia, ja, a, A->ia, A->ja, true, m->M, &need)) return e;
  if (int e = ensure_red_part(need)) return e;
  m->lmin = lmin;
  m->lmax = lmax;
  s->mats.push_back(std::move(m));
  return (int)s->mats.size() - 1;
}

int evsl_amd_server_submit_slice(evslAmdServer* s, int mat, double a, double b, int maxit,
                                 double tol, const polparams* pol_in, const double* vinit) {
  if (int e = ensure_started()) return e;
  if (!s || mat < 0 || mat >= (int)s->mats.size() || !(a < b) || maxit <= 0 || !vinit) {
    set_error_msg("evsl_amd_server_submit_slice: invalid arguments");
    return EVSL_ERR_ARG;
  }
  SrvMatrix& m = *s->mats[mat];
  auto q = std::make_unique<SrvRequest>();
  q->id = (int)s->reqs.size();
  q->kind = 0;
  q->mat = mat;
  q->t_submit = evsl_timer();
  q->intv[0] = a;
  q->intv[1] = b;
  q->intv[2] = m.lmin;
  q->intv[3] = m.lmax;
  q->maxit = std::min(maxit, m.M.n);
  q->tol = tol;
  if (pol_in) {
    q->pol = *pol_in;
    q->pol.mu = nullptr;
  } else {
    set_pol_def(&q->pol);
    q->pol.damping = 2;
    q->pol.thresh_int = 0.8;
    q->pol.thresh_ext = 0.2;
    q->pol.max_deg = 3000;
  }
  if (find_pol(q->intv, &q->pol) != 0) {
    set_error_msg("evsl_amd_server_submit_slice: find_pol failed for the interval");
    return -1;
  }
  const double* v0;
  DevBuf tmp;
  if (int e = as_device_vector(vinit, m.M.n, tmp, &v0)) return e;
  EVSL_HIP(q->v0.alloc((size_t)m.M.n * sizeof(double)));
  EVSL_HIP(hipMemcpy(q->v0.p, v0, (size_t)m.M.n * sizeof(double), hipMemcpyDeviceToDevice));
  q->mem = slice_basis_bytes(m.M.n, q->maxit, s->mixed, &q->pol);
  s->queue.push_back(q->id);
  s->reqs.push_back(std::move(q));
  return s->reqs.back()->id;
}

int evsl_amd_server_submit_dos(evslAmdServer* s, int mat, double a, double b, int Mdeg,
                               int damping, int nvec, int vec_kind, unsigned long long seed) {
  if (int e = ensure_started()) return e;
  if (!s || mat < 0 || mat >= (int)s->mats.size() || Mdeg <= 0 || nvec <= 0 ||
      (vec_kind != EVSL_VEC_RADEMACHER && vec_kind != EVSL_VEC_GAUSSIAN)) {
    set_error_msg("evsl_amd_server_submit_dos: invalid arguments");
    return EVSL_ERR_ARG;
  }
  SrvMatrix& m = *s->mats[mat];
  if (a >= b || a >= m.lmax || b <= m.lmin) {
    set_error_msg("evsl_amd_server_submit_dos: invalid interval");
    return -1;
  }
  auto q = std::make_unique<SrvRequest>();
  q->id = (int)s->reqs.size();
  q->kind = 1;
  q->mat = mat;
  q->t_submit = evsl_timer();
  q->intv[0] = a;
  q->intv[1] = b;
  q->intv[2] = m.lmin;
  q->intv[3] = m.lmax;
  q->Mdeg = Mdeg;
  q->damping = damping;
  q->nvec = nvec;
  q->vec_kind = vec_kind;
  q->seed = seed;
  q->mem = dos_mem(*s->mats[mat], s->mixed);
  s->queue.push_back(q->id);
  s->reqs.push_back(std::move(q));
  return s->reqs.back()->id;
}

int evsl_amd_server_step(evslAmdServer* s) {
  if (int e = ensure_started()) return e;
  if (!s) return EVSL_ERR_ARG;
  if (int e = admit(s)) return e;
  std::vector<int> live;
  for (int mi = 0; mi < (int)s->mats.size(); ++mi)
    if (s->mats[mi]->grp && s->mats[mi]->grp->active() > 0) live.push_back(mi);
  if (ctx().grouped_filter) {
    // the filter passes of every matrix's slices in grouped launches (new
    // slices' start vectors, then every active slice's Lanczos vector), the
    // rest of the step per matrix
    if (int e = grouped_pass(s, live, true)) return e;
    if (int e = grouped_pass(s, live, false)) return e;
    for (int mi : live) {
      Binding bind(s->mats[mi]->M);
      if (int e = s->mats[mi]->grp->step_done()) return e;
    }
  } else {
    for (int mi : live) {
      Binding bind(s->mats[mi]->M);
      if (int e = s->mats[mi]->grp->step()) return e;
    }
  }
  for (int mi : live)
    for (auto& r : s->reqs)
      if (r->kind == 0 && r->state == 1 && r->mat == mi && r->res.done) {
        r->err = r->res.err;
        finish(s, *r);
        r->t_done = r->res.t_done;
      }
  if (int e = dos_steps(s)) return e;
  int left = 0;
  for (auto& r : s->reqs) left += r->state != 2;
  return left;
}

int evsl_amd_server_status(evslAmdServer* s, int id, double* t_submit, double* t_start,
                           double* t_done) {
  SrvRequest* q = get(s, id);
  if (!q) return EVSL_ERR_ARG;
  if (t_submit) *t_submit = q->t_submit;
  if (t_start) *t_start = q->t_start;
  if (t_done) *t_done = q->t_done;
  return q->state;
}

int evsl_amd_server_slice_result(evslAmdServer* s, int id, int* nev, double** lam, double** Y,
                                 double** res) {
  SrvRequest* q = get(s, id);
  if (!q || q->kind != 0 || !nev || !lam || !Y || !res) return EVSL_ERR_ARG;
  if (q->state != 2) return EVSL_ERR_ARG;
  *nev = q->res.nev;
  *lam = q->res.lam;
  *Y = q->res.W;
  *res = q->res.res;
  q->res.lam = q->res.W = q->res.res = nullptr;  // ownership passes to the caller
  return q->err;
}

int evsl_amd_server_dos_result(evslAmdServer* s, int id, double* mu, double* ecnt) {
  SrvRequest* q = get(s, id);
  if (!q || q->kind != 1 || !mu || !ecnt || q->state != 2) return EVSL_ERR_ARG;
  std::memcpy(mu, q->mu.data(), q->mu.size() * sizeof(double));
  *ecnt = q->ecnt;
  return q->err;
}

int evsl_amd_server_stats(evslAmdServer* s, int* queued, int* running, size_t* mem_in_use,
                          size_t* mem_peak) {
  if (!s) return EVSL_ERR_ARG;
  int qd = 0, rn = 0;
  for (auto& r : s->reqs) {
    qd += r->state == 0;
    rn += r->state == 1;
  }
  if (queued) *queued = qd;
  if (running) *running = rn;
  if (mem_in_use) *mem_in_use = s->in_use;
  if (mem_peak) *mem_peak = s->peak;
  return 0;
}

}  // extern "C"
