// grouped.cpp -- ChebAv over the column blocks of several matrices in
// grouped launches (the request loop's filter passes).
//
// Each matrix contributes its columns at one precision, split into blocks of
// at most 8 columns.  Blocks are classed by (precision, width, column-index
// type); within a class they are ordered by decreasing series degree and
// their row blocks numbered one after the other, so the launch at degree k
// covers the prefix of the blocks whose series has not ended.  Every block
// runs the one-matrix fused step (the same kernel body, the fixed lane
// layout), so each column gets the bits its one-matrix filter gives it:
// grouping changes the schedule, never a result.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "internal.h"

namespace evsl {

namespace {

constexpr int kGroupMaxWidth = 8;

struct Block {
  const GroupedCols* part = nullptr;
  int c0 = 0, B = 0, deg = 0;
  size_t buf = 0;  // workspace slot
};

// class key: precision, width, 16-bit indices
int class_key(bool f32, int B, bool i16) { return (f32 ? 1000 : 0) + (i16 ? 100 : 0) + B; }

}  // namespace

int grouped_filter(const std::vector<GroupedCols>& parts, GroupedWork& W) {
  Context& c = ctx();
  // ---- blocks of at most 8 columns; degree-0 filters directly -------------
  std::vector<Block> blocks;
  for (const GroupedCols& p : parts) {
    const int nc = (int)p.cols.size();
    for (int c0 = 0; c0 < nc; c0 += kGroupMaxWidth) {
      Block b;
      b.part = &p;
      b.c0 = c0;
      b.B = std::min(kGroupMaxWidth, nc - c0);
      for (int j = 0; j < b.B; ++j) b.deg = std::max(b.deg, p.cols[c0 + j].pol->deg);
      if (b.deg == 0) {
        for (int j = 0; j < b.B; ++j) {
          const auto& col = p.cols[c0 + j];
          EVSL_HIP(evsl_dev::scal_copy(p.op.A.n, col.pol->mu[0], col.in, col.out, c.stream));
        }
        continue;
      }
      blocks.push_back(b);
    }
  }
  if (blocks.empty()) return 0;
  // ---- classes, ordered by decreasing degree within each ------------------
  std::map<int, std::vector<int>> classes;
  for (int i = 0; i < (int)blocks.size(); ++i) {
    const Block& b = blocks[i];
    classes[class_key(b.part->op.f32, b.B, b.part->op.A.jd != nullptr)].push_back(i);
  }
  while (W.bufs.size() < blocks.size()) W.bufs.push_back(std::make_unique<GroupedWork::Bufs>());
  std::vector<evsl_dev::ChebGroupEntry> tab;
  std::vector<int> owner;
  struct Launch {
    bool f32, i16;
    int B;
    size_t tab0, own0;
    std::vector<int> deg, blk_end;  // per entry in order: degree, prefix row blocks
  };
  std::vector<Launch> launches;
  size_t slot = 0;
  int maxdeg = 0;
  for (auto& kv : classes) {
    std::vector<int>& ids = kv.second;
    std::stable_sort(ids.begin(), ids.end(),
                     [&](int x, int y) { return blocks[x].deg > blocks[y].deg; });
    Launch L;
    const Block& b0 = blocks[ids[0]];
    L.f32 = b0.part->op.f32;
    L.i16 = b0.part->op.A.jd != nullptr;
    L.B = b0.B;
    L.tab0 = tab.size();
    L.own0 = owner.size();
    const int lay = evsl_dev::spmm_table_layout(L.B, L.f32, evsl_dev::kCheb);
    const int V = lay / 10, R = lay % 10;
    const int wslot = evsl_dev::width_slot(L.B / V / R);
    if (wslot < 0) return EVSL_ERR_ARG;
    const size_t vs = L.f32 ? 4 : 8;
    int blk0 = 0;
    for (int id : ids) {
      Block& b = blocks[id];
      b.buf = slot++;
      const GroupedCols& p = *b.part;
      const size_t nb = (size_t)p.op.A.n * b.B;
      GroupedWork::Bufs& bw = *W.bufs[b.buf];
      EVSL_HIP(bw.X0.alloc(nb * vs));
      EVSL_HIP(bw.ring.alloc(3 * nb * vs));
      EVSL_HIP(bw.Y.alloc(nb * sizeof(double)));
      // coefficient table, kept on the device while the same filters return
      // (a filter is known by its address, its coefficient array and degree)
      std::vector<GroupedWork::PolKey> key(b.B);
      for (int j = 0; j < b.B; ++j) {
        const polparams* q = p.cols[b.c0 + j].pol;
        key[j] = {q, q->mu, q->deg};
      }
      if (key != bw.mus_key || b.deg != bw.mus_deg) {
        std::vector<double> t((size_t)(b.deg + 1) * b.B, 0.0);
        for (int j = 0; j < b.B; ++j) {
          const polparams* q = p.cols[b.c0 + j].pol;
          for (int d = 0; d <= q->deg; ++d) t[(size_t)d * b.B + j] = q->mu[d];
        }
        EVSL_HIP(bw.mus.alloc(t.size() * sizeof(double)));
        EVSL_HIP(hipMemcpy(bw.mus.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        bw.mus_key = key;
        bw.mus_deg = b.deg;
      }
      // interleave the input columns
      evsl_dev::ColPtrs cp{};
      for (int j = 0; j < b.B; ++j) cp.p[j] = const_cast<double*>(p.cols[b.c0 + j].in);
      EVSL_HIP(evsl_dev::pack_cols(p.op.A.n, b.B, cp, bw.X0.p, L.f32, c.stream));
      const evsl_dev::Csr& A = p.op.A;
      const int nblk = A.nblk[wslot];
      evsl_dev::ChebGroupEntry e{};
      e.ia = A.ia;
      e.ja = A.ja;
      e.jd = A.jd;
      e.a = L.f32 ? (const void*)A.a32 : (const void*)A.a64;
      e.rbp = A.rb[wslot];
      e.x0 = bw.X0.p;
      e.ring = bw.ring.p;
      e.y = bw.Y.as<double>();
      e.mus = bw.mus.as<double>();
      e.t1 = p.op.t1;
      e.t2 = p.op.t2;
      e.cc = p.op.cc;
      e.n = p.op.A.n;
      e.blk0 = blk0;
      e.deg = b.deg;
      if (!e.a || !e.rbp) return EVSL_ERR_ARG;
      tab.push_back(e);
      owner.insert(owner.end(), nblk, (int)(tab.size() - 1 - L.tab0));
      blk0 += nblk;
      L.deg.push_back(b.deg);
      L.blk_end.push_back(blk0);
      maxdeg = std::max(maxdeg, b.deg);
    }
    launches.push_back(std::move(L));
  }
  // ---- tables to the device (the owner map only when it changed) ----------
  EVSL_HIP(W.d_tab.alloc(tab.size() * sizeof(evsl_dev::ChebGroupEntry)));
  EVSL_HIP(hipMemcpy(W.d_tab.p, tab.data(), tab.size() * sizeof(evsl_dev::ChebGroupEntry),
                     hipMemcpyHostToDevice));
  if (owner != W.owner_h) {
    EVSL_HIP(W.d_owner.alloc(owner.size() * sizeof(int)));
    EVSL_HIP(hipMemcpy(W.d_owner.p, owner.data(), owner.size() * sizeof(int),
                       hipMemcpyHostToDevice));
    W.owner_h = owner;
  }
  // ---- one grouped launch per degree and class ------------------------------
  const auto* d_tab = W.d_tab.as<evsl_dev::ChebGroupEntry>();
  const int* d_owner = W.d_owner.as<int>();
  for (int k = 1; k <= maxdeg; ++k) {
    for (const Launch& L : launches) {
      // entries with deg >= k form a prefix
      int ne = 0;
      while (ne < (int)L.deg.size() && L.deg[ne] >= k) ++ne;
      if (ne == 0) continue;
      const int total = L.blk_end[ne - 1];
      EVSL_HIP(evsl_dev::cheb_group_step(L.f32, L.B, L.i16, d_tab + L.tab0, d_owner + L.own0,
                                         total, k, c.stream));
      const size_t vs = L.f32 ? 4 : 8;
      for (int i = 0; i < ne; ++i) {
        const evsl_dev::ChebGroupEntry& e = tab[L.tab0 + i];
        const int yf = evsl_dev::y_flush(k, e.deg);
        c.astats.filter_bytes += 4.0 * (e.n + 1) +
                                 evsl_dev::cheb_vec_bytes(yf, k == 1, (int)vs) * L.B * e.n;
      }
      c.astats.n_filter_steps += 1;
    }
  }
  // matrix bytes per launch (each entry's CSR once per degree of its series)
  for (const Block& b : blocks) {
    const GroupedCols& p = *b.part;
    c.astats.filter_bytes +=
        (double)b.deg * p.op.A.nnz * ((p.op.f32 ? 4 : 8) + (p.op.A.jd ? 2 : 4));
    c.astats.n_filter_cols += (size_t)b.deg * b.B;
  }
  // ---- outputs ---------------------------------------------------------------
  for (const Block& b : blocks) {
    const GroupedCols& p = *b.part;
    evsl_dev::ColPtrs cp{};
    for (int j = 0; j < b.B; ++j) cp.p[j] = p.cols[b.c0 + j].out;
    EVSL_HIP(evsl_dev::unpack_cols(p.op.A.n, b.B, W.bufs[b.buf]->Y.p, false, cp, c.stream));
  }
  return 0;
}

}  // namespace evsl
