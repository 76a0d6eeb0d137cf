// spmm_kernel.h -- the fused Chebyshev-step CSR SpMM kernel (device code),
// its explicit-layout launcher and the fixed layout policy.  Included by the
// HIP translation units only: kernels.hip (fixed-policy launches, grouped
// KPM) and spmm_layouts.hip, built once per precision (every candidate layout
// of the per-matrix layout tuner).  The kernel and launchers sit in an
// unnamed namespace, so each unit holds its own instances; only the two
// explicit-layout entry points at the end are shared.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dev.h"

namespace evsl_dev {
namespace {

constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int xcd = bid % kXcds, pos = bid / kXcds;
  const int per = nblk / kXcds, rem = nblk % kXcds;
  return xcd * per + min(xcd, rem) + pos;
}

// ---------------------------------------------------------------------
// fused Chebyshev step SpMM
// ---------------------------------------------------------------------
// V contiguous values of one row of an interleaved block (one vector load)
template <typename VT, int V>
struct alignas(sizeof(VT) * V) VecT {
  VT v[V];
};

// Lane layout: L = B/V lanes share a row, each owning V adjacent columns
// (c0 = sub*V), so a block of 256 lanes covers NG = 256/L rows and every x
// access is one V-wide vector load.  LDS = true stages the block's ja/a
// through LDS with coalesced loads; LDS = false reads them from global memory
// (the L lanes of a row load the same address).  The per-column summation
// order is the same for every (V, LDS), so results are bitwise independent
// of the variant (the KPM block partials aside, whose tree depends on NG).
// NT marks the streams a launch touches last -- xkm1, the flushed y and the
// xkp1 store -- non-temporal, so that they do not push the gathered x rows
// out of L2; it pays when the launch's working set exceeds the Infinity
// Cache (chosen per launch, streaming_launch below) and leaves the
// arithmetic unchanged.
template <typename T, int V>
struct ExtVec {
  typedef T type __attribute__((ext_vector_type(V)));
};
template <typename T>
struct ExtVec<T, 1> {
  typedef T type;
};
// Column indices of the filter kernels, by index mode IX:
//   0: ja (int32 columns), 1: jd (16-bit offsets from the row),
//   2, 3: pattern codes (Csr::pc), their table offsets staged as 16-bit (2)
//   or 32-bit (3) offsets from the row.
// Pattern tables of one precision (null pc: not in use).
struct PatArgs {
  const unsigned char* pc;
  const int* off;
  const void* val;
  const void* dg;
  const unsigned char* dpos;  // position of the diagonal in its row (kPatDiag: none)
};

// The work of one row block `blk`: shared by the one-matrix kernel
// (k_cheb_spmm) and the grouped kernels (k_kpm_group, k_cheb_group), so that
// a block's arithmetic is the same in all of them.
template <typename VT, int B, int V, int MODE, bool LDS, int IX, bool NT>
__device__ __forceinline__ void cheb_spmm_block(
    const int blk, const int* __restrict__ ia, const int* __restrict__ ja,
    const short* __restrict__ jd, const VT* __restrict__ av, const int2* __restrict__ rbp,
    const VT* __restrict__ xk, const VT* __restrict__ xkm1, VT* __restrict__ xkp1,
    double* __restrict__ y, const VT* __restrict__ v0, const double* __restrict__ scol,
    const double* __restrict__ mu0col, double t_, double cc_, int first, int yflush,
    double* __restrict__ part, int* s_ia = nullptr, PatArgs pat = PatArgs{}) {
  static_assert(B % V == 0, "V must divide B");
  static_assert(IX >= 0 && IX <= 3, "index mode");
  static_assert(IX < 2 || LDS, "pattern codes are decoded at staging");
  constexpr bool kRel = IX != 0;  // staged columns are offsets from the row
  constexpr bool kPat = IX >= 2;  // one-byte pattern codes
  const VT* __restrict__ pval = static_cast<const VT*>(pat.val);
  const VT* __restrict__ pdg = static_cast<const VT*>(pat.dg);
  constexpr int L = B / V;
  constexpr int NG = kWG / L;  // lanes >= NG*L idle when L does not divide 256
  constexpr int U = (V * sizeof(VT) >= 32) ? 4 : 8;  // gathers in flight per lane
  using Vec = VecT<VT, V>;
  using VecY = VecT<double, V>;
  // staging area for ja/a, reused for the group reductions once the row
  // walk is over
  // column indices: int32, or 16-bit offsets from the row, ja = row + jd
  // (the pattern modes stage their table offsets the same way)
  using IT = typename std::conditional<IX == 1 || IX == 2, short, int>::type;
  const IT* __restrict__ jx = nullptr;
  if constexpr (IX == 1)
    jx = jd;
  else if constexpr (IX == 0)
    jx = ja;
  constexpr size_t kStage = LDS ? (size_t)kNnzCap * (sizeof(IT) + sizeof(VT)) : 0;
  constexpr size_t kRed = (size_t)kWG * V * sizeof(double);
  constexpr size_t kLong = (size_t)kLongStripes * B * sizeof(double);
  constexpr size_t kSmem0 = kStage > kRed ? kStage : kRed;
  __shared__ __align__(16) unsigned char smem[kSmem0 > kLong ? kSmem0 : kLong];
  IT* s_ja = reinterpret_cast<IT*>(smem);
  VT* s_a = reinterpret_cast<VT*>(smem + (size_t)kNnzCap * sizeof(IT));
  double* s_red = reinterpret_cast<double*>(smem);

  // block table entries carry (first row, its row pointer): one load for both ends
  const int2 e0 = rbp[blk], e1 = rbp[blk + 1];
  const int r0 = e0.x, r1 = e1.x, p0 = e0.y, p1 = e1.y;
  const int nz = p1 - p0;
  const int tid = threadIdx.x;
  const int grp = tid / L, sub = tid % L, c0 = sub * V;
  const bool act = tid < NG * L;
  const VT t = (VT)t_, cc = (VT)cc_;
  // kRes: v0 holds the residual block R and scol the per-column factors
  // 2*T_k(theta_c); mu0col[0..1] = (mu_k, mu_{k-1}) shared by all columns.
  // y accumulates the last yflush terms T_{k-1}, T_k, T_{k+1} of the series
  // (0: y untouched); s, s1, s2 are their coefficients.  kCheb keeps the
  // per-degree coefficient rows B apart (scol - B, scol - 2B), kRes the
  // (mu_k, mu_1) pairs 2 apart (mu0col - 2, mu0col - 4).  y and its
  // coefficients are FP64 whatever the vector type: with FP32 vectors the
  // sum of a long series would otherwise carry the rounding of every one
  // of its additions, noise that the filter does not damp (each term is
  // widened to FP64 and added there).
  // With V > 1, kCheb's per-column coefficients live in LDS (s_cf: s, s1,
  // s2, mu0 for the B columns), read by the epilogue: held in registers
  // over the whole row walk they take 4 V FP64 registers per lane, which
  // set the occupancy of the V = 2 and V = 4 layouts (76 -> 60 and 126 -> 94
  // VGPRs at 8 FP32 columns); at V = 1 the 8 registers cost no occupancy
  // and the LDS reads in every row's epilogue would (the FP64 one-column
  // step ran 14% slower).  kRes's are the same for every column (scalars).
  // Thread c < B loads column c's four here and publishes them to s_cf
  // just before the block's first barrier, so that the loads overlap the
  // staging loads instead of waiting ahead of them.
  constexpr bool kCfLds = MODE == kCheb && V > 1;
  __shared__ double s_cf[kCfLds ? 4 * B : 1];
  double cfv[4] = {0.0, 0.0, 0.0, 0.0};
  if (kCfLds && tid < B) {
    cfv[0] = scol[tid];
    if (yflush >= 2) cfv[1] = scol[tid - B];
    if (yflush >= 3) cfv[2] = scol[tid - 2 * B];
    if (first) cfv[3] = mu0col[tid];
  }
  auto publish_cf = [&]() {
    if (kCfLds && tid < B) {
#pragma unroll
      for (int i = 0; i < 4; ++i) s_cf[i * B + tid] = cfv[i];
    }
  };
  double s[V], s1[V], s2[V], mu0[V];
  VT g[V];
#pragma unroll
  for (int j = 0; j < V; ++j) {
    s[j] = s1[j] = s2[j] = 0;
    mu0[j] = 0;
    g[j] = 0;
    if (MODE == kCheb && !kCfLds) {
      s[j] = scol[c0 + j];
      if (yflush >= 2) s1[j] = scol[c0 + j - B];
      if (yflush >= 3) s2[j] = scol[c0 + j - 2 * B];
      if (first) mu0[j] = mu0col[c0 + j];
    } else if (MODE == kRes) {
      g[j] = (VT)scol[c0 + j];
      s[j] = mu0col[0];
      if (yflush >= 2) s1[j] = mu0col[-2];
      if (yflush >= 3) s2[j] = mu0col[-4];
      if (first) mu0[j] = mu0col[1];
    }
  }
  const bool ymode = MODE == kCheb || MODE == kRes;
  const bool need_xr = first || MODE == kKpm || cc_ != 0.0 || (ymode && yflush >= 2);
  double dacc[V];
#pragma unroll
  for (int j = 0; j < V; ++j) dacc[j] = 0.0;

  // epilogue operands of a row, loaded ahead of the gathers
  struct Ops {
    Vec xr, km1, rr, v;
    VecY yo;
  };
  auto ld = [](const VT* p, size_t i) { return *reinterpret_cast<const Vec*>(p + i); };
  // the same load / store, non-temporal when NT (vectors of VT or of double)
  auto ld_last = [](const auto* p, size_t i) {
    using ET = typename std::remove_cv<typename std::remove_reference<decltype(*p)>::type>::type;
    using RV = VecT<ET, V>;
    if constexpr (NT) {
      using EV = typename ExtVec<ET, V>::type;
      const EV e = __builtin_nontemporal_load(reinterpret_cast<const EV*>(p + i));
      RV r;
      __builtin_memcpy(&r, &e, sizeof(RV));
      return r;
    } else {
      return *reinterpret_cast<const RV*>(p + i);
    }
  };
  auto st_last = [](auto* p, size_t i, const auto& v) {
    using ET = typename std::remove_reference<decltype(*p)>::type;
    using RV = VecT<ET, V>;
    static_assert(sizeof(v) == sizeof(RV), "store width");
    if constexpr (NT) {
      using EV = typename ExtVec<ET, V>::type;
      EV e;
      __builtin_memcpy(&e, &v, sizeof(RV));
      __builtin_nontemporal_store(e, reinterpret_cast<EV*>(p + i));
    } else {
      *reinterpret_cast<RV*>(p + i) = v;
    }
  };
  auto prefetch = [&](int r) {
    Ops o{};
    const size_t i = (size_t)r * B + c0;
    if (MODE == kPlain) return o;
    if (need_xr) o.xr = ld(xk, i);
    if (xkm1) o.km1 = ld_last(xkm1, i);
    if (ymode && yflush > 0 && !first) o.yo = ld_last(y, i);
    if (MODE == kRes) o.rr = ld(v0, i);
    if (MODE == kKpm) o.v = ld(v0, i);
    return o;
  };
  auto epilogue = [&](int r, const VT* acc, const Ops& o) {
    const size_t i = (size_t)r * B + c0;
    Vec w;
#pragma unroll
    for (int j = 0; j < V; ++j) {
      if (MODE == kPlain) {
        w.v[j] = acc[j];
        continue;
      }
      VT wj = t * (acc[j] - cc * o.xr.v[j]) - o.km1.v[j];
      if (MODE == kRes) wj += g[j] * o.rr.v[j];
      w.v[j] = wj;
    }
    st_last(xkp1, i, w);
    if (ymode) {
      if (yflush > 0) {
        // the same additions, in the same order, as one update per degree
        VecY yn;
#pragma unroll
        for (int j = 0; j < V; ++j) {
          double cs = s[j], cs1 = s1[j], cs2 = s2[j], cm = mu0[j];
          if (kCfLds) {
            cs = s_cf[c0 + j];
            cs1 = s_cf[B + c0 + j];
            cs2 = s_cf[2 * B + c0 + j];
            cm = s_cf[3 * B + c0 + j];
          }
          double yv = first ? cm * (double)o.xr.v[j] : o.yo.v[j];
          if (yflush >= 3) yv = yv + cs2 * (double)o.km1.v[j];
          if (yflush >= 2) yv = yv + cs1 * (double)o.xr.v[j];
          yn.v[j] = yv + cs * (double)w.v[j];
        }
        st_last(y, i, yn);
      }
    } else if (MODE == kKpm) {
#pragma unroll
      for (int j = 0; j < V; ++j) dacc[j] += (double)w.v[j] * (double)o.v.v[j];
    }
  };

  if (nz <= kNnzCap) {
    // row pointers and epilogue operands of this lane's first row go out
    // with the staging loads
    const int rf = r0 + grp;
    int qf0 = 0, qf1 = 0;
    Ops of{};
    if (act && rf < r1) {
      qf0 = ia[rf] - p0;
      qf1 = ia[rf + 1] - p0;
      of = prefetch(rf);
    }
    const IT* jsrc = jx + p0;
    const VT* asrc = av + p0;
    // layouts whose lane groups walk several rows (R > 1) stage the block's
    // row pointers too, in the launch's dynamic LDS (s_ia, kRowsCap + 1
    // ints; a block holds at most kRowsCap rows), so each further row's
    // bounds come from LDS instead of two global loads ahead of its gathers
    // (2-9% at 8-32 columns on the 3-D Laplacian and Anderson matrices);
    // R = 1 launches pass no buffer and keep their LDS footprint
    const bool ia_lds = LDS && s_ia != nullptr && r1 - r0 <= kRowsCap;
    if (LDS) {
      // a block holds at most kNnzCap = kStageIt * kWG nonzeros and
      // kRowsCap + 1 <= 2 kWG row pointers: each thread issues all its
      // staging loads before it writes any of them to LDS, so that they are
      // in flight together (one memory round trip per block, not one per
      // kWG nonzeros)
      constexpr int kStageIt = kNnzCap / kWG;
      static_assert(kNnzCap % kWG == 0 && kRowsCap + 1 <= 2 * kWG, "staging shape");
      IT jv[kStageIt];
      VT avv[kStageIt];
      int iv[2] = {0, 0};
      // pattern modes: thread t also places row r0 + t's diagonal value
      int dsl = -1;
      VT dval = 0;
      if constexpr (kPat) {
        // one-byte codes, then their (offset, value) from the tables, which
        // the block first copies to LDS (their loads go out beside the
        // code loads); a diagonal code stages offset 0, and its value is
        // written by the thread of its row (position within the row from
        // dpos)
        __shared__ IT s_toff[kPatCodes + 1];
        __shared__ VT s_tval[kPatCodes + 1];
        unsigned char cv[kStageIt];
#pragma unroll
        for (int u = 0; u < kStageIt; ++u) {
          const int p = tid + u * kWG;
          cv[u] = p < nz ? pat.pc[p0 + p] : kPatDiag;
        }
        static_assert(kPatCodes + 1 <= kWG, "one table entry per thread");
        static_assert(kRowsCap <= kWG, "one row of the block per thread");
        if (tid <= kPatCodes) {
          s_toff[tid] = (IT)pat.off[tid];
          s_tval[tid] = pval[tid];
        }
        if (tid < r1 - r0) {
          const int r = r0 + tid;
          const unsigned char dp = pat.dpos[r];
          if (dp != kPatDiag) {
            dsl = ia[r] - p0 + dp;
            dval = pdg[r];
          }
        }
        __syncthreads();  // the tables
#pragma unroll
        for (int u = 0; u < kStageIt; ++u) {
          const bool d = cv[u] == kPatDiag;
          jv[u] = d ? (IT)0 : s_toff[cv[u]];
          avv[u] = d ? (VT)0 : s_tval[cv[u]];
        }
      } else {
#pragma unroll
        for (int u = 0; u < kStageIt; ++u) {
          const int p = tid + u * kWG;
          if (p < nz) {
            jv[u] = jx[p0 + p];
            avv[u] = av[p0 + p];
          }
        }
      }
      if (ia_lds) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (tid + u * kWG <= r1 - r0) iv[u] = ia[r0 + tid + u * kWG];
      }
#pragma unroll
      for (int u = 0; u < kStageIt; ++u) {
        const int p = tid + u * kWG;
        if (p < nz) {
          s_ja[p] = jv[u];
          if (!kPat || jv[u] != (IT)0) s_a[p] = avv[u];
        }
      }
      if (kPat && dsl >= 0) s_a[dsl] = dval;
      if (ia_lds) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (tid + u * kWG <= r1 - r0) s_ia[tid + u * kWG] = iv[u] - p0;
      }
      publish_cf();
      __syncthreads();
      jsrc = s_ja;
      asrc = s_a;
    } else if (kCfLds) {
      publish_cf();
      __syncthreads();
    }
    for (int r = rf; act && r < r1; r += NG) {
      int q0 = qf0, q1 = qf1;
      Ops o = of;
      if (r != rf) {
        if (ia_lds) {
          q0 = s_ia[r - r0];
          q1 = s_ia[r - r0 + 1];
        } else {
          q0 = ia[r] - p0;
          q1 = ia[r + 1] - p0;
        }
        o = prefetch(r);
      }
      VT acc[V];
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] = 0;
      // U independent gathers in flight per lane; zero padding leaves the
      // sequential summation order (and its rounding) unchanged
      for (int q = q0; q < q1; q += U) {
        int jj[U];
        VT aa[U];
        Vec xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const bool ok = q + u < q1;
          jj[u] = ok ? (kRel ? r : 0) + (int)jsrc[q + u] : 0;
          aa[u] = ok ? asrc[q + u] : (VT)0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (q + u < q1) {
            xx[u] = ld(xk, (size_t)jj[u] * B + c0);
          } else {
#pragma unroll
            for (int j = 0; j < V; ++j) xx[u].v[j] = (VT)0;
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < V; ++j) acc[j] += aa[u] * xx[u].v[j];
      }
      epilogue(r, acc, o);
    }
    if (MODE == kKpm && LDS) __syncthreads();  // s_red aliases the staging area
  } else {
    // one long row, split into kLongStripes fixed stripes (stripe s sums the
    // nonzeros p0 + s, p0 + s + kLongStripes, ... in order), whose sums are
    // then added in stripe order in double: the same arithmetic per column
    // for every lane layout and block width, so tuning and batching keep
    // their bits on such rows too.  Lane group g takes stripes g, g + NG, ...
    double* s_str = reinterpret_cast<double*>(smem);  // kLongStripes x B
    for (int st = grp; act && st < kLongStripes; st += NG) {
      VT acc[V];
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] = 0;
      for (int p = p0 + st; p < p1; p += kLongStripes) {
        VT a;
        int col;
        if constexpr (kPat) {
          const unsigned char cd = pat.pc[p];
          const bool d = cd == kPatDiag;
          a = d ? pdg[r0] : pval[cd];
          col = r0 + (d ? 0 : pat.off[cd]);
        } else {
          a = av[p];
          col = (kRel ? r0 : 0) + (int)jx[p];
        }
        const Vec x = ld(xk, (size_t)col * B + c0);
#pragma unroll
        for (int j = 0; j < V; ++j) acc[j] += a * x.v[j];
      }
#pragma unroll
      for (int j = 0; j < V; ++j) s_str[st * B + c0 + j] = (double)acc[j];
    }
    publish_cf();
    __syncthreads();
    if (act && grp == 0) {
      VT tot[V];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        double sj = 0.0;
#pragma unroll 1
        for (int st = 0; st < kLongStripes; ++st) sj += s_str[st * B + c0 + j];
        tot[j] = (VT)sj;
      }
      epilogue(r0, tot, prefetch(r0));
    }
    __syncthreads();
  }

  if (MODE == kKpm) {  // L is a power of two here (NG*L == kWG)
    static_assert(MODE != kKpm || (kWG % L) == 0, "KPM blocks need B/V | 256");
#pragma unroll
    for (int j = 0; j < V; ++j) s_red[tid * V + j] = dacc[j];
    __syncthreads();
    for (int st = NG / 2; st > 0; st >>= 1) {
      if (grp < st) {
#pragma unroll
        for (int j = 0; j < V; ++j) s_red[tid * V + j] += s_red[(tid + st * L) * V + j];
      }
      __syncthreads();
    }
    if (grp == 0) {
#pragma unroll
      for (int j = 0; j < V; ++j) part[(size_t)blk * B + c0 + j] = s_red[tid * V + j];
    }
  }
}

template <typename VT, int B, int V, int MODE, bool LDS, int IX, bool NT = false>
__global__ __launch_bounds__(kWG) void k_cheb_spmm(
    const int* __restrict__ ia, const int* __restrict__ ja, const short* __restrict__ jd,
    const VT* __restrict__ av,
    const int2* __restrict__ rbp, int nblk, const VT* __restrict__ xk,
    const VT* __restrict__ xkm1, VT* __restrict__ xkp1, double* __restrict__ y,
    const VT* __restrict__ v0, const double* __restrict__ scol,
    const double* __restrict__ mu0col, double t_, double cc_, int first, int yflush,
    double* __restrict__ part, int stage_ia, PatArgs pat) {
  extern __shared__ int s_dyn[];  // kRowsCap + 1 ints when stage_ia
  cheb_spmm_block<VT, B, V, MODE, LDS, IX, NT>(xcd_remap(blockIdx.x, nblk), ia, ja, jd, av, rbp,
                                               xk, xkm1, xkp1, y, v0, scol, mu0col, t_, cc_,
                                               first, yflush, part, stage_ia ? s_dyn : nullptr,
                                               pat);
}

// the index mode of a matrix's filter launches at width B
inline int index_mode(const Csr& A, int B, int vsize) {
  return pattern_launch(A, B, vsize) ? (A.pshort ? 2 : 3) : (A.jd ? 1 : 0);
}

template <typename VT>
inline PatArgs pat_args(const Csr& A, int B) {
  if (!pattern_launch(A, B, (int)sizeof(VT))) return PatArgs{};
  if constexpr (sizeof(VT) == 4)
    return PatArgs{A.pc, A.poff, A.pval32, A.dg32, A.pdpos};
  else
    return PatArgs{A.pc, A.poff, A.pval64, A.dg64, A.pdpos};
}

// Non-temporal streams (the NT instances) exist for the ChebAv and residual
// steps of 8 or more columns.  They are chosen when one launch's working set
// -- the matrix, and the vector blocks xk, xkm1, xkp1 (and R) plus the FP64
// y -- exceeds 320 MB: beyond the 256 MB Infinity Cache the block no longer stays
// resident from launch to launch, and keeping the last-use streams out of L2
// saves the x rows re-read across planes.  Measured on the 100^3 Laplacian
// (µs per launch, plain -> non-temporal): FP32 32-column ChebAv 137.5 -> 124.8
// and residual step 149.0 -> 141.4, FP32 16-column residual 81.9 -> 71.4, FP64
// 8/16/32-column ChebAv 71.5 -> 67.8, 131.5 -> 118.3, 242.2 -> 232.5, FP64
// 8-column residual 85.0 -> 72.7; below the threshold the hints cost 3-30%
// (FP32 8- and 16-column ChebAv at 174 / 302 MB, FP64 4-column at 197 MB).
template <int MODE, int B>
constexpr bool kStreamingCapable = (MODE == kCheb || MODE == kRes) && B >= 8;

inline bool streaming_launch(const Csr& A, int B, int vsize, int mode) {
  if (!((mode == kCheb || mode == kRes) && B >= 8)) return false;
  // xk, xkm1, xkp1 (and R) of the vector type, y in FP64
  const double streams = mode == kRes ? 4.0 : 3.0;
  const double mat = pattern_launch(A, B, vsize) ? (double)A.nnz + (double)A.n * (vsize + 1)
                          : (double)A.nnz * (vsize + (A.jd ? 2 : 4));
  const double ws = streams * (double)A.n * B * vsize + 8.0 * A.n * B + mat + 4.0 * (A.n + 1);
  return ws > 320e6;
}

// One fused step with an explicit lane layout (V columns per lane, LDS
// staging on/off, R rows per lane group); the row-block table is the one
// built for B/V/R lanes per row.
template <typename VT, int B, int V, int MODE, bool LDS, int R = 1>
hipError_t launch_variant(const Csr& A, const void* xk, const void* xkm1, void* xkp1, void* y,
                          const void* v0, const double* s, const double* mu0, double t,
                          double cc, bool first, int yflush, double* part, hipStream_t st) {
  const VT* av;
  if constexpr (sizeof(VT) == 4)
    av = A.a32;
  else
    av = A.a64;
  if (!av) return hipErrorInvalidValue;
  static_assert((B / V) % R == 0, "R must divide the lanes per row");
  const int slot = width_slot(B / V / R);
  if (slot < 0 || !A.rb[slot]) return hipErrorInvalidValue;
  const int nblk = A.nblk[slot];
  const int2* rb = reinterpret_cast<const int2*>(A.rb[slot]);
  const int f = first ? 1 : 0;
  // lane groups walk R rows: the row pointers go to dynamic LDS
  const int stage_ia = (LDS && R > 1) ? 1 : 0;
  const size_t dyn = stage_ia ? (size_t)(kRowsCap + 1) * sizeof(int) : 0;
  const PatArgs pat = pat_args<VT>(A, B);
  if (pat.pc && (!pat.val || !pat.dg)) return hipErrorInvalidValue;
  const int ix = index_mode(A, B, (int)sizeof(VT));
#define EVSL_SPMM_LAUNCH(IX_, NT_)                                                             \
  hipLaunchKernelGGL((k_cheb_spmm<VT, B, V, MODE, LDS, IX_, NT_>), dim3(nblk), dim3(kWG), dyn, \
                     st, A.ia, A.ja, A.jd, av, rb, nblk, (const VT*)xk, (const VT*)xkm1,     \
                     (VT*)xkp1, (double*)y, (const VT*)v0, s, mu0, t, cc, f, yflush, part,    \
                     stage_ia, pat)
#define EVSL_SPMM_BY_IX(NT_)    \
  switch (ix) {                 \
    case 3:                     \
      EVSL_SPMM_LAUNCH(3, NT_); \
      break;                    \
    case 2:                     \
      EVSL_SPMM_LAUNCH(2, NT_); \
      break;                    \
    case 1:                     \
      EVSL_SPMM_LAUNCH(1, NT_); \
      break;                    \
    default:                    \
      EVSL_SPMM_LAUNCH(0, NT_); \
  }
  if constexpr (kStreamingCapable<MODE, B>) {
    if (streaming_launch(A, B, (int)sizeof(VT), MODE)) {
      EVSL_SPMM_BY_IX(true)
      return hipGetLastError();
    }
  }
  EVSL_SPMM_BY_IX(false)
#undef EVSL_SPMM_BY_IX
#undef EVSL_SPMM_LAUNCH
  return hipGetLastError();
}

// Layout of the fused filter kernel for each (value size, block width,
// mode), as 10*V + R: V columns per lane and R rows per lane group.  With
// V > 1 a lane takes V adjacent columns, so a block of 256 lanes covers more
// than 256/B rows; with R > 1 a block covers R times the rows its lane groups
// can hold at once and each group walks R of them, amortising the block's
// table loads, staging and barrier (the row-block table is then the one built
// for (B/V)/R lanes per row).  Chosen by a sweep of every valid (V, R) on the
// 7-point 100^3 Laplacian on gfx950 (LDS staging won for every width); odd
// widths only admit V = R = 1.  Time saved by the chosen R over the best
// R = 1 layout: FP32 ChebAv 7% at 4 columns, 9% at 8, 15% at 16, 19% at 32;
// FP64 ChebAv 2% at 4, 7% at 8, 11% at 16 and 32; the FP32 residual step
// 7-8% at 2, 4, 16 and 32.  The FP32 KPM and FP64 residual rows (neither is
// on a solver's hot path) keep the layouts of an earlier V-only sweep.
//                                    B =  2   4   6   8  16  32
constexpr int kSpmmLayout[2][3][6] = {{{21, 22, 21, 22, 44, 44},    // FP32 ChebAv
                                       {21, 21, 21, 21, 21, 41},    // FP32 KPM
                                       {12, 14, 21, 22, 44, 44}},   // FP32 residual
                                      {{21, 12, 21, 22, 24, 24},    // FP64 ChebAv
                                       {11, 12, 21, 41, 41, 41},    // FP64 KPM
                                       {21, 11, 21, 22, 21, 22}}};  // FP64 residual

constexpr int spmm_layout(int vsize, int B, int mode) {
  const int w = B == 2 ? 0 : B == 4 ? 1 : B == 6 ? 2 : B == 8 ? 3 : B == 16 ? 4 : B == 32 ? 5 : -1;
  return w < 0 ? 11 : kSpmmLayout[vsize == 8][mode - kCheb][w];
}
constexpr int spmm_vec(int vsize, int B, int mode) { return spmm_layout(vsize, B, mode) / 10; }
constexpr int spmm_rows(int vsize, int B, int mode) { return spmm_layout(vsize, B, mode) % 10; }

template <typename VT, int B, int MODE>
struct SpmmPolicy {
  static constexpr int V = spmm_vec((int)sizeof(VT), B, MODE == kPlain ? kCheb : MODE);
  static constexpr int R = spmm_rows((int)sizeof(VT), B, MODE == kPlain ? kCheb : MODE);
  static constexpr bool LDS = true;
};

}  // namespace

// explicit-layout launches (spmm_layouts.hip, one object per precision)
hipError_t cheb_spmm_layout_f32(int layout, const Csr& A, int B, Mode mode, const void* xk,
                                const void* xkm1, void* xkp1, void* y, const void* v0,
                                const double* s, const double* mu0, double t, double cc,
                                bool first, int yflush, double* part, hipStream_t st);
hipError_t cheb_spmm_layout_f64(int layout, const Csr& A, int B, Mode mode, const void* xk,
                                const void* xkm1, void* xkp1, void* y, const void* v0,
                                const double* s, const double* mu0, double t, double cc,
                                bool first, int yflush, double* part, hipStream_t st);

}  // namespace evsl_dev
