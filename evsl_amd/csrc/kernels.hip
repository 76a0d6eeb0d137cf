// kernels.hip -- CDNA4 (gfx950) kernels of the filtered-Lanczos / KPM path.
//
// Design notes (wave64, 256-thread workgroups):
//  * cheb_spmm (kernel in spmm_kernel.h; fixed-policy launches and the
//    grouped KPM launch here, the tuner's candidate layouts in
//    spmm_layouts.hip): CSR row blocks are staged through LDS with coalesced
//    loads of ja/a, then each (row, column) lane walks its row in LDS and
//    gathers the B contiguous values of x it needs.  The Chebyshev three-term
//    recurrence (and either the ChebAv accumulation or the KPM moment dot) is
//    fused into the epilogue, so A is read once per degree for all B columns.
//    Workgroup ids are remapped so that each of the 8 XCDs sweeps a
//    contiguous range of rows (its own L2 then holds the neighbouring x
//    segments it re-reads).
//  * Reductions are two-level and deterministic: per-block partials, then a
//    separate small launch sums them in a fixed order.
//  * Tall-skinny Gram-Schmidt: V^T w by column groups of 8 with per-chunk
//    partials; V h with h staged in LDS.  Both stream V once (HBM-bound).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dev.h"
#include "spmm_kernel.h"

namespace evsl_dev {
namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// sum over a 256-thread block; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh /*>=4*/) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

// KPM steps of several matrices in one launch: global row block gb (after
// the XCD remap) belongs to matrix owner[gb], whose entry names its CSR, its
// row-block table, its vector blocks and its recurrence factors; step k
// reads T_k and T_{k-1} from the matrix's ring and writes T_{k+1} and the
// block's moment partials, as the one-matrix kKpm launch does.
template <typename VT, int B, int V>
__global__ __launch_bounds__(kWG) void k_kpm_group(const KpmGroupEntry* __restrict__ tab,
                                                   const int* __restrict__ owner, int total,
                                                   int k) {
  const int gb = xcd_remap(blockIdx.x, total);
  const KpmGroupEntry& e = tab[owner[gb]];
  k -= e.k0;  // this matrix's own step
  const size_t nb = (size_t)e.n * B;
  const VT* x0 = static_cast<const VT*>(e.x0);
  VT* ring = static_cast<VT*>(e.ring);
  const VT* xk = k == 0 ? x0 : ring + (size_t)((k - 1) % 3) * nb;
  const VT* xkm1 = k == 0 ? nullptr : (k == 1 ? x0 : ring + (size_t)((k - 2) % 3) * nb);
  VT* xkp1 = ring + (size_t)(k % 3) * nb;
  cheb_spmm_block<VT, B, V, kKpm, true, 0, false>(
      gb - e.blk0, e.ia, e.ja, nullptr, static_cast<const VT*>(e.a),
      reinterpret_cast<const int2*>(e.rbp), xk, xkm1, xkp1, nullptr, x0, nullptr, nullptr,
      k == 0 ? e.t0 : e.t, e.cc, 0, 0, e.part);
}

// the y-flush schedule of evsl_dev::y_flush, for the grouped launch (each
// entry flushes by its own series degree m)
__device__ __forceinline__ int y_flush_dev(int k, int m) {
  if (k == 1) return 1;
  const int r = (k - 1) % 3;
  if (r == 0) return 3;
  return k == m ? r : 0;
}

// ChebAv steps of several column blocks (of different matrices) in one
// launch: global row block gb (after the XCD remap) belongs to entry
// owner[gb]; the block runs the one-matrix kCheb step at k with that entry's
// operands, in the lane layout of the fixed policy, so every column gets the
// bits of its one-matrix filter.
template <typename VT, int B, bool I16>
__global__ __launch_bounds__(kWG) void k_cheb_group(const ChebGroupEntry* __restrict__ tab,
                                                    const int* __restrict__ owner, int total,
                                                    int k) {
  extern __shared__ int s_dyn[];  // the block's row pointers when R > 1
  using P = SpmmPolicy<VT, B, kCheb>;
  const int gb = xcd_remap(blockIdx.x, total);
  const ChebGroupEntry& e = tab[owner[gb]];
  const size_t nb = (size_t)e.n * B;
  const VT* x0 = static_cast<const VT*>(e.x0);
  VT* ring = static_cast<VT*>(e.ring);
  VT* xkp1 = ring + (size_t)((k - 1) % 3) * nb;
  const VT* xk = k == 1 ? x0 : ring + (size_t)((k - 2) % 3) * nb;
  const VT* xkm1 = k == 1 ? nullptr : (k == 2 ? x0 : ring + (size_t)((k - 3) % 3) * nb);
  cheb_spmm_block<VT, B, P::V, kCheb, true, I16 ? 1 : 0, false>(
      gb - e.blk0, e.ia, e.ja, e.jd, static_cast<const VT*>(e.a),
      reinterpret_cast<const int2*>(e.rbp), xk, xkm1, xkp1, e.y, nullptr,
      e.mus + (size_t)k * B, e.mus, k == 1 ? e.t1 : e.t2, e.cc, k == 1, y_flush_dev(k, e.deg),
      nullptr, P::R > 1 ? s_dyn : nullptr);
}

// out[c] = sum over blocks b in [b0, b1) of part[b*B + c] for block `bid` of
// `nb` (the range split evenly over the nb blocks; out = scratch of nb * B
// values when nb > 1)
__device__ __forceinline__ void sum_partials_block(const double* __restrict__ part, int nblk,
                                                   int B, double* __restrict__ out, int bid,
                                                   int nb, double* s_red) {
  const int NG = kWG / B;
  const int col = threadIdx.x % B, grp = threadIdx.x / B;
  const bool act = (int)threadIdx.x < NG * B;
  const int per = (nblk + nb - 1) / nb;
  const int b0 = bid * per, b1 = min(nblk, b0 + per);
  double a = 0.0;
  if (act)
    for (int b = b0 + grp; b < b1; b += NG) a += part[(size_t)b * B + col];
  s_red[threadIdx.x] = a;
  __syncthreads();
  if (act && grp == 0) {
    double t = 0.0;
    for (int g = 0; g < NG; ++g) t += s_red[g * B + col];
    out[(size_t)bid * B + col] = t;
  }
}

__global__ __launch_bounds__(kWG) void k_sum_block_partials(const double* __restrict__ part,
                                                            int nblk, int B,
                                                            double* __restrict__ out) {
  __shared__ double s_red[kWG];
  sum_partials_block(part, nblk, B, out, blockIdx.x, gridDim.x, s_red);
}

// The partial sums of a grouped KPM step, per matrix in the order of
// sum_block_partials: level 1 runs entry owner[blockIdx] as block
// blockIdx - first[owner] of its G; level 2 sums the G first-level values of
// each two-level entry.
__global__ __launch_bounds__(kWG) void k_kpm_group_sums(const KpmSumEntry* __restrict__ tab,
                                                        const int* __restrict__ owner,
                                                        const int* __restrict__ first, int B,
                                                        int k, int level) {
  __shared__ double s_red[kWG];
  const int r = owner[blockIdx.x];
  const KpmSumEntry& e = tab[r];
  double* out = e.dots + (size_t)(k - e.k0 + 1) * B;
  if (level == 1) {
    if (e.G == 1)
      sum_partials_block(e.part, e.nblk, B, out, 0, 1, s_red);
    else
      sum_partials_block(e.part, e.nblk, B, e.scratch, blockIdx.x - first[r], e.G, s_red);
  } else {
    sum_partials_block(e.scratch, e.G, B, out, 0, 1, s_red);
  }
}

// ---------------------------------------------------------------------
// BLAS-1 style kernels with fixed-grid partial reductions
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kWG) void k_dot_partial(size_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     double* __restrict__ part) {
  __shared__ double sh[4];
  double a = 0.0;
  for (size_t i = (size_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (size_t)gridDim.x * kWG)
    a += x[i] * y[i];
  a = block_sum(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ __launch_bounds__(kWG) void k_sum_scalar(const double* __restrict__ part, int cnt,
                                                    double* __restrict__ out) {
  __shared__ double sh[4];
  double a = 0.0;
  for (int i = threadIdx.x; i < cnt; i += kWG) a += part[i];
  a = block_sum(a, sh);
  if (threadIdx.x == 0) *out = a;
}

__global__ __launch_bounds__(kWG) void k_axpy_dot(size_t n, double a,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y,
                                                  const double* __restrict__ z,
                                                  double* __restrict__ part) {
  __shared__ double sh[4];
  double d = 0.0;
  for (size_t i = (size_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (size_t)gridDim.x * kWG) {
    const double v = y[i] + a * x[i];
    y[i] = v;
    if (z) d += v * z[i];
  }
  if (z) {
    d = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
  }
}

template <bool SELF>
__global__ __launch_bounds__(kWG) void k_axpy_dev(size_t n, double sign,
                                                  const double* __restrict__ a_d,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y,
                                                  const double* __restrict__ z,
                                                  double* __restrict__ part) {
  __shared__ double sh[4];
  const double a = sign * (*a_d);
  double d = 0.0;
  for (size_t i = (size_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (size_t)gridDim.x * kWG) {
    const double v = y[i] + a * x[i];
    y[i] = v;
    d += v * (SELF ? v : z[i]);
  }
  d = block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = d;
}

__global__ void k_scal(size_t n, double a, double* __restrict__ x) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

__global__ void k_scal_copy(size_t n, double a, const double* __restrict__ x,
                            double* __restrict__ y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}

__global__ void k_fill(size_t n, double v, double* __restrict__ x) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = v;
}

__global__ void k_set_identity(int k, double* __restrict__ Z) {
  const size_t tot = (size_t)k * k;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x)
    Z[i] = (i % (size_t)(k + 1) == 0) ? 1.0 : 0.0;
}

__global__ void k_fill_u32(size_t n, unsigned v, unsigned* __restrict__ x) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = v;
}

__global__ void k_f64_to_f32(size_t n, const double* __restrict__ x, float* __restrict__ y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

__global__ void k_f32_to_f64(size_t n, const float* __restrict__ x, double* __restrict__ y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}

// x *= a and x32 = (float)x (a basis column and its FP32 copy)
__global__ void k_scal_f32copy(size_t n, double a, double* __restrict__ x,
                               float* __restrict__ x32) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double v = x[i] * a;
    x[i] = v;
    x32[i] = (float)v;
  }
}

// ---------------------------------------------------------------------
// tall-skinny Gram-Schmidt kernels
// ---------------------------------------------------------------------
constexpr int kGtCols = 8;     // columns per workgroup in V^T w
constexpr int kGtRows = 4096;  // rows per chunk in V^T w

// part[c * nch + chunk] = sum over the chunk's rows of V(r, c) * w(r).
// One workgroup per (chunk, 8-column group) on a 1-D grid of nch * ncg: the
// XCD remap gives each XCD a contiguous range of that index, and the column
// group varies fastest, so the workgroups that read one chunk of w run on
// the same XCD together and w is fetched into its L2 once, not once per
// column group.
__global__ __launch_bounds__(kWG) void k_gemv_t_partial(size_t n, int k,
                                                        const double* __restrict__ V,
                                                        size_t ld,
                                                        const double* __restrict__ w,
                                                        double* __restrict__ part, int nch,
                                                        int ncg) {
  __shared__ double sh[4][kGtCols];
  const int lb = xcd_remap(blockIdx.x, nch * ncg);
  const int ch = lb / ncg;
  const int c0 = (lb % ncg) * kGtCols;
  const int nc = min(kGtCols, k - c0);
  double acc[kGtCols];
#pragma unroll
  for (int c = 0; c < kGtCols; ++c) acc[c] = 0.0;
  const size_t rbeg = (size_t)ch * kGtRows;
  const size_t rend = min(n, rbeg + kGtRows);
  const double* Vc = V + (size_t)c0 * ld;
  if (nc == kGtCols) {
    for (size_t r = rbeg + threadIdx.x; r < rend; r += kWG) {
      const double wr = w[r];
#pragma unroll
      for (int c = 0; c < kGtCols; ++c) acc[c] += Vc[(size_t)c * ld + r] * wr;
    }
  } else {
    for (size_t r = rbeg + threadIdx.x; r < rend; r += kWG) {
      const double wr = w[r];
      for (int c = 0; c < nc; ++c) acc[c] += Vc[(size_t)c * ld + r] * wr;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kGtCols; ++c) {
    const double v = wave_sum(acc[c]);
    if (lane == 0) sh[wid][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    const int c = threadIdx.x;
    part[(size_t)(c0 + c) * nch + ch] = sh[0][c] + sh[1][c] + sh[2][c] + sh[3][c];
  }
}

// h[c] = sum_chunk part[c*nch + chunk]; one 64-lane wave per column
__global__ __launch_bounds__(64) void k_gemv_t_reduce(const double* __restrict__ part,
                                                      int nch, double* __restrict__ h) {
  const int c = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < nch; i += 64) a += part[(size_t)c * nch + i];
  a = wave_sum(a);
  if (threadIdx.x == 0) h[c] = a;
}

constexpr int kHTile = 2048;  // h entries staged in LDS per pass

// w -= V h; part[blk] = partial ||w||^2   (SUB) or y = V h (!SUB)
template <bool SUB>
__global__ __launch_bounds__(kWG) void k_gemv_n(size_t n, int k, const double* __restrict__ V,
                                                size_t ld, const double* __restrict__ h,
                                                double* __restrict__ w,
                                                double* __restrict__ part) {
  __shared__ double s_h[kHTile];
  __shared__ double sh[4];
  const size_t r = (size_t)blockIdx.x * kWG + threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int cbase = 0; cbase < k; cbase += kHTile) {
    const int nc = min(kHTile, k - cbase);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += kWG) s_h[c] = h[cbase + c];
    __syncthreads();
    if (r < n) {
      const double* Vr = V + (size_t)cbase * ld + r;
      int c = 0;
      // two groups of four columns per trip: eight loads in flight, the
      // same additions in the same order
#pragma unroll 2
      for (; c + 4 <= nc; c += 4) {
        a0 += Vr[(size_t)c * ld] * s_h[c];
        a1 += Vr[(size_t)(c + 1) * ld] * s_h[c + 1];
        a2 += Vr[(size_t)(c + 2) * ld] * s_h[c + 2];
        a3 += Vr[(size_t)(c + 3) * ld] * s_h[c + 3];
      }
      for (; c < nc; ++c) a0 += Vr[(size_t)c * ld] * s_h[c];
    }
  }
  const double acc = (a0 + a1) + (a2 + a3);
  double d = 0.0;
  if (r < n) {
    if (SUB) {
      const double v = w[r] - acc;
      w[r] = v;
      d = v * v;
    } else {
      w[r] = acc;
    }
  }
  if (SUB) {
    d = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
  }
}

// The same two passes against an FP32 copy of the basis (V32, ld32 even,
// columns 8-byte aligned), products and sums in FP64: each thread takes a
// pair of adjacent rows with one 8-byte load per column, so a wave moves
// 512 bytes per load as the FP64 kernels do, and the basis is read at half
// their bytes.  A row pair that ends past n (odd n) takes its one row alone;
// rows >= n of V32 (padding) are never read.
__global__ __launch_bounds__(kWG) void k_gemv_t_partial_f32(size_t n, int k,
                                                            const float* __restrict__ V,
                                                            size_t ld,
                                                            const double* __restrict__ w,
                                                            double* __restrict__ part, int nch,
                                                            int ncg) {
  __shared__ double sh[4][kGtCols];
  const int lb = xcd_remap(blockIdx.x, nch * ncg);
  const int ch = lb / ncg;
  const int c0 = (lb % ncg) * kGtCols;
  const int nc = min(kGtCols, k - c0);
  double acc[kGtCols];
#pragma unroll
  for (int c = 0; c < kGtCols; ++c) acc[c] = 0.0;
  const size_t rbeg = (size_t)ch * kGtRows;  // even
  const size_t rend = min(n, rbeg + kGtRows);
  const float* Vc = V + (size_t)c0 * ld;
  for (size_t r = rbeg + 2 * threadIdx.x; r < rend; r += 2 * kWG) {
    const double w0 = w[r];
    if (r + 1 < rend) {
      const double w1 = w[r + 1];
      if (nc == kGtCols) {
#pragma unroll
        for (int c = 0; c < kGtCols; ++c) {
          const float2 v = *reinterpret_cast<const float2*>(Vc + (size_t)c * ld + r);
          acc[c] += (double)v.x * w0;
          acc[c] += (double)v.y * w1;
        }
      } else {
        for (int c = 0; c < nc; ++c) {
          const float2 v = *reinterpret_cast<const float2*>(Vc + (size_t)c * ld + r);
          acc[c] += (double)v.x * w0;
          acc[c] += (double)v.y * w1;
        }
      }
    } else {
      for (int c = 0; c < nc; ++c) acc[c] += (double)Vc[(size_t)c * ld + r] * w0;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kGtCols; ++c) {
    const double v = wave_sum(acc[c]);
    if (lane == 0) sh[wid][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    const int c = threadIdx.x;
    part[(size_t)(c0 + c) * nch + ch] = sh[0][c] + sh[1][c] + sh[2][c] + sh[3][c];
  }
}

// w -= V32 h; part[blk] = partial ||w||^2 over the block's 2 kWG rows
__global__ __launch_bounds__(kWG) void k_gemv_n_sub_f32(size_t n, int k,
                                                        const float* __restrict__ V, size_t ld,
                                                        const double* __restrict__ h,
                                                        double* __restrict__ w,
                                                        double* __restrict__ part) {
  __shared__ double s_h[kHTile];
  __shared__ double sh[4];
  const size_t r = 2 * ((size_t)blockIdx.x * kWG + threadIdx.x);
  const bool pair = r + 1 < n;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;  // row r, columns c .. c + 3
  double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;  // row r + 1
  for (int cbase = 0; cbase < k; cbase += kHTile) {
    const int nc = min(kHTile, k - cbase);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += kWG) s_h[c] = h[cbase + c];
    __syncthreads();
    if (pair) {
      const float* Vr = V + (size_t)cbase * ld + r;
      int c = 0;
      // two groups of four columns per trip: eight 8-byte loads in flight
#pragma unroll 2
      for (; c + 4 <= nc; c += 4) {
        const float2 v0 = *reinterpret_cast<const float2*>(Vr + (size_t)c * ld);
        const float2 v1 = *reinterpret_cast<const float2*>(Vr + (size_t)(c + 1) * ld);
        const float2 v2 = *reinterpret_cast<const float2*>(Vr + (size_t)(c + 2) * ld);
        const float2 v3 = *reinterpret_cast<const float2*>(Vr + (size_t)(c + 3) * ld);
        a0 += (double)v0.x * s_h[c];
        a1 += (double)v1.x * s_h[c + 1];
        a2 += (double)v2.x * s_h[c + 2];
        a3 += (double)v3.x * s_h[c + 3];
        b0 += (double)v0.y * s_h[c];
        b1 += (double)v1.y * s_h[c + 1];
        b2 += (double)v2.y * s_h[c + 2];
        b3 += (double)v3.y * s_h[c + 3];
      }
      for (; c < nc; ++c) {
        const float2 v0 = *reinterpret_cast<const float2*>(Vr + (size_t)c * ld);
        a0 += (double)v0.x * s_h[c];
        b0 += (double)v0.y * s_h[c];
      }
    } else if (r < n) {
      const float* Vr = V + (size_t)cbase * ld + r;
      int c = 0;
      for (; c + 4 <= nc; c += 4) {
        a0 += (double)Vr[(size_t)c * ld] * s_h[c];
        a1 += (double)Vr[(size_t)(c + 1) * ld] * s_h[c + 1];
        a2 += (double)Vr[(size_t)(c + 2) * ld] * s_h[c + 2];
        a3 += (double)Vr[(size_t)(c + 3) * ld] * s_h[c + 3];
      }
      for (; c < nc; ++c) a0 += (double)Vr[(size_t)c * ld] * s_h[c];
    }
  }
  double d = 0.0;
  if (r < n) {
    const double v = w[r] - ((a0 + a1) + (a2 + a3));
    w[r] = v;
    d = v * v;
  }
  if (pair) {
    const double v = w[r + 1] - ((b0 + b1) + (b2 + b3));
    w[r + 1] = v;
    d += v * v;
  }
  d = block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = d;
}

// ---------------------------------------------------------------------
// FP64 GEMM (column-major): C = A * B, 64x64 tiles, k-step 16
// ---------------------------------------------------------------------
constexpr int kTM = 64, kTN = 64, kTK = 16;

__global__ __launch_bounds__(kWG) void k_gemm_nn(size_t n, int m, int K,
                                                 const double* __restrict__ A, size_t lda,
                                                 const double* __restrict__ B, size_t ldb,
                                                 double* __restrict__ C, size_t ldc) {
  __shared__ double As[kTK][kTM];
  __shared__ double Bs[kTK][kTN + 1];
  const size_t row0 = (size_t)blockIdx.x * kTM;
  const int col0 = blockIdx.y * kTN;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double c[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) c[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += kTK) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int e = threadIdx.x + l * kWG;
      const int rr = e % kTM, kk = e / kTM;
      const size_t gr = row0 + rr;
      const int gk = k0 + kk;
      As[kk][rr] = (gr < n && gk < K) ? A[(size_t)gk * lda + gr] : 0.0;
      const int kb = e % kTK, cb = e / kTK;
      const int gc = col0 + cb, gkb = k0 + kb;
      Bs[kb][cb] = (gc < m && gkb < K) ? B[(size_t)gc * ldb + gkb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) c[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gc = col0 + ty + 16 * j;
    if (gc >= m) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t gr = row0 + tx + 16 * i;
      if (gr < n) C[(size_t)gc * ldc + gr] = c[i][j];
    }
  }
}

// ---------------------------------------------------------------------
// per-column reductions (column-major), grid (chunks, m)
// ---------------------------------------------------------------------
constexpr int kColRows = 8192;

template <int KIND>  // 0: X.Y ; 1: ||W - t U||^2
__global__ __launch_bounds__(kWG) void k_col_partial(size_t n, const double* __restrict__ X,
                                                     size_t ldx, const double* __restrict__ Y,
                                                     size_t ldy, const double* __restrict__ t,
                                                     double* __restrict__ part, int nch) {
  __shared__ double sh[4];
  const int ch = blockIdx.x, j = blockIdx.y;
  const double* x = X + (size_t)j * ldx;
  const double* y = Y + (size_t)j * ldy;
  const double tj = KIND == 1 ? t[j] : 0.0;
  const size_t rbeg = (size_t)ch * kColRows, rend = min(n, rbeg + kColRows);
  double a = 0.0;
  for (size_t r = rbeg + threadIdx.x; r < rend; r += kWG) {
    if (KIND == 0) {
      a += x[r] * y[r];
    } else {
      const double d = x[r] - tj * y[r];
      a += d * d;
    }
  }
  a = block_sum(a, sh);
  if (threadIdx.x == 0) part[(size_t)j * nch + ch] = a;
}

__global__ void k_col_scale_rsqrt(size_t n, double* __restrict__ X, size_t ldx,
                                  const double* __restrict__ s2) {
  const int j = blockIdx.y;
  const double s = s2[j] > 0.0 ? 1.0 / sqrt(s2[j]) : 0.0;
  double* x = X + (size_t)j * ldx;
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (size_t)gridDim.x * blockDim.x)
    x[r] *= s;
}

// U(:,j) = p[j] * U(:,j) + Y(:,j)
__global__ void k_col_scale_add(size_t n, double* __restrict__ U, size_t ldu,
                                const double* __restrict__ p, const double* __restrict__ Y,
                                size_t ldy) {
  const int j = blockIdx.y;
  const double pj = p[j];
  double* u = U + (size_t)j * ldu;
  const double* yv = Y + (size_t)j * ldy;
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (size_t)gridDim.x * blockDim.x)
    u[r] = pj * u[r] + yv[r];
}

// R(:,j) = (W(:,j) - th[j] * U(:,j)) * inv_dd
__global__ void k_col_resid(size_t n, const double* __restrict__ W, size_t ldw,
                            const double* __restrict__ U, size_t ldu,
                            const double* __restrict__ th, double inv_dd,
                            double* __restrict__ R, size_t ldr) {
  const int j = blockIdx.y;
  const double tj = th[j];
  const double* w = W + (size_t)j * ldw;
  const double* u = U + (size_t)j * ldu;
  double* rr = R + (size_t)j * ldr;
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (size_t)gridDim.x * blockDim.x)
    rr[r] = (w[r] - tj * u[r]) * inv_dd;
}

__global__ void k_gather_cols(size_t nrow, const double* __restrict__ src, size_t lds,
                              const int* __restrict__ idx, double* __restrict__ dst,
                              size_t ldd) {
  const int j = blockIdx.y;
  const double* s = src + (size_t)idx[j] * lds;
  double* d = dst + (size_t)j * ldd;
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrow;
       r += (size_t)gridDim.x * blockDim.x)
    d[r] = s[r];
}

// ---------------------------------------------------------------------
// QL rotation application: thread per row of Z, one sweep at a time; the
// rotated pair shares column i with the next rotation, so that value is
// carried in a register and each rotation costs one load and one store.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kWG) void k_apply_rotations(int k, double* __restrict__ Z,
                                                         int nsweep,
                                                         const int* __restrict__ top,
                                                         const int* __restrict__ cnt,
                                                         const long long* __restrict__ off,
                                                         const double* __restrict__ cs) {
  const int r = blockIdx.x * kWG + threadIdx.x;
  if (r >= k) return;
  double* z = Z + r;
  const size_t ldz = (size_t)k;
  for (int sw = 0; sw < nsweep; ++sw) {
    const int ihi = top[sw], c = cnt[sw];
    const double* rot = cs + 2 * off[sw];
    double carry = z[(size_t)(ihi + 1) * ldz];
    int j = 0;
    for (; j + 4 <= c; j += 4) {
      const int i = ihi - j;
      const double z0 = z[(size_t)i * ldz];
      const double z1 = z[(size_t)(i - 1) * ldz];
      const double z2 = z[(size_t)(i - 2) * ldz];
      const double z3 = z[(size_t)(i - 3) * ldz];
      double cc = rot[2 * j], ss = rot[2 * j + 1];
      z[(size_t)(i + 1) * ldz] = ss * z0 + cc * carry;
      carry = cc * z0 - ss * carry;
      cc = rot[2 * j + 2];
      ss = rot[2 * j + 3];
      z[(size_t)i * ldz] = ss * z1 + cc * carry;
      carry = cc * z1 - ss * carry;
      cc = rot[2 * j + 4];
      ss = rot[2 * j + 5];
      z[(size_t)(i - 1) * ldz] = ss * z2 + cc * carry;
      carry = cc * z2 - ss * carry;
      cc = rot[2 * j + 6];
      ss = rot[2 * j + 7];
      z[(size_t)(i - 2) * ldz] = ss * z3 + cc * carry;
      carry = cc * z3 - ss * carry;
    }
    for (; j < c; ++j) {
      const int i = ihi - j;
      const double zi = z[(size_t)i * ldz];
      const double cc = rot[2 * j], ss = rot[2 * j + 1];
      z[(size_t)(i + 1) * ldz] = ss * zi + cc * carry;
      carry = cc * zi - ss * carry;
    }
    z[(size_t)(ihi - c + 1) * ldz] = carry;
  }
}

// ---------------------------------------------------------------------
// row-major block helpers
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kWG) void k_rm_col_nrm2(size_t n, int B,
                                                     const double* __restrict__ X,
                                                     double* __restrict__ part) {
  __shared__ double s_red[kWG];
  const int col = threadIdx.x % B, grp = threadIdx.x / B, NG = kWG / B;
  double a = 0.0;
  for (size_t r = (size_t)blockIdx.x * NG + grp; r < n; r += (size_t)gridDim.x * NG) {
    const double v = X[r * B + col];
    a += v * v;
  }
  s_red[threadIdx.x] = a;
  __syncthreads();
  for (int st = NG / 2; st > 0; st >>= 1) {
    if (grp < st) s_red[threadIdx.x] += s_red[threadIdx.x + st * B];
    __syncthreads();
  }
  if (grp == 0) part[(size_t)blockIdx.x * B + col] = s_red[col];
}

__global__ void k_rm_col_scale_rsqrt(size_t n, int B, double* __restrict__ X,
                                     const double* __restrict__ s2) {
  const size_t tot = n * (size_t)B;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const double s = s2[i % B];
    X[i] = s > 0.0 ? X[i] / sqrt(s) : 0.0;
  }
}

// counter-based generator: splitmix64 finaliser of (seed, column, row)
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__global__ void k_rm_random(size_t n, int B, int kind, unsigned long long seed, int col0,
                            int ncols, double* __restrict__ X) {
  const size_t tot = n * (size_t)B;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % B);
    const size_t r = i / B;
    double v = 0.0;
    if (c < ncols) {
      const unsigned long long key =
          mix64(seed + 0x9e3779b97f4a7c15ULL * (unsigned long long)(col0 + c + 1)) ^ (r * 0xd1b54a32d192ed03ULL);
      const unsigned long long h = mix64(key);
      if (kind == 0) {  // Rademacher
        v = (h >> 63) ? 1.0 : -1.0;
      } else {          // Gaussian, Box-Muller on two 53-bit uniforms
        const unsigned long long h2 = mix64(key ^ 0x632be59bd9b4e019ULL);
        const double u1 = ((h >> 11) + 1) * (1.0 / 9007199254740993.0);
        const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);
        v = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      }
    }
    X[i] = v;
  }
}

template <typename T>
__global__ void k_pack_cols(size_t n, int B, ColPtrs cols, T* __restrict__ X) {
  const size_t tot = n * (size_t)B;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % B);
    X[i] = (T)cols.p[c][i / B];
  }
}

template <typename T>
__global__ void k_unpack_cols(size_t n, int B, const T* __restrict__ X, ColPtrs cols) {
  const size_t tot = n * (size_t)B;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % B);
    cols.p[c][i / B] = (double)X[i];
  }
}

// dst (n x B2, interleaved) = the first B2 columns of src (n x B, interleaved)
template <typename T>
__global__ void k_narrow_cols(size_t n, int B, int B2, const T* __restrict__ src,
                              T* __restrict__ dst) {
  const size_t tot = n * (size_t)B2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / B2;
    dst[i] = src[r * B + (i - r * B2)];
  }
}

// cols.p[c - c0][r] = X[r*B + c] for the columns c0 <= c < c1 of an interleaved block
template <typename T>
__global__ void k_unpack_range(size_t n, int B, int c0, int c1, const T* __restrict__ X,
                               ColPtrs cols) {
  const int w = c1 - c0;
  const size_t tot = n * (size_t)w;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / w;
    const int c = (int)(i - r * w);
    cols.p[c][r] = (double)X[r * B + c0 + c];
  }
}

template <bool TO_CM>
__global__ void k_rm_cm(size_t n, int B, const double* __restrict__ src,
                        double* __restrict__ dst, size_t ld) {
  const size_t tot = n * (size_t)B;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % B);
    const size_t r = i / B;
    if (TO_CM)
      dst[(size_t)c * ld + r] = src[i];
    else
      dst[i] = src[(size_t)c * ld + r];
  }
}

// Gram panel: part[s][j][i] = sum over rows of chunk s of A(r,i) * B(r,j)
constexpr int kGnRows = 8192;  // rows per split
__global__ __launch_bounds__(kWG) void k_gemm_tn_partial(size_t n, int m, int p,
                                                         const double* __restrict__ A,
                                                         size_t lda,
                                                         const double* __restrict__ Bm,
                                                         size_t ldb,
                                                         double* __restrict__ part) {
  __shared__ double As[kTK][kTM + 1];
  __shared__ double Bs[kTK][kTN + 1];
  const int i0 = blockIdx.x * kTM, j0 = blockIdx.y * kTN;
  const size_t rbeg = (size_t)blockIdx.z * kGnRows;
  const size_t rend = min(n, rbeg + kGnRows);
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double c[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) c[i][j] = 0.0;
  for (size_t r0 = rbeg; r0 < rend; r0 += kTK) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int e = threadIdx.x + l * kWG;
      const int rr = e % kTK, cc = e / kTK;  // 16 rows x 64 columns
      const size_t gr = r0 + rr;
      const int gi = i0 + cc, gj = j0 + cc;
      As[rr][cc] = (gr < rend && gi < m) ? A[(size_t)gi * lda + gr] : 0.0;
      Bs[rr][cc] = (gr < rend && gj < p) ? Bm[(size_t)gj * ldb + gr] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) c[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
  double* out = part + (size_t)blockIdx.z * m * p;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gj = j0 + ty + 16 * j;
    if (gj >= p) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int gi = i0 + tx + 16 * i;
      if (gi < m) out[(size_t)gj * m + gi] = c[i][j];
    }
  }
}

__global__ void k_sum_splits(int nsplit, size_t cnt, const double* __restrict__ part,
                             int m, double* __restrict__ C, size_t ldc) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (size_t)gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int s = 0; s < nsplit; ++s) a += part[(size_t)s * cnt + e];
    const size_t i = e % m, j = e / m;
    C[j * ldc + i] = a;
  }
}

// out[k] = (a[k] - (ja[k] == row ? cc : 0)) * scale, one thread per row
__global__ void k_shift_diag(int n, const double* __restrict__ d, double cc, double scale,
                             float* __restrict__ out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const double v = d[r] - cc;
    out[r] = (float)(v * scale);
  }
}

__global__ void k_shift_scale(int n, const int* __restrict__ ia, const int* __restrict__ ja,
                              const double* __restrict__ a, double cc, double scale,
                              float* __restrict__ out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    for (int k = ia[r]; k < ia[r + 1]; ++k) {
      const double v = (ja[k] == r) ? a[k] - cc : a[k];
      out[k] = (float)(v * scale);
    }
  }
}

inline unsigned grid_for(size_t n, int per_block, unsigned cap) {
  size_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

template <typename VT, int MODE>
hipError_t launch_cheb_B(const Csr& A, int B, const void* xk, const void* xkm1, void* xkp1,
                         void* y, const void* v0, const double* s, const double* mu0,
                         double t, double cc, bool first, int yflush, double* part,
                         hipStream_t st) {
#define EVSL_CHEB_CASE(BB)                                                                 \
  case BB:                                                                                 \
    return launch_variant<VT, BB, SpmmPolicy<VT, BB, MODE>::V, MODE,                      \
                          SpmmPolicy<VT, BB, MODE>::LDS, SpmmPolicy<VT, BB, MODE>::R>(        \
        A, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
  if constexpr (MODE == kKpm) {
    switch (B) {
      EVSL_CHEB_CASE(1)
      EVSL_CHEB_CASE(2)
      EVSL_CHEB_CASE(4)
      EVSL_CHEB_CASE(8)
      EVSL_CHEB_CASE(16)
      EVSL_CHEB_CASE(32)
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (B) {
      EVSL_CHEB_CASE(1)
      EVSL_CHEB_CASE(2)
      EVSL_CHEB_CASE(3)
      EVSL_CHEB_CASE(4)
      EVSL_CHEB_CASE(5)
      EVSL_CHEB_CASE(6)
      EVSL_CHEB_CASE(7)
      EVSL_CHEB_CASE(8)
      EVSL_CHEB_CASE(16)
      EVSL_CHEB_CASE(32)
      default:
        return hipErrorInvalidValue;
    }
  }
#undef EVSL_CHEB_CASE
}

// STREAM-style copy with 16-byte accesses: the measured HBM reference the
// filter kernel is compared with
// The bandwidth yardstick: each thread keeps kCopyUnroll 16-byte loads in
// flight (coalesced: word j of a thread sits kWG words after word j-1), and
// loads and stores are non-temporal, so the copy streams through the caches
// the way a well-formed bulk kernel does.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCopyUnroll = 4;
__global__ __launch_bounds__(kWG) void k_stream_copy(size_t n16, const u32x4* __restrict__ src,
                                                     u32x4* __restrict__ dst) {
  const size_t step = (size_t)gridDim.x * kWG * kCopyUnroll;
  for (size_t base = (size_t)blockIdx.x * kWG * kCopyUnroll + threadIdx.x; base < n16;
       base += step) {
    u32x4 v[kCopyUnroll];
#pragma unroll
    for (int j = 0; j < kCopyUnroll; ++j) {
      const size_t i = base + (size_t)j * kWG;
      if (i < n16) v[j] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int j = 0; j < kCopyUnroll; ++j) {
      const size_t i = base + (size_t)j * kWG;
      if (i < n16) __builtin_nontemporal_store(v[j], dst + i);
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
hipError_t cheb_group_step(bool fp32, int B, bool i16, const ChebGroupEntry* tab,
                           const int* owner, int total, int k, hipStream_t st) {
  if (total <= 0) return hipSuccess;
  if (k < 1) return hipErrorInvalidValue;
  auto go = [&](auto vt_tag, auto b_tag, auto i16_tag) {
    using VT = decltype(vt_tag);
    constexpr int BB = decltype(b_tag)::value;
    constexpr bool I = decltype(i16_tag)::value;
    const size_t dyn =
        SpmmPolicy<VT, BB, kCheb>::R > 1 ? (size_t)(kRowsCap + 1) * sizeof(int) : 0;
    hipLaunchKernelGGL((k_cheb_group<VT, BB, I>), dim3(total), dim3(kWG), dyn, st, tab, owner,
                       total, k);
    return hipGetLastError();
  };
  auto width = [&](auto vt_tag, auto i16_tag) -> hipError_t {
    switch (B) {
#define EVSL_GROUP_CASE(BB) \
  case BB:                  \
    return go(vt_tag, std::integral_constant<int, BB>{}, i16_tag);
      EVSL_GROUP_CASE(1)
      EVSL_GROUP_CASE(2)
      EVSL_GROUP_CASE(3)
      EVSL_GROUP_CASE(4)
      EVSL_GROUP_CASE(5)
      EVSL_GROUP_CASE(6)
      EVSL_GROUP_CASE(7)
      EVSL_GROUP_CASE(8)
#undef EVSL_GROUP_CASE
      default:
        return hipErrorInvalidValue;
    }
  };
  using T16 = std::integral_constant<bool, true>;
  using T32 = std::integral_constant<bool, false>;
  if (fp32) return i16 ? width(0.0f, T16{}) : width(0.0f, T32{});
  return i16 ? width(0.0, T16{}) : width(0.0, T32{});
}

hipError_t stream_copy(size_t bytes, const void* src, void* dst, hipStream_t st) {
  const size_t n16 = bytes / 16;
  if (n16 == 0) return hipSuccess;
  const unsigned grid = grid_for((n16 + kCopyUnroll - 1) / kCopyUnroll, kWG, 65536);
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kWG), 0, st, n16, (const u32x4*)src,
                     (u32x4*)dst);
  return hipGetLastError();
}

// V (lanes_only) or V*R of the policy for (B, precision, mode); -1 if unsupported
static int policy_factor(int B, bool fp32, Mode mode, bool lanes_only) {
  auto vec = [&](auto vt_tag, auto mode_tag) -> int {
    using VT = decltype(vt_tag);
    constexpr int M = decltype(mode_tag)::value;
    switch (B) {
#define EVSL_VEC_CASE(BB) \
  case BB:                \
    return lanes_only ? SpmmPolicy<VT, BB, M>::V                                 \
                      : SpmmPolicy<VT, BB, M>::V * SpmmPolicy<VT, BB, M>::R;
      EVSL_VEC_CASE(1)
      EVSL_VEC_CASE(2)
      EVSL_VEC_CASE(3)
      EVSL_VEC_CASE(4)
      EVSL_VEC_CASE(5)
      EVSL_VEC_CASE(6)
      EVSL_VEC_CASE(7)
      EVSL_VEC_CASE(8)
      EVSL_VEC_CASE(16)
      EVSL_VEC_CASE(32)
#undef EVSL_VEC_CASE
      default:
        return -1;
    }
  };
  using P = std::integral_constant<int, kPlain>;
  using C = std::integral_constant<int, kCheb>;
  using K = std::integral_constant<int, kKpm>;
  using R = std::integral_constant<int, kRes>;
  if (fp32)
    return mode == kPlain ? vec(0.0f, P{}) : mode == kCheb ? vec(0.0f, C{})
           : mode == kKpm ? vec(0.0f, K{}) : vec(0.0f, R{});
  return mode == kPlain ? vec(0.0, P{}) : mode == kCheb ? vec(0.0, C{})
         : mode == kKpm ? vec(0.0, K{}) : vec(0.0, R{});
}

int cheb_spmm_vec(int B, bool fp32, Mode mode) { return policy_factor(B, fp32, mode, true); }

int cheb_spmm_rows(int B, bool fp32, Mode mode) {
  const int v = policy_factor(B, fp32, mode, true), vr = policy_factor(B, fp32, mode, false);
  return v > 0 ? vr / v : -1;
}

bool cheb_spmm_streaming(const Csr& A, int B, bool fp32, Mode mode) {
  return streaming_launch(A, B, fp32 ? 4 : 8, mode);
}

int cheb_spmm_nblk(const Csr& A, int B, bool fp32, Mode mode) {
  const int VR = policy_factor(B, fp32, mode, false);
  const int slot = VR > 0 ? width_slot(B / VR) : -1;
  return slot < 0 ? -1 : A.nblk[slot];
}

int spmm_table_layout(int B, bool fp32, Mode mode) {
  const int v = policy_factor(B, fp32, mode, true), vr = policy_factor(B, fp32, mode, false);
  return v > 0 ? 10 * v + vr / v : -1;
}

hipError_t cheb_spmm_layout(int layout, const Csr& A, int B, bool fp32, Mode mode,
                            const void* xk, const void* xkm1, void* xkp1, void* y,
                            const void* v0, const double* s, const double* mu0, double t,
                            double cc, bool first, int yflush, double* part, hipStream_t st) {
  if (A.n <= 0) return hipSuccess;
  if (yflush < 0 || yflush > 3 || (first && yflush != 1) || (yflush == 3 && !xkm1))
    return hipErrorInvalidValue;
  return fp32 ? cheb_spmm_layout_f32(layout, A, B, mode, xk, xkm1, xkp1, y, v0, s, mu0, t, cc,
                                     first, yflush, part, st)
              : cheb_spmm_layout_f64(layout, A, B, mode, xk, xkm1, xkp1, y, v0, s, mu0, t, cc,
                                     first, yflush, part, st);
}

hipError_t cheb_spmm(const Csr& A, int B, bool fp32, Mode mode, const void* xk,
                     const void* xkm1, void* xkp1, void* y, const void* v0, const double* s,
                     const double* mu0, double t, double cc, bool first, int yflush,
                     double* part, hipStream_t st) {
  if (A.n <= 0) return hipSuccess;
  if (yflush < 0 || yflush > 3 || (first && yflush != 1) || (yflush == 3 && !xkm1))
    return hipErrorInvalidValue;
  // ChebAv and residual steps of the tuned widths go through the matrix's
  // layout tuner (same bits, the fastest layout once it has measured)
  if (A.tune && (mode == kCheb || mode == kRes) && tune_width_index(B) >= 0)
    return cheb_spmm_tuned(A, B, fp32, mode, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first,
                           yflush, part, st);
  if (fp32) {
    switch (mode) {
      case kPlain:
        return launch_cheb_B<float, kPlain>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
      case kCheb:
        return launch_cheb_B<float, kCheb>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
      case kKpm:
        return launch_cheb_B<float, kKpm>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
      case kRes:
        return launch_cheb_B<float, kRes>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
    }
  } else {
    switch (mode) {
      case kPlain:
        return launch_cheb_B<double, kPlain>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
      case kCheb:
        return launch_cheb_B<double, kCheb>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
      case kKpm:
        return launch_cheb_B<double, kKpm>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
      case kRes:
        return launch_cheb_B<double, kRes>(A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first, yflush, part, st);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t kpm_group_step(bool fp32, int B, const KpmGroupEntry* tab, const int* owner,
                          int total, int k, const KpmSumEntry* sums, const int* owner1,
                          const int* first1, int nl1, const int* owner2, int nl2,
                          hipStream_t st) {
  if (total <= 0 || nl1 <= 0) return hipErrorInvalidValue;
#define EVSL_KPM_GROUP(VT_, B_)                                                           \
  hipLaunchKernelGGL((k_kpm_group<VT_, B_, SpmmPolicy<VT_, B_, kKpm>::V>), dim3(total),   \
                     dim3(kWG), 0, st, tab, owner, total, k)
#define EVSL_KPM_GROUP_B(VT_)                  \
  switch (B) {                                 \
    case 1: EVSL_KPM_GROUP(VT_, 1); break;     \
    case 2: EVSL_KPM_GROUP(VT_, 2); break;     \
    case 4: EVSL_KPM_GROUP(VT_, 4); break;     \
    case 8: EVSL_KPM_GROUP(VT_, 8); break;     \
    case 16: EVSL_KPM_GROUP(VT_, 16); break;   \
    case 32: EVSL_KPM_GROUP(VT_, 32); break;   \
    default: return hipErrorInvalidValue;      \
  }
  if (fp32) {
    EVSL_KPM_GROUP_B(float);
  } else {
    EVSL_KPM_GROUP_B(double);
  }
#undef EVSL_KPM_GROUP_B
#undef EVSL_KPM_GROUP
  hipLaunchKernelGGL(k_kpm_group_sums, dim3(nl1), dim3(kWG), 0, st, sums, owner1, first1, B, k,
                     1);
  if (nl2 > 0)
    hipLaunchKernelGGL(k_kpm_group_sums, dim3(nl2), dim3(kWG), 0, st, sums, owner2,
                       (const int*)nullptr, B, k, 2);
  return hipGetLastError();
}

hipError_t sum_block_partials(const double* part, int nblk, int B, double* out,
                              double* scratch, hipStream_t st) {
  if (B < 1 || B > 32) return hipErrorInvalidValue;
  const int G = kSumScratch / 32;  // first-level blocks (<= 256)
  if (sum_levels_blocks(nblk, B) == 1 || !scratch) {
    hipLaunchKernelGGL(k_sum_block_partials, dim3(1), dim3(kWG), 0, st, part, nblk, B, out);
  } else {
    hipLaunchKernelGGL(k_sum_block_partials, dim3(G), dim3(kWG), 0, st, part, nblk, B, scratch);
    hipLaunchKernelGGL(k_sum_block_partials, dim3(1), dim3(kWG), 0, st,
                       (const double*)scratch, G, B, out);
  }
  return hipGetLastError();
}

hipError_t dot_partial(size_t n, const double* x, const double* y, double* part,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(kWG), 0, st, n, x, y, part);
  return hipGetLastError();
}

hipError_t sum_scalar(const double* part, int cnt, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_scalar, dim3(1), dim3(kWG), 0, st, part, cnt, out);
  return hipGetLastError();
}

hipError_t axpy_dot(size_t n, double a, const double* x, double* y, const double* z,
                    double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dot, dim3(kRedBlocks), dim3(kWG), 0, st, n, a, x, y, z, part);
  return hipGetLastError();
}

hipError_t axpy_dev_nrm(size_t n, double sign, const double* a_d, const double* x,
                        double* y, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dev<true>, dim3(kRedBlocks), dim3(kWG), 0, st, n, sign, a_d, x,
                     y, (const double*)nullptr, part);
  return hipGetLastError();
}

hipError_t axpy_dev_dot(size_t n, double sign, const double* a_d, const double* x,
                        double* y, const double* z, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dev<false>, dim3(kRedBlocks), dim3(kWG), 0, st, n, sign, a_d, x,
                     y, z, part);
  return hipGetLastError();
}

hipError_t scal(size_t n, double a, double* x, hipStream_t st) {
  hipLaunchKernelGGL(k_scal, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, a, x);
  return hipGetLastError();
}

hipError_t scal_copy(size_t n, double a, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_scal_copy, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, a, x, y);
  return hipGetLastError();
}

hipError_t fill(size_t n, double v, double* x, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, v, x);
  return hipGetLastError();
}

hipError_t fill_u32(size_t n, unsigned v, unsigned* x, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, v, x);
  return hipGetLastError();
}

hipError_t shift_diag_f32(int n, const double* d, double cc, double scale, float* out,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_shift_diag, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, d, cc,
                     scale, out);
  return hipGetLastError();
}

hipError_t shift_scale_f32(const Csr& A, double cc, double scale, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_shift_scale, dim3(grid_for(A.n, kWG, 4096)), dim3(kWG), 0, st, A.n, A.ia,
                     A.ja, A.a64, cc, scale, out);
  return hipGetLastError();
}

hipError_t set_identity(int k, double* Z, hipStream_t st) {
  const size_t tot = (size_t)k * k;
  hipLaunchKernelGGL(k_set_identity, dim3(grid_for(tot, kWG, 4096)), dim3(kWG), 0, st, k, Z);
  return hipGetLastError();
}

hipError_t f64_to_f32(size_t n, const double* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, x, y);
  return hipGetLastError();
}

hipError_t f32_to_f64(size_t n, const float* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, x, y);
  return hipGetLastError();
}

int gemv_t_chunks(size_t n) { return (int)((n + kGtRows - 1) / kGtRows); }

hipError_t gemv_t(size_t n, int k, const double* V, size_t ld, const double* w,
                  double* part, double* h, hipStream_t st) {
  if (k <= 0) return hipSuccess;
  const int nch = gemv_t_chunks(n);
  const int ncg = (k + kGtCols - 1) / kGtCols;
  hipLaunchKernelGGL(k_gemv_t_partial, dim3((unsigned)nch * ncg), dim3(kWG), 0, st, n, k, V,
                     ld, w, part, nch, ncg);
  hipLaunchKernelGGL(k_gemv_t_reduce, dim3(k), dim3(64), 0, st, part, nch, h);
  return hipGetLastError();
}

hipError_t gemv_n_sub_nrm(size_t n, int k, const double* V, size_t ld, const double* h,
                          double* w, double* part, hipStream_t st) {
  const unsigned g = (unsigned)((n + kWG - 1) / kWG);
  hipLaunchKernelGGL(k_gemv_n<true>, dim3(g), dim3(kWG), 0, st, n, k, V, ld, h, w, part);
  return hipGetLastError();
}

hipError_t gemv_t_f32(size_t n, int k, const float* V, size_t ld, const double* w,
                      double* part, double* h, hipStream_t st) {
  if (k <= 0) return hipSuccess;
  const int nch = gemv_t_chunks(n);
  const int ncg = (k + kGtCols - 1) / kGtCols;
  hipLaunchKernelGGL(k_gemv_t_partial_f32, dim3((unsigned)nch * ncg), dim3(kWG), 0, st, n, k,
                     V, ld, w, part, nch, ncg);
  hipLaunchKernelGGL(k_gemv_t_reduce, dim3(k), dim3(64), 0, st, part, nch, h);
  return hipGetLastError();
}

int gemv_n_sub_f32_blocks(size_t n) { return (int)((n + 2 * kWG - 1) / (2 * kWG)); }

hipError_t gemv_n_sub_nrm_f32(size_t n, int k, const float* V, size_t ld, const double* h,
                              double* w, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_gemv_n_sub_f32, dim3((unsigned)gemv_n_sub_f32_blocks(n)), dim3(kWG), 0,
                     st, n, k, V, ld, h, w, part);
  return hipGetLastError();
}

hipError_t scal_f32copy(size_t n, double a, double* x, float* x32, hipStream_t st) {
  hipLaunchKernelGGL(k_scal_f32copy, dim3(grid_for(n, kWG, 4096)), dim3(kWG), 0, st, n, a, x,
                     x32);
  return hipGetLastError();
}

hipError_t gemv_n(size_t n, int k, const double* V, size_t ld, const double* h, double* y,
                  hipStream_t st) {
  const unsigned g = (unsigned)((n + kWG - 1) / kWG);
  hipLaunchKernelGGL(k_gemv_n<false>, dim3(g), dim3(kWG), 0, st, n, k, V, ld, h, y,
                     (double*)nullptr);
  return hipGetLastError();
}

hipError_t gemm_nn(size_t n, int m, int k, const double* A, size_t lda, const double* B,
                   size_t ldb, double* C, size_t ldc, hipStream_t st) {
  if (n == 0 || m <= 0) return hipSuccess;
  dim3 grid((unsigned)((n + kTM - 1) / kTM), (m + kTN - 1) / kTN);
  hipLaunchKernelGGL(k_gemm_nn, grid, dim3(kWG), 0, st, n, m, k, A, lda, B, ldb, C, ldc);
  return hipGetLastError();
}

static int col_chunks(size_t n) { return (int)((n + kColRows - 1) / kColRows); }

hipError_t col_dot(size_t n, int m, const double* X, size_t ldx, const double* Y, size_t ldy,
                   double* part, double* out, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  const int nch = col_chunks(n);
  hipLaunchKernelGGL(k_col_partial<0>, dim3(nch, m), dim3(kWG), 0, st, n, X, ldx, Y, ldy,
                     (const double*)nullptr, part, nch);
  hipLaunchKernelGGL(k_gemv_t_reduce, dim3(m), dim3(64), 0, st, part, nch, out);
  return hipGetLastError();
}

hipError_t col_res2(size_t n, int m, const double* W, size_t ldw, const double* U, size_t ldu,
                    const double* t, double* part, double* out, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  const int nch = col_chunks(n);
  hipLaunchKernelGGL(k_col_partial<1>, dim3(nch, m), dim3(kWG), 0, st, n, W, ldw, U, ldu, t,
                     part, nch);
  hipLaunchKernelGGL(k_gemv_t_reduce, dim3(m), dim3(64), 0, st, part, nch, out);
  return hipGetLastError();
}

hipError_t col_scale_rsqrt(size_t n, int m, double* X, size_t ldx, const double* s2,
                           hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_col_scale_rsqrt, dim3(grid_for(n, kWG, 256), m), dim3(kWG), 0, st, n,
                     X, ldx, s2);
  return hipGetLastError();
}

hipError_t col_scale_add(size_t n, int m, double* U, size_t ldu, const double* p, const double* Y,
                         size_t ldy, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_col_scale_add, dim3(grid_for(n, kWG, 256), m), dim3(kWG), 0, st, n, U,
                     ldu, p, Y, ldy);
  return hipGetLastError();
}

hipError_t col_resid(size_t n, int m, const double* W, size_t ldw, const double* U, size_t ldu,
                     const double* th, double inv_dd, double* R, size_t ldr, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_col_resid, dim3(grid_for(n, kWG, 256), m), dim3(kWG), 0, st, n, W, ldw,
                     U, ldu, th, inv_dd, R, ldr);
  return hipGetLastError();
}

hipError_t gather_cols(size_t nrow, int m, const double* src, size_t lds, const int* idx,
                       double* dst, size_t ldd, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_cols, dim3(grid_for(nrow, kWG, 256), m), dim3(kWG), 0, st,
                     nrow, src, lds, idx, dst, ldd);
  return hipGetLastError();
}

hipError_t apply_rotations(int k, double* Z, int nsweep, const int* top, const int* cnt,
                           const long long* off, const double* cs, hipStream_t st) {
  if (k <= 0 || nsweep <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_rotations, dim3((k + kWG - 1) / kWG), dim3(kWG), 0, st, k, Z,
                     nsweep, top, cnt, off, cs);
  return hipGetLastError();
}

hipError_t rm_col_nrm2(size_t n, int B, const double* X, double* part, double* out,
                       hipStream_t st) {
  if (B < 1 || B > kWG || (kWG % B) != 0) return hipErrorInvalidValue;
  const int NG = kWG / B;
  const unsigned g = grid_for(n, NG, 1024);
  hipLaunchKernelGGL(k_rm_col_nrm2, dim3(g), dim3(kWG), 0, st, n, B, X, part);
  hipLaunchKernelGGL(k_sum_block_partials, dim3(1), dim3(kWG), 0, st, (const double*)part,
                     (int)g, B, out);
  return hipGetLastError();
}

hipError_t rm_col_scale_rsqrt(size_t n, int B, double* X, const double* s2, hipStream_t st) {
  hipLaunchKernelGGL(k_rm_col_scale_rsqrt, dim3(grid_for(n * B, kWG, 8192)), dim3(kWG), 0, st,
                     n, B, X, s2);
  return hipGetLastError();
}

hipError_t rm_random(size_t n, int B, int kind, unsigned long long seed, int col0, int ncols,
                     double* X, hipStream_t st) {
  hipLaunchKernelGGL(k_rm_random, dim3(grid_for(n * B, kWG, 8192)), dim3(kWG), 0, st, n, B,
                     kind, seed, col0, ncols, X);
  return hipGetLastError();
}

hipError_t pack_cols(size_t n, int B, const ColPtrs& cols, void* X, bool fp32, hipStream_t st) {
  if (B <= 0 || B > kMaxCols) return hipErrorInvalidValue;
  const unsigned g = grid_for(n * B, kWG, 8192);
  if (fp32)
    hipLaunchKernelGGL(k_pack_cols<float>, dim3(g), dim3(kWG), 0, st, n, B, cols, (float*)X);
  else
    hipLaunchKernelGGL(k_pack_cols<double>, dim3(g), dim3(kWG), 0, st, n, B, cols, (double*)X);
  return hipGetLastError();
}

hipError_t unpack_cols(size_t n, int B, const void* X, bool fp32, const ColPtrs& cols,
                       hipStream_t st) {
  if (B <= 0 || B > kMaxCols) return hipErrorInvalidValue;
  const unsigned g = grid_for(n * B, kWG, 8192);
  if (fp32)
    hipLaunchKernelGGL(k_unpack_cols<float>, dim3(g), dim3(kWG), 0, st, n, B, (const float*)X,
                       cols);
  else
    hipLaunchKernelGGL(k_unpack_cols<double>, dim3(g), dim3(kWG), 0, st, n, B,
                       (const double*)X, cols);
  return hipGetLastError();
}

hipError_t narrow_cols(size_t n, int B, int B2, const void* src, void* dst, bool fp32,
                       hipStream_t st) {
  if (B2 <= 0 || B2 > B) return hipErrorInvalidValue;
  const unsigned grid = grid_for(n * (size_t)B2, kWG, 8192);
  if (fp32)
    hipLaunchKernelGGL(k_narrow_cols<float>, dim3(grid), dim3(kWG), 0, st, n, B, B2,
                       (const float*)src, (float*)dst);
  else
    hipLaunchKernelGGL(k_narrow_cols<double>, dim3(grid), dim3(kWG), 0, st, n, B, B2,
                       (const double*)src, (double*)dst);
  return hipGetLastError();
}

hipError_t unpack_range(size_t n, int B, int c0, int c1, const void* X, bool fp32,
                        const ColPtrs& cols, hipStream_t st) {
  if (c0 < 0 || c1 > B || c0 >= c1 || c1 - c0 > kMaxCols) return hipErrorInvalidValue;
  const unsigned grid = grid_for(n * (size_t)(c1 - c0), kWG, 8192);
  if (fp32)
    hipLaunchKernelGGL(k_unpack_range<float>, dim3(grid), dim3(kWG), 0, st, n, B, c0, c1,
                       (const float*)X, cols);
  else
    hipLaunchKernelGGL(k_unpack_range<double>, dim3(grid), dim3(kWG), 0, st, n, B, c0, c1,
                       (const double*)X, cols);
  return hipGetLastError();
}

hipError_t rm_to_cm(size_t n, int B, const double* X, double* P, size_t ld, hipStream_t st) {
  hipLaunchKernelGGL(k_rm_cm<true>, dim3(grid_for(n * B, kWG, 8192)), dim3(kWG), 0, st, n, B,
                     X, P, ld);
  return hipGetLastError();
}

hipError_t cm_to_rm(size_t n, int B, const double* P, size_t ld, double* X, hipStream_t st) {
  hipLaunchKernelGGL(k_rm_cm<false>, dim3(grid_for(n * B, kWG, 8192)), dim3(kWG), 0, st, n, B,
                     P, X, ld);
  return hipGetLastError();
}

size_t gemm_tn_part_elems(size_t n, int m, int p) {
  const size_t nsplit = (n + kGnRows - 1) / kGnRows;
  return nsplit * (size_t)m * p;
}

hipError_t gemm_tn(size_t n, int m, int p, const double* A, size_t lda, const double* B,
                   size_t ldb, double* C, size_t ldc, double* part, hipStream_t st) {
  if (m <= 0 || p <= 0) return hipSuccess;
  const int nsplit = (int)((n + kGnRows - 1) / kGnRows);
  dim3 grid((m + kTM - 1) / kTM, (p + kTN - 1) / kTN, nsplit);
  hipLaunchKernelGGL(k_gemm_tn_partial, grid, dim3(kWG), 0, st, n, m, p, A, lda, B, ldb, part);
  const size_t cnt = (size_t)m * p;
  hipLaunchKernelGGL(k_sum_splits, dim3(grid_for(cnt, kWG, 4096)), dim3(kWG), 0, st, nsplit, cnt,
                     part, m, C, ldc);
  return hipGetLastError();
}

}  // namespace evsl_dev
