// dev.h -- host-callable launchers for the HIP kernels in kernels.hip.
//
// Everything here runs asynchronously on the given stream.  Pointers are
// device pointers unless named *_h.  Vector blocks are stored row-major
// ("interleaved"): entry (row r, column c) of a B-column block lives at
// x[r * B + c], so the B values a sparse row gathers sit in one contiguous
// segment.  Lanczos bases are column-major with leading dimension ld.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstddef>
#include <cstdint>

namespace evsl_dev {

// Row blocks for the LDS-staged CSR kernels: block b covers rows
// [rb[b], rb[b+1]) and either holds at most kNnzCap nonzeros and 256/B rows
// (one (row, column) lane each for a B-column block), or is a single row
// longer than kNnzCap.  One table per supported block width.
constexpr int kWG = 256;
constexpr int kNnzCap = 2048;
constexpr int kRowsCap = 256;
// a row longer than kNnzCap is summed in this many fixed stripes
constexpr int kLongStripes = 32;
constexpr int kNumWidths = 10;  // B = 1..8, 16, 32
inline int width_slot(int B) { return B <= 8 ? B - 1 : (B == 16 ? 8 : (B == 32 ? 9 : -1)); }
inline int width_of_slot(int s) { return s < 8 ? s + 1 : (s == 8 ? 16 : 32); }
inline int rows_cap_for(int B) { return kWG / B; }

struct SpmmTuner;

constexpr int kPatCodes = 255;
constexpr unsigned char kPatDiag = 255;

struct Csr {
  int n = 0, nnz = 0;
  const int* ia = nullptr;
  const int* ja = nullptr;
  // 16-bit column offsets ja - row, when every row's columns lie within
  // +-32767 of it (the filter kernels then read these instead of ja)
  const short* jd = nullptr;
  const double* a64 = nullptr;
  const float* a32 = nullptr;
  // Pattern codes, when the matrix's off-diagonal entries take at most
  // kPatCodes distinct (column offset, value) pairs (stencil and lattice
  // operators): one byte per nonzero, pc[k] < kPatCodes naming the pair
  // (poff[pc[k]], pval[pc[k]]) and pc[k] == kPatDiag the row's diagonal,
  // whose value is dg[row].  The filter kernels then stream one byte per
  // nonzero instead of an index and a value, with the same products in the
  // same order.  pval32/dg32 always go with a32 (both shifted or neither).
  // pdpos[row]: the diagonal's position within its row (kPatDiag: none).
  const unsigned char* pc = nullptr;
  const unsigned char* pdpos = nullptr;
  const int* poff = nullptr;
  bool pshort = false;  // every poff fits int16
  const double* pval64 = nullptr;
  const double* dg64 = nullptr;
  const float* pval32 = nullptr;
  const float* dg32 = nullptr;
  // per block width: nblk+1 pairs (first row of the block, its row pointer)
  const int* rb[kNumWidths] = {};
  int nblk[kNumWidths] = {};
  // the matrix's layout tuner (host object; null: the fixed layout policy)
  SpmmTuner* tune = nullptr;
};

// Launches whose vector rows are at most this many bytes (B * value size)
// read pattern codes; wider ones read the index and value arrays.  Decoding
// costs a table lookup per nonzero ahead of the gathers: on the 100^3
// Laplacian it paid at 1 and 2 FP32 columns and 1 FP64 column (11.2 against
// 13.7 us, 16.2 against 19.1, 16.2 against 19.3) and cost or tied at 4-32
// columns, where the vectors carry most of a launch's bytes.
constexpr int kPatMaxRowBytes = 8;
inline bool pattern_launch(const Csr& A, int B, int vsize) {
  return A.pc != nullptr && B * vsize <= kPatMaxRowBytes;
}

enum Mode { kPlain = 0, kCheb = 1, kKpm = 2, kRes = 3 };

// One Chebyshev degree step over a B-column block (B in {1..8,16,32}; the
// KPM mode needs B in {1,2,4,8,16,32}):
//   acc = A * xk
//   kPlain: xkp1 = acc
//   kCheb : w = t*(acc - cc*xk) - (xkm1 ? xkm1 : 0); xkp1 = w;
//           y (FP64 in both precisions: each term is widened and added in
//           FP64) accumulates the last yflush terms of sum_k mu_k T_k:
//             first (yflush 1): y = mu0[c]*xk + s[c]*w
//             yflush 3: y += s[c-2B]*xkm1, then += s[c-B]*xk, then += s[c]*w
//             yflush 2: y += s[c-B]*xk, then += s[c]*w;  1: y += s[c]*w
//             yflush 0: y untouched (neither read nor written)
//   kKpm  : w as kCheb; xkp1 = w; part[blk*B + c] = sum_rows w*v0
//   kRes  : w = t*(acc - cc*xk) - xkm1 + s[c]*v0 (v0 = residual block R);
//           xkp1 = w; y as kCheb with coefficients mu0[0] (this degree),
//           mu0[-2], mu0[-4] (the two before) and mu0[1] for xk when first
// s, mu0: per-column coefficients (device; kRes see above).
// fp32 selects float vectors and matrix values (float accumulation); y is
// always a block of doubles.
hipError_t cheb_spmm(const Csr& A, int B, bool fp32, Mode mode, const void* xk,
                     const void* xkm1, void* xkp1, void* y, const void* v0,
                     const double* s, const double* mu0, double t, double cc,
                     bool first, int yflush, double* part, hipStream_t st);

// Lane layouts of the fused filter kernel, as 10*V + R (V columns per lane,
// R rows per lane group), that the layout tuner chooses among for the ChebAv
// and residual steps of kTuneWidths columns; 0 ends a row.  Every layout
// sums each row in the same order, so all give the same bits.
constexpr int kNumTuneWidths = 5;
constexpr int kTuneWidths[kNumTuneWidths] = {2, 4, 8, 16, 32};
constexpr int kTuneCand[2][kNumTuneWidths][6] = {{{11, 12, 21, 0, 0, 0},      // FP64, B = 2
                                                  {11, 12, 14, 21, 22, 0},    // 4
                                                  {11, 12, 21, 22, 24, 0},    // 8
                                                  {11, 12, 21, 22, 24, 0},    // 16
                                                  {11, 12, 21, 22, 24, 0}},   // 32
                                                 {{11, 12, 21, 0, 0, 0},      // FP32, B = 2
                                                  {11, 12, 14, 21, 22, 41},   // 4
                                                  {11, 21, 22, 24, 41, 42},   // 8
                                                  {11, 21, 22, 41, 42, 44},   // 16
                                                  {11, 21, 22, 41, 42, 44}}}; // 32
inline int tune_width_index(int B) {
  return B == 2 ? 0 : B == 4 ? 1 : B == 8 ? 2 : B == 16 ? 3 : B == 32 ? 4 : -1;
}
// the tuner's candidates for (B, precision): count, *list (null if none)
inline int tune_candidates(int B, bool fp32, const int** list) {
  const int w = tune_width_index(B);
  if (w < 0) {
    if (list) *list = nullptr;
    return 0;
  }
  const int* row = kTuneCand[fp32 ? 1 : 0][w];
  int k = 0;
  while (k < 6 && row[k]) ++k;
  if (list) *list = row;
  return k;
}
// layout (10*V + R) of the fixed policy for (B, precision, mode)
int spmm_table_layout(int B, bool fp32, Mode mode);
// one cheb_spmm launch with an explicit candidate layout (kCheb or kRes, B
// in kTuneWidths); hipErrorInvalidValue for a layout not in the list
hipError_t cheb_spmm_layout(int layout, const Csr& A, int B, bool fp32, Mode mode,
                            const void* xk, const void* xkm1, void* xkp1, void* y,
                            const void* v0, const double* s, const double* mu0, double t,
                            double cc, bool first, int yflush, double* part, hipStream_t st);
// The per-matrix layout tuner (spmm_tune.cpp).  While a (precision, mode,
// width) key is undecided, each launch of it runs one candidate between two
// events (the ChebAv/residual launches with yflush 0 or 3, the steady state
// of a long series); once every candidate holds its samples and the events
// have completed, the key keeps the candidate of least time per series step
// (two yflush-0 launches and one yflush-3 launch), if it beats the fixed
// policy's layout by 3% (near ties keep the table).  Layouts give the same
// bits, so tuning never changes a result.
SpmmTuner* spmm_tuner_create();
void spmm_tuner_destroy(SpmmTuner* t);
// policy of a key: layout > 0 forces that candidate, 0 the fixed table,
// -1 tunes (the default); resets what was learned.  -1 on a bad key/layout.
int spmm_tuner_set(SpmmTuner* t, int B, bool fp32, Mode mode, int layout);
// layout the key launches now: the forced or chosen one, or the table's
// while undecided
int spmm_tuner_layout(SpmmTuner* t, int B, bool fp32, Mode mode);
// 1 when the key launches a fixed layout (forced, table, or tuned and
// decided), 0 while it is still measuring
int spmm_tuner_decided(SpmmTuner* t, int B, bool fp32, Mode mode);
// cheb_spmm through the tuner (called by cheb_spmm for tunable keys)
hipError_t cheb_spmm_tuned(const Csr& A, int B, bool fp32, Mode mode, const void* xk,
                           const void* xkm1, void* xkp1, void* y, const void* v0,
                           const double* s, const double* mu0, double t, double cc,
                           bool first, int yflush, double* part, hipStream_t st);

// y terms flushed by launch k of a degree-m series (k = 1..m): the first
// launch sets y, then every third launch adds three terms, and the last adds
// what is pending
inline int y_flush(int k, int m) {
  if (k == 1) return 1;
  const int r = (k - 1) % 3;
  if (r == 0) return 3;
  return k == m ? r : 0;
}

// algorithmic vector bytes per row and column of one kCheb launch with
// vectors of vs bytes (kRes: add vs for the residual block): xk (gathered
// once), xkm1, xkp1, and the FP64 y in and out when it is flushed; the
// first launch reads xk and writes xkp1 and y
inline double cheb_vec_bytes(int yflush, bool first, int vs) {
  return first ? 2.0 * vs + 8.0 : 3.0 * vs + (yflush ? 16.0 : 0.0);
}

// Device scratch of the allocation-free filter entry point (noalloc.hip): a
// file-scope device array of kNoAllocScratchBytes per GPU, held until the
// process ends.  noalloc_scratch returns its address on the current device
// (looked up once per device).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
hipError_t noalloc_scratch(void** ptr);

// dst = src for bytes / 16 16-byte words (the bandwidth reference kernel)
hipError_t stream_copy(size_t bytes, const void* src, void* dst, hipStream_t st);

// columns per lane of the cheb_spmm kernel used for this width, precision
// and mode (-1 for an unsupported width)
int cheb_spmm_vec(int B, bool fp32, Mode mode);
// rows per lane group of that kernel (row blocks of R x the lane groups)
int cheb_spmm_rows(int B, bool fp32, Mode mode);

// number of row blocks (KPM partials per column) one cheb_spmm launch of
// this width, precision and mode uses; -1 for an unsupported width
int cheb_spmm_nblk(const Csr& A, int B, bool fp32, Mode mode);
// whether a launch at (B, precision, mode) on A uses the non-temporal
// instance (k_cheb_spmm<..., NT = true>)
bool cheb_spmm_streaming(const Csr& A, int B, bool fp32, Mode mode);

// One matrix of a grouped KPM launch (a device-resident table entry): its
// CSR (int32 column indices), its row-block table for the launch's lane
// layout ((row, ia[row]) pairs), the start block v0 and the ring of three
// T_k blocks (n x B interleaved, element type of the launch), the block
// partials of the moment dots, and the recurrence
// T_{k+1} = t (A - cc I) T_k - T_{k-1} with factor t0 at k = 0.  A launch
// at step k advances this matrix's recurrence at its own step k - k0.
struct KpmGroupEntry {
  const int* ia;
  const int* ja;
  const void* a;
  const int* rbp;
  const void* x0;
  void* ring;
  double* part;
  double t0, t, cc;
  int n, blk0;  // rows; first global row block of this matrix
  int k0, pad;
};
// One column block of a grouped ChebAv launch (a device table entry): its
// matrix's CSR (16-bit column offsets jd when the launch's instance reads
// them), values (FP32 shifted copy or FP64) and row-block table for the
// launch's lane layout, its start block x0 (n x B interleaved), ring of three
// blocks (T_k written to slot (k-1) % 3), FP64 output block y, coefficient
// table mus (row d: the B columns' mu_d, zero beyond a column's degree), the
// recurrence factors (t1 at k = 1, t2 after; cc the shift left to the
// epilogue) and the block's series degree deg (its last launch flushes y).
struct ChebGroupEntry {
  const int* ia;
  const int* ja;
  const short* jd;
  const void* a;
  const int* rbp;
  const void* x0;
  void* ring;
  double* y;
  const double* mus;
  double t1, t2, cc;
  int n, blk0, deg, pad;
};
// Step k of a grouped ChebAv launch over the first `total` row blocks of the
// entries of one (precision, width B in 1..8, index type) class; owner[gb]
// is the entry of global row block gb.  Entries are ordered by decreasing
// degree, so the launch at step k covers the prefix of entries with
// deg >= k.  Lane layout: the fixed policy's for (B, precision, kCheb).
hipError_t cheb_group_step(bool fp32, int B, bool i16, const ChebGroupEntry* tab,
                           const int* owner, int total, int k, hipStream_t st);

// The partial sums of one grouped matrix: nblk block partials (B per block)
// summed as sum_block_partials does (G first-level blocks into scratch, then
// one; or directly when G == 1) into dots + (k+1)*B.
struct KpmSumEntry {
  const double* part;
  double* scratch;
  double* dots;
  int nblk, G;
  int k0, pad;
};
// One grouped KPM step k over `total` row blocks (owner[gb] = entry of row
// block gb) at block width B (lane layout of the kKpm policy), then its
// partial sums: level 1 over nl1 blocks (owner1/first1), level 2 over the
// two-level entries listed in owner2.
hipError_t kpm_group_step(bool fp32, int B, const KpmGroupEntry* tab, const int* owner,
                          int total, int k, const KpmSumEntry* sums, const int* owner1,
                          const int* first1, int nl1, const int* owner2, int nl2,
                          hipStream_t st);

// out[c] = sum_b part[b*B + c] over nblk row blocks, in a fixed order; uses
// scratch (>= kSumScratch doubles) for a first level when nblk is large
constexpr int kSumScratch = 256 * 32;
// first-level blocks of sum_block_partials for (nblk, B): 1 when it sums
// directly, else kSumScratch / 32
inline int sum_levels_blocks(int nblk, int B) {
  return ((size_t)nblk * B <= 16384) ? 1 : kSumScratch / 32;
}
hipError_t sum_block_partials(const double* part, int nblk, int B, double* out,
                              double* scratch, hipStream_t st);

// Fixed-grid reductions: partial buffers hold kRedBlocks entries.
constexpr int kRedBlocks = 1024;
// part[b] = partial sum of x.y
hipError_t dot_partial(size_t n, const double* x, const double* y, double* part,
                       hipStream_t st);
// out = sum part[0..cnt)  (single block)
hipError_t sum_scalar(const double* part, int cnt, double* out, hipStream_t st);
// y += a*x (a on host); part = partial of y.z   (z may be null: no dot)
hipError_t axpy_dot(size_t n, double a, const double* x, double* y, const double* z,
                    double* part, hipStream_t st);
// y += sign * (*a_d) * x; part = partial of y.y
hipError_t axpy_dev_nrm(size_t n, double sign, const double* a_d, const double* x,
                        double* y, double* part, hipStream_t st);
// y += sign * (*a_d) * x; part = partial of y.z
hipError_t axpy_dev_dot(size_t n, double sign, const double* a_d, const double* x,
                        double* y, const double* z, double* part, hipStream_t st);
// x *= a
hipError_t scal(size_t n, double a, double* x, hipStream_t st);
// y = a * x
hipError_t scal_copy(size_t n, double a, const double* x, double* y, hipStream_t st);
// x[i] = v
hipError_t fill(size_t n, double v, double* x, hipStream_t st);
hipError_t fill_u32(size_t n, unsigned v, unsigned* x, hipStream_t st);
// FP32 copy of (A - cc I) * scale (diagonal entries shifted); rows must
// hold a stored diagonal entry
hipError_t shift_scale_f32(const Csr& A, double cc, double scale, float* out, hipStream_t st);
// out[r] = (float)((d[r] - cc) * scale): the diagonal of the shifted FP32
// copy, with shift_scale_f32's arithmetic
hipError_t shift_diag_f32(int n, const double* d, double cc, double scale, float* out,
                          hipStream_t st);
// Z (k x k, column-major) = identity
hipError_t set_identity(int k, double* Z, hipStream_t st);
// precision conversions
hipError_t f64_to_f32(size_t n, const double* x, float* y, hipStream_t st);
hipError_t f32_to_f64(size_t n, const float* x, double* y, hipStream_t st);

// Classical Gram-Schmidt pass against the k columns of V (column-major, ld):
//   h = V^T w          (gemv_t: partials part[k][nch] then reduce)
//   w -= V h, part2 = partial of ||w||^2
int gemv_t_chunks(size_t n);
hipError_t gemv_t(size_t n, int k, const double* V, size_t ld, const double* w,
                  double* part, double* h, hipStream_t st);
hipError_t gemv_n_sub_nrm(size_t n, int k, const double* V, size_t ld, const double* h,
                          double* w, double* part, hipStream_t st);
// y = V h (no subtraction)
hipError_t gemv_n(size_t n, int k, const double* V, size_t ld, const double* h, double* y,
                  hipStream_t st);
// The same pass against an FP32 copy of the basis (ld even, columns 8-byte
// aligned), FP64 products and sums; gemv_n_sub_nrm_f32 leaves
// gemv_n_sub_f32_blocks(n) partials of ||w||^2 (one per 512 rows).
hipError_t gemv_t_f32(size_t n, int k, const float* V, size_t ld, const double* w,
                      double* part, double* h, hipStream_t st);
int gemv_n_sub_f32_blocks(size_t n);
hipError_t gemv_n_sub_nrm_f32(size_t n, int k, const float* V, size_t ld, const double* h,
                              double* w, double* part, hipStream_t st);
// x *= a, x32 = (float)x
hipError_t scal_f32copy(size_t n, double a, double* x, float* x32, hipStream_t st);

// C(n x m) = A(n x k) * B(k x m), all column-major, FP64
hipError_t gemm_nn(size_t n, int m, int k, const double* A, size_t lda, const double* B,
                   size_t ldb, double* C, size_t ldc, hipStream_t st);

// per-column reductions over m columns of n rows (column-major):
//   col_dot:  out[j] = X(:,j).Y(:,j)
//   col_res:  out[j] = || W(:,j) - t[j] U(:,j) ||^2
hipError_t col_dot(size_t n, int m, const double* X, size_t ldx, const double* Y,
                   size_t ldy, double* part, double* out, hipStream_t st);
hipError_t col_res2(size_t n, int m, const double* W, size_t ldw, const double* U,
                    size_t ldu, const double* t, double* part, double* out,
                    hipStream_t st);
// X(:,j) *= f(s[j]) with f = 1/sqrt(s) (normalise by squared norms)
hipError_t col_scale_rsqrt(size_t n, int m, double* X, size_t ldx, const double* s2,
                           hipStream_t st);
// U(:,j) = p[j] * U(:,j) + Y(:,j)
hipError_t col_scale_add(size_t n, int m, double* U, size_t ldu, const double* p, const double* Y,
                         size_t ldy, hipStream_t st);
// R(:,j) = (W(:,j) - th[j] * U(:,j)) * inv_dd
hipError_t col_resid(size_t n, int m, const double* W, size_t ldw, const double* U, size_t ldu,
                     const double* th, double inv_dd, double* R, size_t ldr, hipStream_t st);
// dst(:,j) = src(:, idx[j]) for j < m  (rows nrow)
hipError_t gather_cols(size_t nrow, int m, const double* src, size_t lds, const int* idx,
                       double* dst, size_t ldd, hipStream_t st);

// Apply the plane rotations recorded by the implicit-QL tridiagonal solver to
// the k x k column-major matrix Z (thread per row).  Sweep s covers rotation
// indices i = top[s], top[s]-1, ..., top[s]-cnt[s]+1 with cosines/sines
// cs[2*(off[s]+j)], cs[2*(off[s]+j)+1].
hipError_t apply_rotations(int k, double* Z, int nsweep, const int* top, const int* cnt,
                           const long long* off, const double* cs, hipStream_t st);

// block (row-major, B columns) helpers
// out[c] = ||X(:,c)||^2 partials -> finalised into out (B values)
hipError_t rm_col_nrm2(size_t n, int B, const double* X, double* part, double* out,
                       hipStream_t st);
// X(:,c) *= 1/sqrt(s2[c]) (zero columns left zero)
hipError_t rm_col_scale_rsqrt(size_t n, int B, double* X, const double* s2,
                              hipStream_t st);
// x(r*B+c) for r<n, c<B from Rademacher / Gaussian streams keyed by (seed, col0+c, r)
hipError_t rm_random(size_t n, int B, int kind, unsigned long long seed, int col0,
                     int ncols, double* X, hipStream_t st);

// Device addresses of up to kMaxCols FP64 column vectors, handed to the
// pack/unpack kernels by value (a kernel argument: no device table to upload)
constexpr int kMaxCols = 32;
struct ColPtrs {
  double* p[kMaxCols];
};
// Interleave B column vectors cols.p[0..B) into a row-major block X
// (r*B + c), converting to float if fp32; and the inverse.
hipError_t pack_cols(size_t n, int B, const ColPtrs& cols, void* X, bool fp32, hipStream_t st);
// dst (n x B2) = the first B2 columns of the interleaved n x B block src
// (separate buffers)
hipError_t narrow_cols(size_t n, int B, int B2, const void* src, void* dst, bool fp32,
                       hipStream_t st);
// cols.p[c - c0] (FP64 device vectors) = column c of the interleaved block, c0 <= c < c1
hipError_t unpack_range(size_t n, int B, int c0, int c1, const void* X, bool fp32,
                        const ColPtrs& cols, hipStream_t st);
hipError_t unpack_cols(size_t n, int B, const void* X, bool fp32, const ColPtrs& cols,
                       hipStream_t st);
// row-major block <-> column-major panel (ld) of B columns starting at col0
hipError_t rm_to_cm(size_t n, int B, const double* X, double* P, size_t ld, hipStream_t st);
hipError_t cm_to_rm(size_t n, int B, const double* P, size_t ld, double* X, hipStream_t st);

// C(m x p) = A(n x m)^T * B(n x p), all column-major FP64; tall-skinny
// (split over rows, partials summed by a second launch).  part must hold
// gemm_tn_part_elems(n, m, p) doubles.
size_t gemm_tn_part_elems(size_t n, int m, int p);
hipError_t gemm_tn(size_t n, int m, int p, const double* A, size_t lda, const double* B,
                   size_t ldb, double* C, size_t ldc, double* part, hipStream_t st);

}  // namespace evsl_dev
