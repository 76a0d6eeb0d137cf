// runtime.cpp -- context, memory helpers, operator registration, RNG,
// timing and statistics of the drop-in C interface.
//
// Mirrors SRC/evsl.c (EVSLStart/SetAMatrix/SetAMatvec), SRC/evsl_memory.c,
// SRC/timing.c, SRC/stats.c and SRC/vect.c:14-61 of EVSL 1.1.1.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "internal.h"

evslStat evslstat;

namespace evsl {

static std::string g_err;

void set_error(const char* where, hipError_t e) {
  if (g_err.empty()) g_err = std::string(where) + ": " + hipGetErrorString(e);
}
void set_error_msg(const std::string& msg) {
  if (g_err.empty()) g_err = msg;
}
const char* last_error() { return g_err.c_str(); }

hipError_t DevBuf::alloc(size_t nbytes) {
  if (nbytes <= bytes && p) return hipSuccess;
  release();
  if (nbytes == 0) return hipSuccess;
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    p = nullptr;
    return e;
  }
  bytes = nbytes;
  return hipSuccess;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

evsl_dev::Csr DevMatrix::view() const {
  evsl_dev::Csr c;
  c.n = n;
  c.nnz = nnz;
  c.ia = ia;
  c.ja = ja;
  c.jd = jd;
  c.a64 = a64;
  c.a32 = a32;
  if (pc) {
    c.pc = pc;
    c.pdpos = pdpos;
    c.poff = poff;
    c.pshort = pshort;
    c.pval64 = pval64;
    c.dg64 = dg64;
    c.pval32 = pval32;
    c.dg32 = dg32;
  }
  for (int w = 0; w < evsl_dev::kNumWidths; ++w) {
    c.rb[w] = rb[w];
    c.nblk[w] = nblk[w];
  }
  c.tune = ctx().spmm_tuning ? tune : nullptr;
  return c;
}

void DevMatrix::release() {
  if (owns_csr) {
    if (ia) (void)hipFree(ia);
    if (ja) (void)hipFree(ja);
    if (a64) (void)hipFree(a64);
  }
  for (int w = 0; w < evsl_dev::kNumWidths; ++w)
    if (rb[w]) (void)hipFree(rb[w]);
  if (jd) (void)hipFree(jd);
  if (a32) (void)hipFree(a32);
  if (a32s) (void)hipFree(a32s);
  for (void* p : {(void*)pc, (void*)pdpos, (void*)poff, (void*)pval64, (void*)dg64, (void*)pval32,
                  (void*)dg32, (void*)pval32s, (void*)dg32s})
    if (p) (void)hipFree(p);
  evsl_dev::spmm_tuner_destroy(tune);
  *this = DevMatrix();
}

std::vector<int> make_row_blocks(const int* ia, int n, int rows_cap) {
  std::vector<int> rb;
  rb.reserve(n / std::max(1, rows_cap) + 2);
  int r = 0;
  while (r < n) {
    rb.push_back(r);
    const int len0 = ia[r + 1] - ia[r];
    if (len0 > evsl_dev::kNnzCap) {  // a long row gets a block of its own
      r += 1;
      continue;
    }
    int rows = 0, nz = 0;
    while (r < n && rows < rows_cap) {
      const int len = ia[r + 1] - ia[r];
      if (len > evsl_dev::kNnzCap || nz + len > evsl_dev::kNnzCap) break;
      nz += len;
      ++rows;
      ++r;
    }
  }
  rb.push_back(n);
  return rb;
}

static Context g_ctx;
Context& ctx() { return g_ctx; }

// Called first by every device entry point: the error message then belongs
// to the first failure of the current call.
void clear_error() { g_err.clear(); }

int ensure_started() {
  Context& c = g_ctx;
  g_err.clear();
  if (c.started) return 0;
  return EVSLStart();
}

double* red_part() { return g_ctx.red_part.as<double>(); }
double* red_scalar(int i) { return g_ctx.red_scalars.as<double>() + i; }

int read_scalars(const double* d, int cnt, double* out) {
  Context& c = g_ctx;
  EVSL_HIP(hipMemcpyAsync(c.pinned, d, cnt * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  EVSL_HIP(hipStreamSynchronize(c.stream));
  std::memcpy(out, c.pinned, cnt * sizeof(double));
  return 0;
}

int ensure_a32() {
  Context& c = g_ctx;
  if (c.op != 1) return EVSL_ERR_NOMAT;
  if (c.A.a32) return 0;
  EVSL_HIP(hipMalloc((void**)&c.A.a32, (size_t)std::max(c.A.nnz, 1) * sizeof(float)));
  EVSL_HIP(evsl_dev::f64_to_f32(c.A.nnz, c.A.a64, c.A.a32, c.stream));
  if (c.A.pc) {  // the pattern tables in FP32: (float) of each value, as a32
    std::vector<float> pv(evsl_dev::kPatCodes + 1, 0.0f);
    for (int i = 0; i < c.A.npat; ++i) pv[i] = (float)c.A.pval_h[i];
    EVSL_HIP(hipMalloc((void**)&c.A.pval32, pv.size() * sizeof(float)));
    EVSL_HIP(hipMemcpy(c.A.pval32, pv.data(), pv.size() * sizeof(float), hipMemcpyHostToDevice));
    EVSL_HIP(hipMalloc((void**)&c.A.dg32, (size_t)std::max(c.A.n, 1) * sizeof(float)));
    EVSL_HIP(evsl_dev::f64_to_f32(c.A.n, c.A.dg64, c.A.dg32, c.stream));
  }
  return 0;
}

int apply_A(const double* x, double* y) {
  Context& c = g_ctx;
  if (c.op == 1) {
    EVSL_HIP(evsl_dev::cheb_spmm(c.A.view(), 1, false, evsl_dev::kPlain, x, nullptr, y,
                                 nullptr, nullptr, nullptr, nullptr, 1.0, 0.0, false, 0,
                                 nullptr, c.stream));
    return 0;
  }
  if (c.op == 2) {
    EVSL_HIP(hipStreamSynchronize(c.stream));
    c.func(const_cast<double*>(x), y, c.func_data);
    EVSL_HIP(hipDeviceSynchronize());
    return 0;
  }
  set_error_msg("no operator: call SetAMatrix or SetAMatvec first");
  return EVSL_ERR_NOMAT;
}

// A device operator from a device CSR (arrays already on the device): the
// row-block tables of every block width, 16-bit column offsets when they fit,
// and whether every row stores its diagonal; *red_need = the reduction
// scratch (doubles) its launches need
namespace {

// Pattern codes of a host CSR: every off-diagonal entry's (column offset,
// value) pair gets a code in [0, kPatCodes) in order of first appearance
// (values told apart by their bits), the diagonal entry kPatDiag with its
// value in dg and its position within the row in dpos.  False when there
// are more than kPatCodes distinct pairs, a row stores its diagonal twice
// or past its 254th entry.
bool encode_pattern(int n, const int* ia, const int* ja, const double* a,
                    std::vector<unsigned char>& code, std::vector<int>& off,
                    std::vector<double>& val, std::vector<double>& dg,
                    std::vector<unsigned char>& dpos) {
  constexpr int kSlots = 1024;  // open addressing, four times the table
  struct Slot {
    long long off;
    std::uint64_t bits;
    int code;
  };
  std::vector<Slot> h(kSlots, Slot{0, 0, -1});
  code.resize((size_t)std::max(ia[n], 1));
  dg.assign((size_t)n, 0.0);
  dpos.assign((size_t)n, evsl_dev::kPatDiag);
  off.clear();
  val.clear();
  for (int r = 0; r < n; ++r) {
    bool diag = false;
    for (int k = ia[r]; k < ia[r + 1]; ++k) {
      const long long o = (long long)ja[k] - r;
      if (o == 0) {
        if (diag || k - ia[r] >= evsl_dev::kPatDiag) return false;
        diag = true;
        dg[r] = a[k];
        dpos[r] = (unsigned char)(k - ia[r]);
        code[k] = evsl_dev::kPatDiag;
        continue;
      }
      std::uint64_t bits;
      std::memcpy(&bits, &a[k], sizeof(bits));
      const std::uint64_t mix =
          (std::uint64_t)o * 0x9E3779B97F4A7C15ull ^ (bits * 0xC2B2AE3D27D4EB4Full);
      int s = (int)((mix >> 32) & (kSlots - 1));
      while (h[s].code >= 0 && !(h[s].off == o && h[s].bits == bits)) s = (s + 1) & (kSlots - 1);
      if (h[s].code < 0) {
        if ((int)off.size() == evsl_dev::kPatCodes) return false;
        h[s] = Slot{o, bits, (int)off.size()};
        off.push_back((int)o);
        val.push_back(a[k]);
      }
      code[k] = (unsigned char)h[s].code;
    }
  }
  return true;
}

}  // namespace

int build_device_matrix(int n, int nnz, int* ia_d, int* ja_d, double* a_d, const int* ia_h,
                        const int* ja_h, bool owns, DevMatrix& M, size_t* red_need) {
  M.release();
  M.n = n;
  M.nnz = nnz;
  M.ia = ia_d;
  M.ja = ja_d;
  M.a64 = a_d;
  M.owns_csr = owns;
  std::vector<int> ia_tmp, ja_tmp;
  if (!ia_h) {
    ia_tmp.resize(n + 1);
    EVSL_HIP(hipMemcpy(ia_tmp.data(), ia_d, (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    ia_h = ia_tmp.data();
  }
  if (!ja_h) {
    ja_tmp.resize(std::max(nnz, 1));
    EVSL_HIP(hipMemcpy(ja_tmp.data(), ja_d, (size_t)nnz * sizeof(int), hipMemcpyDeviceToHost));
    ja_h = ja_tmp.data();
  }
  bool full = true;
  for (int r = 0; r < n && full; ++r) {
    bool has = false;
    for (int k = ia_h[r]; k < ia_h[r + 1]; ++k)
      if (ja_h[k] == r) {
        has = true;
        break;
      }
    full = has;
  }
  M.full_diag = full;
  M.tune = evsl_dev::spmm_tuner_create();
  if (!M.tune) {
    set_error_msg("out of host memory for the layout tuner");
    return EVSL_ERR_DEVICE;
  }
  // 16-bit column offsets when every |ja - row| fits (banded and stencil
  // matrices in a local ordering): 2 bytes less per nonzero and filter step
  {
    bool fits = true;
    for (int r = 0; r < n && fits; ++r)
      for (int k = ia_h[r]; k < ia_h[r + 1]; ++k) {
        const long d = (long)ja_h[k] - r;
        if (d < -32768 || d > 32767) {
          fits = false;
          break;
        }
      }
    if (fits && nnz > 0) {
      std::vector<short> jd((size_t)nnz);
      for (int r = 0; r < n; ++r)
        for (int k = ia_h[r]; k < ia_h[r + 1]; ++k) jd[k] = (short)(ja_h[k] - r);
      EVSL_HIP(hipMalloc((void**)&M.jd, (size_t)nnz * sizeof(short)));
      EVSL_HIP(hipMemcpy(M.jd, jd.data(), (size_t)nnz * sizeof(short), hipMemcpyHostToDevice));
    }
  }
  // pattern codes when the off-diagonal entries take few distinct
  // (offset, value) pairs: 1 byte per nonzero for the filter kernels
  if (ctx().pattern_codes && nnz > 0) {
    std::vector<double> a_h((size_t)nnz);
    EVSL_HIP(hipMemcpy(a_h.data(), a_d, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<unsigned char> code, dpos;
    std::vector<int> off;
    std::vector<double> val, dg;
    if (encode_pattern(n, ia_h, ja_h, a_h.data(), code, off, val, dg, dpos)) {
      M.npat = (int)off.size();
      M.pshort = true;
      for (int o : off) M.pshort = M.pshort && o >= -32768 && o <= 32767;
      M.pval_h = val;
      off.resize(evsl_dev::kPatCodes + 1, 0);
      val.resize(evsl_dev::kPatCodes + 1, 0.0);
      EVSL_HIP(hipMalloc((void**)&M.pc, (size_t)nnz));
      EVSL_HIP(hipMemcpy(M.pc, code.data(), (size_t)nnz, hipMemcpyHostToDevice));
      EVSL_HIP(hipMalloc((void**)&M.poff, off.size() * sizeof(int)));
      EVSL_HIP(hipMemcpy(M.poff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
      EVSL_HIP(hipMalloc((void**)&M.pval64, val.size() * sizeof(double)));
      EVSL_HIP(hipMemcpy(M.pval64, val.data(), val.size() * sizeof(double),
                         hipMemcpyHostToDevice));
      EVSL_HIP(hipMalloc((void**)&M.pdpos, (size_t)n));
      EVSL_HIP(hipMemcpy(M.pdpos, dpos.data(), (size_t)n, hipMemcpyHostToDevice));
      EVSL_HIP(hipMalloc((void**)&M.dg64, (size_t)n * sizeof(double)));
      EVSL_HIP(hipMemcpy(M.dg64, dg.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  size_t need = (size_t)evsl_dev::kRedBlocks * 64;
  for (int w = 0; w < evsl_dev::kNumWidths; ++w) {
    const int B = evsl_dev::width_of_slot(w);
    std::vector<int> rb = make_row_blocks(ia_h, n, evsl_dev::rows_cap_for(B));
    M.nblk[w] = (int)rb.size() - 1;
    std::vector<int> rbp(2 * rb.size());  // (row, ia[row]) pairs
    for (size_t i = 0; i < rb.size(); ++i) {
      rbp[2 * i] = rb[i];
      rbp[2 * i + 1] = ia_h[rb[i]];
    }
    EVSL_HIP(hipMalloc((void**)&M.rb[w], rbp.size() * sizeof(int)));
    EVSL_HIP(hipMemcpy(M.rb[w], rbp.data(), rbp.size() * sizeof(int), hipMemcpyHostToDevice));
    need = std::max(need, (size_t)(M.nblk[w] + 1) * B);
  }
  // reduction scratch: the largest per-row-block partial buffer, the
  // first-level sums behind it, and one partial per 256 rows
  *red_need = std::max(need, (size_t)n / evsl_dev::kWG + 1);
  return 0;
}

// the context's reduction scratch holds at least `need` doubles (plus the
// first-level sums behind them)
int ensure_red_part(size_t need) {
  Context& c = g_ctx;
  const size_t bytes = (need + evsl_dev::kSumScratch) * sizeof(double);
  if (c.red_part.bytes >= bytes) return 0;
  EVSL_HIP(c.red_part.alloc(bytes));
  return 0;
}

static int register_device_csr(int n, int nnz, int* ia_d, int* ja_d, double* a_d,
                               const int* ia_h, const int* ja_h, bool owns) {
  Context& c = g_ctx;
  size_t need = 0;
  if (int e = build_device_matrix(n, nnz, ia_d, ja_d, a_d, ia_h, ja_h, owns, c.A, &need)) {
    c.n = 0;
    c.op = 0;
    return e;
  }
  c.n = n;
  c.op = 1;
  return ensure_red_part(need);
}

int ensure_a32_shifted(double cc, double scale, bool* ok) {
  Context& c = g_ctx;
  *ok = false;
  if (c.op != 1 || !c.A.full_diag) return 0;
  if (c.A.a32s && c.A.shift_cc == cc && c.A.shift_scale == scale) {
    *ok = true;
    return 0;
  }
  if (!c.A.a32s)
    EVSL_HIP(hipMalloc((void**)&c.A.a32s, (size_t)std::max(c.A.nnz, 1) * sizeof(float)));
  EVSL_HIP(evsl_dev::shift_scale_f32(c.A.view(), cc, scale, c.A.a32s, c.stream));
  if (c.A.pc) {
    // the pattern tables of the shifted copy, with the arithmetic of
    // shift_scale_f32: (float)(v * scale) off the diagonal (every table
    // entry), (float)((d - cc) * scale) on it
    std::vector<float> pv(evsl_dev::kPatCodes + 1, 0.0f);
    for (int i = 0; i < c.A.npat; ++i) pv[i] = (float)(c.A.pval_h[i] * scale);
    if (!c.A.pval32s)
      EVSL_HIP(hipMalloc((void**)&c.A.pval32s, pv.size() * sizeof(float)));
    EVSL_HIP(hipMemcpy(c.A.pval32s, pv.data(), pv.size() * sizeof(float), hipMemcpyHostToDevice));
    if (!c.A.dg32s) EVSL_HIP(hipMalloc((void**)&c.A.dg32s, (size_t)std::max(c.A.n, 1) * sizeof(float)));
    EVSL_HIP(evsl_dev::shift_diag_f32(c.A.n, c.A.dg64, cc, scale, c.A.dg32s, c.stream));
  }
  c.A.shift_cc = cc;
  c.A.shift_scale = scale;
  *ok = true;
  return 0;
}

bool filter_op_ready(bool f32) {
  const Context& c = g_ctx;
  if (c.op != 1) return false;
  if (!f32) return true;
  return c.A.full_diag ? c.A.a32s != nullptr : c.A.a32 != nullptr;
}

int make_filter_op(double cc, double dd, bool f32, FilterOp* op) {
  Context& c = g_ctx;
  op->A = c.A.view();
  op->f32 = f32;
  op->t1 = 1.0 / dd;
  op->t2 = 2.0 / dd;
  op->cc = cc;
  if (!f32) return 0;
  const double scale = std::ldexp(1.0, -(int)std::lround(std::log2(dd)));
  bool ok = false;
  if (int e = ensure_a32_shifted(cc, scale, &ok)) return e;
  if (ok) {
    op->A.a32 = c.A.a32s;
    op->A.pval32 = c.A.pval32s;
    op->A.dg32 = c.A.dg32s;
    op->t1 /= scale;
    op->t2 /= scale;
    op->cc = 0.0;
    return 0;
  }
  if (int e = ensure_a32()) return e;  // no stored diagonal: unshifted FP32 values
  op->A.a32 = c.A.a32;
  op->A.pval32 = c.A.pval32;  // made with a32 (the view above predates them)
  op->A.dg32 = c.A.dg32;
  return 0;
}

double* red_scratch() {
  return g_ctx.red_part.as<double>() + (g_ctx.red_part.bytes / sizeof(double) - evsl_dev::kSumScratch);
}

}  // namespace evsl

using namespace evsl;

extern "C" {

// ---- memory (SRC/evsl_memory.c:9-149) -----------------------------------
void* _evsl_Malloc(size_t nbytes) { return std::malloc(nbytes); }
void* _evsl_Calloc(size_t count, size_t nbytes) { return std::calloc(count, nbytes); }
void* _evsl_Realloc(void* ptr, size_t nbytes) { return std::realloc(ptr, nbytes); }
void _evsl_Free(void* ptr) { std::free(ptr); }

void* _evsl_Malloc_device(size_t nbytes) {
  void* p = nullptr;
  if (nbytes == 0) return nullptr;
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    set_error("hipMalloc", e);
    return nullptr;
  }
  return p;
}

void* _evsl_Calloc_device(size_t count, size_t nbytes) {
  void* p = _evsl_Malloc_device(count * nbytes);
  if (p && hipMemset(p, 0, count * nbytes) != hipSuccess) {
    (void)hipFree(p);
    return nullptr;
  }
  return p;
}

void* _evsl_Realloc_device(void* old_ptr, size_t old_nbytes, size_t new_nbytes) {
  void* p = _evsl_Malloc_device(new_nbytes);
  if (p && old_ptr) {
    (void)hipMemcpy(p, old_ptr, std::min(old_nbytes, new_nbytes), hipMemcpyDeviceToDevice);
    (void)hipFree(old_ptr);
  }
  return p;
}

void _evsl_Free_device(void* ptr) {
  if (ptr) (void)hipFree(ptr);
}

void evsl_memcpy_device_to_host(void* dst, void* src, size_t nbytes) {
  hipError_t e = hipMemcpy(dst, src, nbytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) set_error("hipMemcpy D2H", e);
}
void evsl_memcpy_device_to_device(void* dst, void* src, size_t nbytes) {
  hipError_t e = hipMemcpy(dst, src, nbytes, hipMemcpyDeviceToDevice);
  if (e != hipSuccess) set_error("hipMemcpy D2D", e);
}
void evsl_memcpy_host_to_device(void* dst, void* src, size_t nbytes) {
  hipError_t e = hipMemcpy(dst, src, nbytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) set_error("hipMemcpy H2D", e);
}
void evsl_memset_device(void* ptr, int value, size_t nbytes) {
  hipError_t e = hipMemset(ptr, value, nbytes);
  if (e != hipSuccess) set_error("hipMemset", e);
}

// ---- lifecycle (SRC/evsl.c:24-93) -----------------------------------------
int EVSLStart(void) {
  Context& c = ctx();
  if (c.started) return 0;
  int dev = 0;
  EVSL_HIP(hipGetDevice(&dev));
  EVSL_HIP(hipFree(nullptr));  // create the primary context
  c.stream = nullptr;          // legacy null stream
  EVSL_HIP(c.red_part.alloc((size_t)evsl_dev::kRedBlocks * 64 * sizeof(double)));
  EVSL_HIP(c.red_scalars.alloc(64 * sizeof(double)));
  EVSL_HIP(hipHostMalloc((void**)&c.pinned, 4096 * sizeof(double), hipHostMallocDefault));
  // the allocation-free entry point's scratch on this device: its address
  // is looked up here once (the entry point retries on its first call if
  // this fails, and reports the error there)
  void* scratch = nullptr;
  if (hipError_t e = evsl_dev::noalloc_scratch(&scratch); e != hipSuccess) (void)hipGetLastError();
  c.started = true;
  StatsReset();
  std::memset(&c.astats, 0, sizeof(c.astats));
  return 0;
}

int EVSLFinish(void) {
  Context& c = ctx();
  if (!c.started) return 0;
  (void)hipDeviceSynchronize();
  c.A.release();
  c.red_part.release();
  c.red_scalars.release();
  if (c.pinned) (void)hipHostFree(c.pinned);
  c.pinned = nullptr;
  c.op = 0;
  c.n = 0;
  c.func = nullptr;
  c.func_data = nullptr;
  c.started = false;
  return 0;
}

// Host CSR: uploaded to the device (FP64 values; FP32 copy made on demand).
// The reference keeps the pointer (SRC/evsl.c:95-101); here the device copy
// is taken at this call, so later host edits of A are not seen.
int SetAMatrix(csrMat* A) {
  if (int e = ensure_started()) return e;
  if (!A || A->nrows != A->ncols || A->nrows <= 0) return EVSL_ERR_ARG;
  const int n = A->nrows, nnz = A->ia[n];
  int *ia = nullptr, *ja = nullptr;
  double* a = nullptr;
  EVSL_HIP(hipMalloc((void**)&ia, (n + 1) * sizeof(int)));
  EVSL_HIP(hipMalloc((void**)&ja, (size_t)std::max(nnz, 1) * sizeof(int)));
  EVSL_HIP(hipMalloc((void**)&a, (size_t)std::max(nnz, 1) * sizeof(double)));
  EVSL_HIP(hipMemcpy(ia, A->ia, (n + 1) * sizeof(int), hipMemcpyHostToDevice));
  EVSL_HIP(hipMemcpy(ja, A->ja, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
  EVSL_HIP(hipMemcpy(a, A->a, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice));
  return register_device_csr(n, nnz, ia, ja, a, A->ia, A->ja, true);
}

// A->ia/ja/a are device pointers (as produced by evsl_create_csr_gpu)
int SetAMatrix_device_csr(csrMat* A) {
  if (int e = ensure_started()) return e;
  if (!A || A->nrows != A->ncols || A->nrows <= 0) return EVSL_ERR_ARG;
  int nnz = 0;
  EVSL_HIP(hipMemcpy(&nnz, A->ia + A->nrows, sizeof(int), hipMemcpyDeviceToHost));
  return register_device_csr(A->nrows, nnz, A->ia, A->ja, A->a, nullptr, nullptr, false);
}

void evsl_create_csr_gpu(csrMat* Acpu, csrMat* Agpu) {
  const int n = Acpu->nrows, nnz = Acpu->ia[n];
  Agpu->owndata = 1;
  Agpu->nrows = Acpu->nrows;
  Agpu->ncols = Acpu->ncols;
  Agpu->nnz = nnz;
  Agpu->ia = (int*)_evsl_Malloc_device((n + 1) * sizeof(int));
  Agpu->ja = (int*)_evsl_Malloc_device((size_t)std::max(nnz, 1) * sizeof(int));
  Agpu->a = (double*)_evsl_Malloc_device((size_t)std::max(nnz, 1) * sizeof(double));
  evsl_memcpy_host_to_device(Agpu->ia, Acpu->ia, (n + 1) * sizeof(int));
  evsl_memcpy_host_to_device(Agpu->ja, Acpu->ja, (size_t)nnz * sizeof(int));
  evsl_memcpy_host_to_device(Agpu->a, Acpu->a, (size_t)nnz * sizeof(double));
}

void evsl_free_csr_gpu(csrMat* csr) {
  Context& c = ctx();
  if (c.op == 1 && c.A.ia == csr->ia) {  // drop the registration first
    c.A.release();
    c.op = 0;
  }
  _evsl_Free_device(csr->ia);
  _evsl_Free_device(csr->ja);
  _evsl_Free_device(csr->a);
  csr->ia = csr->ja = nullptr;
  csr->a = nullptr;
}

// user operator: func(x, y, data) is called with device vectors
int SetAMatvec(int n, MVFunc func, void* data) {
  if (int e = ensure_started()) return e;
  if (n <= 0 || !func) return EVSL_ERR_ARG;
  Context& c = ctx();
  c.A.release();
  c.n = n;
  c.op = 2;
  EVSL_HIP(c.red_part.alloc(((size_t)evsl_dev::kRedBlocks * 64 + (size_t)n / evsl_dev::kWG + 1 +
                             evsl_dev::kSumScratch) * sizeof(double)));
  c.func = func;
  c.func_data = data;
  return 0;
}

// generalized problems are outside this library's device path
int SetBMatrix(csrMat*) { return EVSL_ERR_UNSUPP; }
int SetBMatvec(int, MVFunc, void*) { return EVSL_ERR_UNSUPP; }
int SetBSol(SolFuncR, void*) { return EVSL_ERR_UNSUPP; }
int SetLTSol(SolFuncR, void*) { return EVSL_ERR_UNSUPP; }
int SetStdEig(void) { return 0; }
int SetGenEig(void) { return EVSL_ERR_UNSUPP; }
void SetDiagScal(double*) {}

// ---- RNG (SRC/vect.c:14-61) ---------------------------------------------
void rand_double(int n, double* v) {
  const double t = ((double)RAND_MAX) / 2.0;
  for (int i = 0; i < n; i++) v[i] = (rand() - t) / t;
}

void randn_double(int n, double* v) {
  const double two_pi = 2.0 * 3.1415926535;
  for (int i = 0; i < n; i++) {
    const double u1 = rand() * (1.0 / RAND_MAX);
    const double u2 = rand() * (1.0 / RAND_MAX);
    v[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
  }
}

// host libc stream, then one upload (the reference CUDA build's default
// generation method, SRC/evsl_cuda.cu:113-116)
void rand_double_device(int n, double* v) {
  std::vector<double> h(n);
  rand_double(n, h.data());
  evsl_memcpy_host_to_device(v, h.data(), (size_t)n * sizeof(double));
}

void randn_double_device(int n, double* v) {
  std::vector<double> h(n);
  randn_double(n, h.data());
  evsl_memcpy_host_to_device(v, h.data(), (size_t)n * sizeof(double));
}

// ---- timing and stats (SRC/timing.c:60-74, SRC/stats.c) --------------------
// CLOCK_MONOTONIC; like the reference's GPU build it first waits for the device
double evsl_timer(void) {
  (void)hipDeviceSynchronize();
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

void StatsReset(void) { std::memset(&evslstat, 0, sizeof(evslstat)); }

void StatsPrint(FILE* f) {
  if (!f) return;
  const evslStat* s = &evslstat;
  std::fprintf(f, " Timing (sec):\n");
  if (s->t_iter) std::fprintf(f, "   Iteration time (tot)     :  %f\n", s->t_iter);
  std::fprintf(f, "   - - - - - - - - - - - - - - - - -\n");
  if (s->n_polAv)
    std::fprintf(f, "   Pol(A)*v                 :  %f (%8zu, avg %f)\n", s->t_polAv,
                 s->n_polAv, s->t_polAv / s->n_polAv);
  if (s->n_mvA)
    std::fprintf(f, "   Matvec matrix A          :  %f (%8zu, avg %f)\n", s->t_mvA, s->n_mvA,
                 s->t_mvA / s->n_mvA);
  if (s->t_reorth) std::fprintf(f, "   Reorthogonalization      :  %f\n", s->t_reorth);
  if (s->t_eig) std::fprintf(f, "   LAPACK eig               :  %f\n", s->t_eig);
  if (s->t_ritz) std::fprintf(f, "   Compute Ritz vectors     :  %f\n", s->t_ritz);
  std::fflush(f);
}

// ---- extension: precision, stats, errors --------------------------------
int evsl_amd_set_filter_precision(int prec) {
  if (prec != EVSL_PREC_FP64 && prec != EVSL_PREC_MIXED) return EVSL_ERR_ARG;
  ctx().filter_prec = prec;
  return 0;
}
int evsl_amd_get_filter_precision(void) { return ctx().filter_prec; }
int evsl_amd_set_refine_passes(int passes) {
  if (passes < 1 || passes > 8) return EVSL_ERR_ARG;
  ctx().refine_passes = passes;
  return 0;
}
int evsl_amd_set_refine_guard(double guard) {
  if (!(guard > 0.0 && guard <= 1.0)) return EVSL_ERR_ARG;
  ctx().refine_guard = guard;
  return 0;
}
int evsl_amd_set_grouped_filter(int on) {
  ctx().grouped_filter = on != 0;
  return 0;
}
int evsl_amd_set_refine_widen(double widen) {
  if (!(widen >= 0.0 && widen <= 1e6)) return EVSL_ERR_ARG;
  ctx().refine_widen = widen;
  return 0;
}
int evsl_amd_set_mixed_capped_fp64(int on) {
  ctx().mixed_capped_fp64 = on != 0;
  return 0;
}
int evsl_amd_set_fp32_basis(int on) {
  ctx().fp32_basis = on != 0;
  return 0;
}
int evsl_amd_set_mixed_degree_range(int min_deg, int max_deg) {
  if (min_deg < 0 || max_deg < min_deg) return EVSL_ERR_ARG;
  ctx().mixed_min_deg = min_deg;
  ctx().mixed_max_deg = max_deg;
  return 0;
}
int evsl_amd_get_stats(evslAmdStats* s) {
  if (!s) return EVSL_ERR_ARG;
  *s = ctx().astats;
  return 0;
}
void evsl_amd_reset_stats(void) { std::memset(&ctx().astats, 0, sizeof(evslAmdStats)); }
const char* evsl_amd_last_error(void) { return last_error(); }
size_t evsl_amd_stats_size(void) { return sizeof(evslAmdStats); }

}  // extern "C"
