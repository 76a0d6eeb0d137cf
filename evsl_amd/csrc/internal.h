// internal.h -- host-side state shared by the C-ABI entry points.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdio>
#include <string>
#include <memory>
#include <vector>

#include "../../include/evsl.h"
#include "../../include/evsl_amd.h"
#include "dev.h"

namespace evsl {

// ---- error handling ------------------------------------------------------
// HIP failures are recorded (first message kept) and turned into
// EVSL_ERR_DEVICE by the entry points; nothing calls exit() or assert().
void set_error(const char* where, hipError_t e);
void set_error_msg(const std::string& msg);
const char* last_error();

#define EVSL_HIP(call)                                   \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) {                              \
      ::evsl::set_error(#call, e_);                      \
      return EVSL_ERR_DEVICE;                            \
    }                                                    \
  } while (0)

// ---- device buffers ----------------------------------------------------
// RAII device allocation used inside the solvers.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  hipError_t alloc(size_t nbytes);
  void release();
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// ---- operator ----------------------------------------------------------
struct DevMatrix {
  int n = 0, nnz = 0;
  int *ia = nullptr, *ja = nullptr;
  short* jd = nullptr;             // 16-bit column offsets (when they fit)
  double* a64 = nullptr;
  float* a32 = nullptr;            // FP32 copy of the values (unshifted)
  float* a32s = nullptr;           // FP32 copy of (A - shift_cc I) * shift_scale
  double shift_cc = 0.0, shift_scale = 0.0;
  bool full_diag = false;          // every row stores its diagonal entry
  int* rb[evsl_dev::kNumWidths] = {};
  int nblk[evsl_dev::kNumWidths] = {};
  evsl_dev::SpmmTuner* tune = nullptr;  // the filter kernel's layout tuner
  bool owns_csr = false;  // ia/ja/a64 allocated here (SetAMatrix)
  // pattern codes (evsl_dev::Csr::pc): one byte per nonzero, tables of at
  // most kPatCodes (column offset, value) pairs, the diagonal per row; the
  // FP32 tables go with a32 (unshifted) and a32s (shifted)
  unsigned char* pc = nullptr;
  unsigned char* pdpos = nullptr;  // the diagonal's position in its row
  int* poff = nullptr;
  bool pshort = false;
  int npat = 0;                  // table entries in use
  std::vector<double> pval_h;    // host copy of the FP64 table values
  double* pval64 = nullptr;
  double* dg64 = nullptr;
  float* pval32 = nullptr;
  float* dg32 = nullptr;
  float* pval32s = nullptr;
  float* dg32s = nullptr;
  // bytes per nonzero of the column index the filter kernels read
  int idx_bytes() const { return pc ? 1 : (jd ? 2 : 4); }
  // matrix bytes one filter launch of B columns streams with values of vs
  // bytes: the codes, the diagonal and its position, or the indices and the
  // values (evsl_dev::pattern_launch)
  double mat_bytes(int vs, int B) const {
    return (pc && B * vs <= evsl_dev::kPatMaxRowBytes) ? (double)nnz + (double)n * (vs + 1)
                                                       : (double)nnz * (vs + (jd ? 2 : 4));
  }
  evsl_dev::Csr view() const;
  void release();
};

// Build the LDS row-block partition (at most rows_cap rows and kNnzCap
// nonzeros per block) from host row pointers.
std::vector<int> make_row_blocks(const int* ia_h, int n, int rows_cap);
// A device operator M from a device CSR: row-block tables of every width,
// 16-bit column offsets when they fit, the stored-diagonal flag; *red_need =
// reduction scratch (doubles) its launches need (see ensure_red_part)
int build_device_matrix(int n, int nnz, int* ia_d, int* ja_d, double* a_d, const int* ia_h,
                        const int* ja_h, bool owns, DevMatrix& M, size_t* red_need);
int ensure_red_part(size_t need);
// make sure the shifted FP32 copy for (cc, scale) exists; false if the
// matrix lacks stored diagonal entries
int ensure_a32_shifted(double cc, double scale, bool* ok);

struct Context {
  bool started = false;
  int n = 0;
  int op = 0;  // 0 none, 1 device CSR, 2 user MVFunc
  DevMatrix A;
  MVFunc func = nullptr;
  void* func_data = nullptr;
  hipStream_t stream = nullptr;  // the legacy null stream (serialises with user code)
  int filter_prec = EVSL_PREC_FP64;
  int refine_passes = 1;
  double refine_guard = 0.8;
  // the first refinement pass filters with the polynomial of the slice's
  // interval widened by refine_widen slice widths on each side (lower
  // degree); 0 uses the slice's own filter
  double refine_widen = 4.0;
  int mixed_min_deg = 32;    // mixed form: filters of degree outside
  int mixed_max_deg = 20000; // [mixed_min_deg, mixed_max_deg] run in FP64
  bool mixed_capped_fp64 = true;  // mixed form: capped filters run in FP64
  // mixed form: the Gram-Schmidt passes of a slice with an FP32 filter read
  // an FP32 copy of its Lanczos basis (FP64 products and sums)
  bool fp32_basis = true;
  bool spmm_tuning = true;   // filter launches go through the matrix's layout tuner
  bool pattern_codes = true; // matrices built from now on use pattern codes when they can
  // the request loop filters the slices of all its matrices in grouped
  // launches (else one matrix at a time)
  bool grouped_filter = true;
  // scratch for reductions
  DevBuf red_part, red_scalars;
  // pinned host scalars for device->host reads
  double* pinned = nullptr;
  evslAmdStats astats;
};

Context& ctx();
// forget the previous call's error message
void clear_error();
int ensure_started();
// make sure the FP32 copy of the matrix values exists
int ensure_a32();
// y = A x for device vectors (CSR kernel or user function)
int apply_A(const double* x, double* y);
// device scratch for reductions (>= kRedBlocks * 64 doubles, plus scalars)
double* red_part();
double* red_scalar(int i);
// kSumScratch doubles at the end of the reduction scratch (two-level sums)
double* red_scratch();
// copy cnt doubles from device to the pinned host buffer and wait
int read_scalars(const double* d, int cnt, double* out);

// ---- tridiagonal eigensolver (tridiag.cpp) ----------------------------
// Implicit QL with a Wilkinson-type shift.  Eigenvalues ascending in d.
// If rec != nullptr the plane rotations are recorded for replay on the
// device (the eigenvectors are then Z = I * rotations, columns permuted by
// rec->perm to ascending order).
struct RotationLog {
  std::vector<int> top, cnt;
  std::vector<long long> off;
  std::vector<double> cs;
  std::vector<int> perm;  // sorted position j holds original column perm[j]
  void clear() {
    top.clear();
    cnt.clear();
    off.clear();
    cs.clear();
    perm.clear();
  }
};
int tridiag_ql(int n, double* d, double* e, RotationLog* rec);
// eigenvalues and eigenvectors of a small symmetric tridiagonal on the host
// (column-major n x n Z); used where T is tiny.
int tridiag_eig_host(int n, const double* diag, const double* sdiag, double* eval,
                     double* Z);
// dense symmetric eigensolver on the host (Householder + QL), column-major
int sym_eig_host(int n, const double* A, int lda, double* Q, int ldq, double* lam);

// Replay a rotation log on the device: Zsel(:, j) = eigenvector of T at
// sorted position sel[j] (kdim x m, column-major, leading dimension kdim).
int tridiag_rotations_to_device(int kdim, const RotationLog& rec, const std::vector<int>& sel,
                                double* Zsel_d);

// ---- polynomial helpers (filters.cpp) ----------------------------------
int dampcf(int m, int damping, double* jac);
int chebxPltd(int m, const double* mu, int npts, const double* xi, double* yi);

// Operands of one filter application y = p((A - cc)/dd) x.  In FP32 the
// matrix is stored pre-shifted and scaled by a power of two,
// (A - cc I) * 2^-s with 2^s ~ dd, so that the shift is exact in FP64 before
// rounding and the epilogue needs no cc*x term (cc = 0, t scaled by 2^s).
struct FilterOp {
  evsl_dev::Csr A;
  double t1 = 0.0, t2 = 0.0, cc = 0.0;
  bool f32 = false;
};
int make_filter_op(double cc, double dd, bool f32, FilterOp* op);
// whether make_filter_op(.., f32, ..) on the current device matrix runs
// without allocating: always in FP64; in FP32 once the matrix's FP32 copy
// exists (a shifted copy is recomputed in place for another cc, dd)
bool filter_op_ready(bool f32);

// ---- filter application (lanczos.cpp) -----------------------------------
struct ChebWork {
  DevBuf bufs;  // 3 n-vectors
  DevBuf mu_d;  // filter coefficients on the device
  DevBuf mu0_d;
  DevBuf b32;   // FP32 work (4 n-vectors) for the mixed filter
  int deg_loaded = -1;
  const double* mu_src = nullptr;
};
// y = p(A) v (device vectors) with pol; FP64 or FP32 per ctx().filter_prec
int cheb_av(const polparams* pol, const double* v, double* y, ChebWork& w, int prec);

// ---- Lanczos building blocks (lanczos.cpp) ----------------------------------
// check_intv of INC/internal_header.h:194-217: -1 empty interval, -2 outside
// [lmin, lmax], 0 valid
int check_intv(const double* intv);
// device pointer for n doubles the caller holds on the host or the device
// (a host array is copied into tmp)
int as_device_vector(const double* p, size_t n, DevBuf& tmp, const double** out);
// Outputs of one slice of a LanczosGroup, with ChebLanNr's ownership: lam,
// res host (evsl_Free), W device (evsl_Free_device)
struct SliceResult {
  int nev = 0;
  double* lam = nullptr;
  double* W = nullptr;
  double* res = nullptr;
  int err = 0;
  int steps = 0;
  bool done = false;
  double t_done = 0.0;
};
// Slices of the current matrix (shared lmin, lmax) advancing through the
// filtered Lanczos recurrence together, one step at a time; a slice added
// between steps joins with the next one.  pol and v0 must stay valid until
// the slice is done (a mixed slice whose refinement stalls restarts from v0
// in FP64).
// the mixed form's rule for a slice's filter: FP32 unless its degree lies
// outside the mixed degree range or it is a capped filter run in FP64
bool mixed_slice_fp32(const polparams* pol);
// device bytes a slice's Lanczos basis takes (FP64, plus the FP32 copy of a
// mixed slice that filters in FP32)
size_t slice_basis_bytes(int n, int maxit, bool mixed, const polparams* pol);

class LanczosGroup {
 public:
  LanczosGroup(bool mixed, FILE* fstats);
  ~LanczosGroup();
  LanczosGroup(const LanczosGroup&) = delete;
  LanczosGroup& operator=(const LanczosGroup&) = delete;
  int add(const double* intv, int maxit, double tol, const double* v0, const polparams* pol,
          SliceResult* out);
  // one Lanczos step of every slice (its filter passes on this matrix)
  int step();
  int active() const;
  // The same step in phases, for a caller that filters the columns of
  // several groups together (grouped launches over their matrices).  Each
  // phase runs with this group's matrix as the context's operator; the
  // caller filters the columns a *_cols phase appends (out = p(A) in) before
  // the matching *_done phase, and accounts the filter time.
  struct FilterCol {
    const double* in;
    double* out;
    const polparams* pol;
    bool f32;
  };
  int start_cols(std::vector<FilterCol>& cols);  // start vectors of new slices
  int start_done();                              // normalise them, activate
  int step_cols(std::vector<FilterCol>& cols);   // the active slices' vectors
  int step_done();  // recurrence, re-orthogonalisation, tests, results
  // filters columns of this group's matrix (bound) with its own blocked
  // filter: tuned lane layouts, columns of lower degree dropped as the series
  // passes their degree
  int filter_cols(const std::vector<FilterCol>& cols);

 private:
  struct Impl;
  std::unique_ptr<Impl> p_;
};
// ---- grouped filter passes over several matrices (grouped.cpp) -------------
// The columns of one matrix at one precision: op from make_filter_op with
// that matrix bound (its pointers and sizes, op.A.n and op.A.nnz, stay valid
// once unbound), and the columns out = p(A) in (every col.f32 == op.f32;
// same cc, dd), each an op.A.n-vector.
struct GroupedCols {
  FilterOp op;
  std::vector<LanczosGroup::FilterCol> cols;
};
// device work of grouped_filter, reused from call to call
struct GroupedWork {
  struct PolKey {
    const polparams* pol;
    const double* mu;
    int deg;
    bool operator==(const PolKey& o) const { return pol == o.pol && mu == o.mu && deg == o.deg; }
    bool operator!=(const PolKey& o) const { return !(*this == o); }
  };
  struct Bufs {
    DevBuf X0, ring, Y, mus;
    std::vector<PolKey> mus_key;
    int mus_deg = -1;
  };
  std::vector<std::unique_ptr<Bufs>> bufs;
  DevBuf d_tab, d_owner;
  std::vector<int> owner_h;
};
// Filters every part's columns with grouped launches: one launch per
// Chebyshev degree and class of column blocks (precision, width, index
// type) covers the row blocks of all matrices whose series has not ended.
// Each column gets the bits of its one-matrix filter.
int grouped_filter(const std::vector<GroupedCols>& parts, GroupedWork& W);

// KPM moments and eigenvalue count from raw[k] = sum over the nvec vectors of
// v^T T_k(A~) v (k >= 1), as kpmdos finishes them
void kpm_moments(int Mdeg, int damping, int nvec, const double* intv, int n, const double* raw,
                 double* mu, double* ecnt);
// workspace of one filtered Lanczos run
struct LanWork {
  DevBuf V, gt_part, h, vrand;
  ChebWork cw;
  int nch = 0;
};
// CGS with the DGKS test, two GEMV passes at most (SRC/misc_la.c:221-261),
// against the k columns of Q (column-major, ld n); on entry old2 = ||v||^2;
// returns ||v|| in *nrm.  L.gt_part and L.h must hold k + 1 columns.
int cgs_dgks(int n, int k, const double* Q, double* v, double old2, double* nrm, LanWork& L);
// ||x|| (one device->host read)
int nrm2_dev(int n, const double* x, double* out);

}  // namespace evsl
