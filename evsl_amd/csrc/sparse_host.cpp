// sparse_host.cpp -- host-side sparse/vector helpers of the C interface:
// COO -> CSR with column-sorted rows, CSR copies, host CSR matvec, sorting
// and linspace (SRC/spmat.c:27-322, SRC/vect.c:82-141 of EVSL 1.1.1), and the
// drivers' Matrix Market reader (TESTS/COMMON/io.c:115-217).
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/evsl.h"

namespace {

void csr_alloc(int nrow, int ncol, int nnz, csrMat* csr) {
  csr->owndata = 1;
  csr->nrows = nrow;
  csr->ncols = ncol;
  csr->nnz = nnz;
  csr->ia = (int*)std::malloc((size_t)(nrow + 1) * sizeof(int));
  csr->ja = (int*)std::malloc((size_t)std::max(nnz, 1) * sizeof(int));
  csr->a = (double*)std::malloc((size_t)std::max(nnz, 1) * sizeof(double));
}

// sort the column indices of every row (values follow); the reference does
// this by a double transposition (SRC/spmat.c:67-85), which yields the same
// order for rows without duplicate columns.
void sort_rows(csrMat* A) {
  std::vector<int> perm;
  std::vector<int> jt;
  std::vector<double> at;
  for (int i = 0; i < A->nrows; ++i) {
    const int b = A->ia[i], e = A->ia[i + 1];
    bool sorted = true;
    for (int k = b + 1; k < e; ++k)
      if (A->ja[k] < A->ja[k - 1]) {
        sorted = false;
        break;
      }
    if (sorted) continue;
    const int len = e - b;
    perm.resize(len);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int x, int y) { return A->ja[b + x] < A->ja[b + y]; });
    jt.resize(len);
    at.resize(len);
    for (int k = 0; k < len; ++k) {
      jt[k] = A->ja[b + perm[k]];
      at[k] = A->a[b + perm[k]];
    }
    std::memcpy(A->ja + b, jt.data(), len * sizeof(int));
    std::memcpy(A->a + b, at.data(), len * sizeof(double));
  }
}

}  // namespace

extern "C" {

// SRC/spmat.c:160-195
int cooMat_to_csrMat(int cooidx, cooMat* coo, csrMat* csr) {
  const int nnz = coo->nnz, nrows = coo->nrows;
  csr_alloc(coo->nrows, coo->ncols, nnz, csr);
  std::memset(csr->ia, 0, (size_t)(nrows + 1) * sizeof(int));
  for (int i = 0; i < nnz; i++) csr->ia[coo->ir[i] - cooidx + 1]++;
  for (int i = 0; i < nrows; i++) csr->ia[i + 1] += csr->ia[i];
  std::vector<int> next(csr->ia, csr->ia + nrows);
  for (int i = 0; i < nnz; i++) {
    const int row = coo->ir[i] - cooidx;
    const int k = next[row]++;
    csr->ja[k] = coo->jc[i] - cooidx;
    csr->a[k] = coo->vv[i];
  }
  sort_rows(csr);
  return 0;
}

void free_csr(csrMat* csr) {
  if (!csr->owndata) return;
  std::free(csr->ia);
  std::free(csr->ja);
  std::free(csr->a);
  csr->ia = csr->ja = nullptr;
  csr->a = nullptr;
}

void free_coo(cooMat* coo) {
  std::free(coo->ir);
  std::free(coo->jc);
  std::free(coo->vv);
  coo->ir = coo->jc = nullptr;
  coo->vv = nullptr;
}

// SRC/spmat.c:119-135
void csr_copy(csrMat* A, csrMat* B, int allocB) {
  const int nrows = A->nrows, nnz = A->ia[nrows];
  if (allocB) {
    csr_alloc(nrows, A->ncols, nnz, B);
  } else {
    B->nrows = nrows;
    B->ncols = A->ncols;
  }
  B->nnz = A->nnz;
  std::memcpy(B->ia, A->ia, (size_t)(nrows + 1) * sizeof(int));
  std::memcpy(B->ja, A->ja, (size_t)nnz * sizeof(int));
  std::memcpy(B->a, A->a, (size_t)nnz * sizeof(double));
}

// SRC/spmat.c:207-222
int arrays_copyto_csrMat(int nrow, int ncol, int* ia, int* ja, double* a, csrMat* A) {
  const int nnz = ia[nrow];
  csr_alloc(nrow, ncol, nnz, A);
  std::memcpy(A->ia, ia, (size_t)(nrow + 1) * sizeof(int));
  std::memcpy(A->ja, ja, (size_t)nnz * sizeof(int));
  std::memcpy(A->a, a, (size_t)nnz * sizeof(double));
  sort_rows(A);
  return 0;
}

// host y = A x for a host csrMat (SRC/spmat.c:309-322)
void matvec_csr(double* x, double* y, void* data) {
  const csrMat* A = (const csrMat*)data;
  for (int i = 0; i < A->nrows; ++i) {
    double r = 0.0;
    for (int k = A->ia[i]; k < A->ia[i + 1]; ++k) r += x[A->ja[k]] * A->a[k];
    y[i] = r;
  }
}

// ascending sort, optionally returning the original positions (SRC/vect.c:119-141)
void sort_double(int n, double* v, int* ind) {
  std::vector<int> p(n);
  std::iota(p.begin(), p.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return v[a] < v[b]; });
  std::vector<double> s(n);
  for (int i = 0; i < n; ++i) s[i] = v[p[i]];
  std::memcpy(v, s.data(), (size_t)n * sizeof(double));
  if (ind)
    for (int i = 0; i < n; ++i) ind[i] = p[i];
}

// SRC/vect.c:82-92
void linspace(double a, double b, int num, double* arr) {
  const double h = (num == 1 ? 0 : (b - a) / (num - 1));
  arr[0] = a;
  arr[num - 1] = b;
  for (int i = 1; i < num - 1; i++) arr[i] = a + i * h;
}

}  // extern "C"

// Matrix Market coordinate reader of the reference drivers
// (TESTS/COMMON/io.c:115-217): real or integer, general or symmetric; for a
// symmetric file the stored triangle is followed by the mirrored
// off-diagonal entries, in file order; indices are shifted by idxout - idxin.
// Returns 0, 1 for an unsupported or malformed file (as the reference), or
// EVSL_ERR_ARG when the file cannot be opened (the reference exits).
extern "C" int evsl_amd_read_coo_mm(const char* path, int idxin, int idxout, cooMat* A) {
  if (!path || !A) return EVSL_ERR_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return EVSL_ERR_ARG;
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  char line[4096];
  if (!std::fgets(line, sizeof line, f)) return 1;
  // banner: %%MatrixMarket matrix coordinate <real|integer> <general|symmetric>
  char tag[64] = {0}, obj[64] = {0}, fmt[64] = {0}, field[64] = {0}, sym[64] = {0};
  if (std::sscanf(line, "%63s %63s %63s %63s %63s", tag, obj, fmt, field, sym) != 5) return 1;
  auto lower = [](char* s) {
    for (; *s; ++s) *s = (char)std::tolower((unsigned char)*s);
  };
  lower(obj);
  lower(fmt);
  lower(field);
  lower(sym);
  if (std::strcmp(tag, "%%MatrixMarket") != 0 || std::strcmp(obj, "matrix") != 0) return 1;
  if (std::strcmp(fmt, "coordinate") != 0) return 1;
  if (std::strcmp(field, "real") != 0 && std::strcmp(field, "integer") != 0) return 1;
  const bool symmetric = std::strcmp(sym, "symmetric") == 0;
  if (!symmetric && std::strcmp(sym, "general") != 0) return 1;
  long nrow = -1, ncol = -1, nnz = -1;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '%') continue;
    if (std::sscanf(line, "%ld %ld %ld", &nrow, &ncol, &nnz) == 3) break;
    nrow = -1;
  }
  if (nrow < 0 || ncol < 0 || nnz < 0) return 1;
  if (nrow != ncol) return 1;
  const long cap = symmetric ? 2 * nnz : nnz;
  std::vector<int> ir((size_t)cap), jc((size_t)cap);
  std::vector<double> vv((size_t)cap);
  for (long k = 0; k < nnz; ++k) {
    if (!std::fgets(line, sizeof line, f)) return 1;
    char* p = line;
    char* e;
    const long r = std::strtol(p, &e, 10);
    if (e == p) return 1;
    p = e;
    const long cidx = std::strtol(p, &e, 10);
    if (e == p) return 1;
    p = e;
    const double v = std::strtod(p, &e);
    if (e == p) return 1;
    ir[k] = (int)r;
    jc[k] = (int)cidx;
    vv[k] = v;
  }
  long j = nnz;
  if (symmetric)
    for (long k = 0; k < nnz; ++k)
      if (ir[k] != jc[k]) {
        ir[j] = jc[k];
        jc[j] = ir[k];
        vv[j] = vv[k];
        ++j;
      }
  const int off = idxout - idxin;
  for (long k = 0; k < j; ++k) {
    ir[k] += off;
    jc[k] += off;
  }
  A->nrows = (int)nrow;
  A->ncols = (int)ncol;
  A->nnz = (int)j;
  A->ir = (int*)std::malloc((size_t)std::max(j, 1L) * sizeof(int));
  A->jc = (int*)std::malloc((size_t)std::max(j, 1L) * sizeof(int));
  A->vv = (double*)std::malloc((size_t)std::max(j, 1L) * sizeof(double));
  if (!A->ir || !A->jc || !A->vv) return EVSL_ERR_ALLOC;
  std::memcpy(A->ir, ir.data(), (size_t)j * sizeof(int));
  std::memcpy(A->jc, jc.data(), (size_t)j * sizeof(int));
  std::memcpy(A->vv, vv.data(), (size_t)j * sizeof(double));
  return 0;
}
