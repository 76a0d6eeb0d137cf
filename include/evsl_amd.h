/*
 * evsl_amd.h -- extensions of the drop-in interface (include/evsl.h) that
 * have no counterpart in EVSL 1.1.1.  They expose the MI355X-specific forms
 * of the same hot path: filter precision, seeded device random vectors for
 * KPM, all slices of one matrix solved as one batched filtered Lanczos, and
 * per-call device statistics.
 */
#ifndef EVSL_AMD_EXT_H
#define EVSL_AMD_EXT_H

#include <stddef.h>
#include "evsl.h"

#define EVSL_PREC_FP64 0  /* filter in FP64 (the reference's arithmetic)      */
#define EVSL_PREC_MIXED 1 /* FP32 filter, FP64 Lanczos + FP64 Rayleigh-Ritz   */

#define EVSL_VEC_RAND_DOUBLE 0 /* libc rand() U[-1,1], the reference stream */
#define EVSL_VEC_RADEMACHER 1  /* seeded device counter RNG, +-1            */
#define EVSL_VEC_GAUSSIAN 2    /* seeded device counter RNG, N(0,1)         */

/* Device-side statistics accumulated since the last reset.  Times are in
 * seconds measured with HIP events on the library's stream. */
typedef struct _evslAmdStats {
  double t_filter;       /* Chebyshev filter (ChebAv / KPM recurrences)   */
  double t_reorth;       /* CGS/DGKS re-orthogonalization                */
  double t_eig;          /* tridiagonal eigensolves (host + rotations)    */
  double t_ritz;         /* Ritz vectors, residuals, Rayleigh-Ritz        */
  size_t n_filter_steps; /* fused SpMM degree steps launched              */
  size_t n_filter_cols;  /* sum over those launches of the block width    */
  size_t n_lanczos;      /* Lanczos steps                                 */
  size_t n_reorth_pass;  /* Gram-Schmidt passes                           */
  double filter_bytes;   /* algorithmic bytes of the Lanczos/KPM filter   */
  double reorth_bytes;   /* algorithmic bytes of the Gram-Schmidt passes  */
  size_t n_refine_cols;  /* Ritz-vector columns filtered by refinement     */
  size_t n_refine_extra; /* refinement passes beyond the requested ones   */
  double refine_bytes;   /* algorithmic bytes of the refinement's filter  */
  size_t n_refine_degcols; /* sum over refinement passes of columns x degree */
  size_t n_fp64_slices;  /* MIXED: slices filtered in FP64 from the start  */
                         /* (degree outside the mixed range, or a capped   */
                         /* filter)                                        */
  size_t n_fp64_reruns;  /* MIXED: slices redone in FP64 after a stalled   */
                         /* refinement                                     */
  size_t n_refine_fp64_cols; /* refinement columns filtered in FP64 (their */
                             /* residual too large for the FP32 recurrence) */
  size_t n_refine_retries; /* MIXED: stalled refinements run again at half  */
                           /* the guard threshold                           */
} evslAmdStats;

#ifdef __cplusplus
extern "C" {
#endif

/* filter precision for ChebLanNr / ChebLanTr / ChebLanNr_batch */
int evsl_amd_set_filter_precision(int prec);
int evsl_amd_get_filter_precision(void);
/* number of filtered Rayleigh-Ritz refinement passes in MIXED mode (>= 1);
 * more run while a Ritz value in [a, b] is near convergence (residual below
 * 1000 tol) and the largest such residual falls by 10% per pass, at most 8 */
int evsl_amd_set_refine_passes(int passes);
/* guard threshold of the refinement: Ritz values of p(A) down to
 * guard * bar enter the FP64 Rayleigh-Ritz space (0 < guard <= 1,
 * default 0.8; a stalled refinement is retried once at half of it) */
int evsl_amd_set_refine_guard(double guard);
/* the first refinement pass filters with find_pol's polynomial for the
 * slice's interval widened by `widen` slice widths on each side (clipped to
 * [lmin, lmax]): a lower degree, enough to damp the FP32 filter's rounding
 * noise far from the interval, which is what the residuals weigh most;
 * further passes, while a Ritz value in [a, b] is near convergence, use the
 * slice's own filter.  0 makes every pass use the slice's filter (default 4) */
int evsl_amd_set_refine_widen(double widen);
/* MIXED mode: only slices whose filter degree lies in [min_deg, max_deg]
 * filter in FP32 (default [32, 20000]); the others run in FP64 -- a
 * low-degree filter carries few of the bytes and separates the refinement's
 * guard set too weakly.  A filter that did not meet find_pol's thresholds
 * (deg = max_deg - 1) also runs in FP64: it separates its interval weakly
 * and its slice tends to end at maxit, where the FP32-filtered basis leaves
 * pairs just above tol (both counted in evslAmdStats.n_fp64_slices).  A
 * mixed slice whose refinement stalls above tol after its extra passes is
 * refined once more with half the guard threshold, and if that stalls too,
 * redone from its start vector in FP64 (evslAmdStats.n_fp64_reruns). */
int evsl_amd_set_mixed_degree_range(int min_deg, int max_deg);
/* MIXED mode: whether a capped filter (deg = max_deg - 1) runs in FP64 (1,
 * the default) or, like the others, in FP32 (0) */
int evsl_amd_set_mixed_capped_fp64(int on);
/* MIXED mode: whether the Gram-Schmidt passes of a slice that filters in
 * FP32 read an FP32 copy of its Lanczos basis (1, the default) or the FP64
 * basis (0).  The copy is written beside each new basis vector (4 bytes per
 * entry more memory); the projections V^T w and w - V h keep FP64 products
 * and sums, so a pass reads half the bytes.  The Ritz vectors and the
 * refinement use the FP64 basis either way. */
int evsl_amd_set_fp32_basis(int on);

int evsl_amd_get_stats(evslAmdStats *s);
void evsl_amd_reset_stats(void);
/* Message of the first failure in the most recent device call ("" if it
 * succeeded; host-only calls leave it unchanged). */
const char *evsl_amd_last_error(void);
/* size in bytes of evslAmdStats (ABI check for bindings) */
size_t evsl_amd_stats_size(void);

/* kpmdos with a choice of random vectors (kind: EVSL_VEC_*); with
 * EVSL_VEC_RAND_DOUBLE it is kpmdos exactly. */
int kpmdos_ex(int Mdeg, int damping, int nvec, double *intv, int vec_kind,
              unsigned long long seed, double *mu, double *ecnt);

/* KPM densities of several matrices in one batched run (requests of uneven
 * size advance together: one grouped launch per Chebyshev degree over the
 * row blocks of every matrix).  Request r is kpmdos (SRC/spslice.c:29-148)
 * on mats[r] (host CSR; the same pointer may appear in several requests)
 * with interval intvs[4r..4r+3], Mdeg, damping and nvec shared, and the
 * seeded device vectors of kpmdos_ex (vec_kind EVSL_VEC_RADEMACHER or
 * EVSL_VEC_GAUSSIAN, seed seeds[r], or r when seeds is NULL): in
 * EVSL_PREC_FP64 its moments equal kpmdos_ex on that matrix bit for bit;
 * EVSL_PREC_MIXED runs the recurrence in FP32 on (A - ctr I) 2^-s with the
 * moment dots accumulated in FP64.  Outputs mu[r*(Mdeg+1) + k], ecnt[r].
 * Independent of SetAMatrix.  Returns 0, -1 for an invalid interval,
 * EVSL_ERR_ARG, EVSL_ERR_UNSUPP (vec_kind) or a device error. */
int kpmdos_batch(int nreq, csrMat *const *mats, const double *intvs, int Mdeg, int damping,
                 int nvec, int vec_kind, const unsigned long long *seeds, int prec, double *mu,
                 double *ecnt);

/* All slices of one matrix as one batched filtered Lanczos run.  Slice s is
 * [intvs[4s], intvs[4s+1]] with global bounds intvs[4s+2..4s+3] (identical
 * for every slice) and filter pols[s] (from find_pol).  Each slice follows
 * ChebLanNr (same recurrence, stopping test, acceptance); the active slices'
 * Lanczos vectors advance together through one blocked SpMM per degree.
 * vinit is shared by all slices (host or device).  Outputs are per slice,
 * with ChebLanNr's ownership: lamo[s], reso[s] host (evsl_Free), Wo[s]
 * device (evsl_Free_device).  With EVSL_PREC_MIXED the filter runs in FP32
 * and the accepted pairs come from an FP64 filtered Rayleigh-Ritz pass. */
int ChebLanNr_batch(int nslices, double *intvs, int maxit, double tol, double *vinit,
                    polparams *pols, int *nevOut, double **lamo, double **Wo, double **reso,
                    FILE *fstats);

/* Y(:, j) = p_j(A) X(:, j) for the ncols columns of the n x ncols
 * column-major X with the blocked filter (the block form of the reference's
 * internal ChebAv, SRC/chebpoly.c:519-598): prec EVSL_PREC_FP64, or
 * EVSL_PREC_MIXED for the FP32 filter on the shifted matrix.  pols[j] (from
 * find_pol) may differ per column but must share cc and dd.  X and Y may be
 * host or device memory. */
int evsl_amd_cheb_av_cols(int ncols, const polparams *pols, const double *X, double *Y,
                          int prec);

/* evsl_amd_cheb_av_cols for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It allocates and frees no
 * device memory: its work buffers live in a static 64 MiB device array that
 * each GPU holds from the library's load until the process ends.  X and Y
 * (n x ncols, column-major) must be device memory; the device matrix must be
 * set (SetAMatrix) and, for EVSL_PREC_MIXED, its FP32 copy made by an
 * earlier mixed-precision call on it.  Columns run in blocks that fit the
 * scratch; one column needs n x (4 x s + 8) + (deg + 1) x 8 bytes (s = 8 in
 * FP64, 4 in MIXED), so n up to about 1.6 million rows in FP64 and 2.7
 * million in MIXED; beyond that EVSL_ERR_ARG.  The result equals
 * evsl_amd_cheb_av_cols's bit for bit.  Calls that overlap on one GPU share
 * the scratch and race; each call leaves it holding its own last block. */
int evsl_amd_cheb_av_cols_noalloc(int ncols, const polparams *pols, const double *X,
                                  double *Y, int prec);

/* A request loop over device-resident matrices (no counterpart in the
 * reference; its batched slicing and DOS forms served as requests arrive).
 * Matrices registered with evsl_amd_server_add_matrix stay in HBM.  Slice
 * requests (ChebLanNr on [a, b] of a registered matrix with its [lmin, lmax],
 * Krylov dimension maxit, filter from find_pol with pol_in's inputs or the
 * drivers' defaults when NULL, start vector vinit host or device) and DOS
 * requests (kpmdos_ex with seeded device vectors; count over [a, b]) are
 * queued by submit; each evsl_amd_server_step admits queued requests in
 * arrival order while their projected memory (n x (maxit+2) x 8 bytes for a
 * slice) fits mem_budget, advances every running slice by one filtered
 * Lanczos step -- the filter passes of the small matrices' slices (launches
 * of at most 32 MB) in grouped launches (one launch per degree covers the
 * row blocks of every such matrix whose filters have not reached it), each
 * larger matrix with its own filter; evsl_amd_set_grouped_filter(0) filters
 * every matrix on its own -- and every running DOS request by up to 64
 * Chebyshev degrees in grouped launches.  A request's result does not depend
 * on the other requests sharing its steps (bit for bit).  step returns the
 * number of requests not yet done (0 when idle) or an error.  status returns
 * 0 queued, 1 running, 2 done, with evsl_timer() stamps.  Results: slice ->
 * lam, res host (evsl_Free), Y device (evsl_Free_device), ownership passes
 * to the caller; DOS -> mu[Mdeg+1], ecnt.  Single-threaded, like the context. */
typedef struct evslAmdServer evslAmdServer;
evslAmdServer *evsl_amd_server_create(size_t mem_budget, int prec);
void evsl_amd_server_destroy(evslAmdServer *s);
int evsl_amd_server_add_matrix(evslAmdServer *s, csrMat *A, double lmin, double lmax);
int evsl_amd_server_submit_slice(evslAmdServer *s, int mat, double a, double b, int maxit,
                                 double tol, const polparams *pol_in, const double *vinit);
int evsl_amd_server_submit_dos(evslAmdServer *s, int mat, double a, double b, int Mdeg,
                               int damping, int nvec, int vec_kind, unsigned long long seed);
int evsl_amd_server_step(evslAmdServer *s);
int evsl_amd_server_status(evslAmdServer *s, int req, double *t_submit, double *t_start,
                           double *t_done);
int evsl_amd_server_slice_result(evslAmdServer *s, int req, int *nev, double **lam,
                                 double **Y, double **res);
int evsl_amd_server_dos_result(evslAmdServer *s, int req, double *mu, double *ecnt);
int evsl_amd_server_stats(evslAmdServer *s, int *queued, int *running, size_t *mem_in_use,
                          size_t *mem_peak);
/* the request loop's slice filters: grouped launches over all matrices (1,
 * the default) or one matrix at a time (0); results are the same bits */
int evsl_amd_set_grouped_filter(int on);

/* Matrix Market coordinate reader of the reference drivers
 * (TESTS/COMMON/io.c:115-217, read_coo_MM): real/integer, general/symmetric
 * (the stored triangle, then its mirrored off-diagonal entries); indices
 * shifted by idxout - idxin; arrays malloc'ed (free_coo).  Returns 0, 1 for
 * an unsupported or malformed file, EVSL_ERR_ARG if it cannot be opened. */
int evsl_amd_read_coo_mm(const char *path, int idxin, int idxout, cooMat *A);

/* Measurement hook: time nsteps launches of the fused filter kernel on the
 * current matrix with HIP events: block width B in {1..8,16,32}, mode 1 a
 * ChebAv step of the filter for spectrum centre cc and half-width dd (in
 * EVSL_PREC_MIXED on the shifted FP32 matrix), mode 2 a KPM step (FP64, B a
 * power of two), mode 3 a step of the refinement's residual recurrence.
 * Returns the mean launch time and the algorithmic bytes one launch moves. */
int evsl_amd_bench_filter(int B, int prec, int mode, int nsteps, double cc, double dd,
                          double *ms_per_step, double *bytes_per_step);

/* Measurement hook: the dense FP64 kernels of the solvers on random data
 * (n rows), reps times between HIP events: kind 0 one Gram-Schmidt pass
 * against k basis vectors (V^T w, then w -= V h with the norm partials),
 * kind 1 the Ritz-vector GEMM V (n x k) Z (k x m), kind 2 the Gram product
 * U^T U of an n x m block, kind 3 the kind 0 pass against an FP32 copy of V
 * (the mixed form's passes).  Returns the mean time and the algorithmic
 * bytes and flops of one repetition. */
int evsl_amd_bench_dense(int kind, int n, int k, int m, int reps, double *ms_per_rep,
                         double *bytes, double *flops);

/* Lane layout of the fused filter kernel (no counterpart in the reference;
 * its CSR matvec has no layout).  A layout is 10*V + R: V adjacent columns
 * per lane and R rows per lane group.  The ChebAv (mode 1) and residual
 * (mode 3) steps of 2, 4, 8, 16 and 32 columns choose theirs per matrix: each
 * device matrix carries a tuner that times the candidate layouts on the
 * solver's own first launches of a (B, prec, mode) key, between HIP events it
 * polls without waiting, and then keeps the fastest.  Every layout sums each
 * row in the same order, so tuning never changes a result.
 *   evsl_amd_spmm_layout: the layout the key launches now on the current
 *     matrix (the fixed policy's while the tuner is still measuring, and for
 *     KPM steps and other widths); *decided (may be NULL) 1 once fixed;
 *     -1 for a bad mode.
 *   evsl_amd_set_spmm_layout: the key's policy on the current matrix --
 *     layout > 0 forces that candidate, 0 the fixed policy, -1 tunes (the
 *     default); resets what the key has measured.  EVSL_ERR_ARG for another
 *     width or mode or a layout outside the candidates, EVSL_ERR_NOMAT
 *     without a device matrix.
 *   evsl_amd_set_spmm_tuning: 0 runs every matrix on the fixed policy, 1
 *     (the default) tunes. */
int evsl_amd_spmm_layout(int B, int prec, int mode, int *decided);
int evsl_amd_set_spmm_layout(int B, int prec, int mode, int layout);
int evsl_amd_set_spmm_tuning(int on);

/* Columns per lane of the fused filter kernel evsl_amd_bench_filter times
 * for (B, prec, mode) on the current matrix (the tuned layout's V, after the
 * tuner has settled; bench_filter lets it settle before timing): its
 * template instance is
 * k_cheb_spmm<float|double, B, V, mode, true, IX, NT> (IX from
 * evsl_amd_spmm_index_mode, NT from evsl_amd_spmm_streaming); -1 for an
 * unsupported B. */
int evsl_amd_spmm_vec(int B, int prec, int mode);
/* 1 if the fused filter kernel for (B, prec, mode) on the current matrix
 * runs its non-temporal instance (k_cheb_spmm<..., IX, true>: one launch's
 * working set exceeds 320 MB), 0 otherwise or without a matrix. */
int evsl_amd_spmm_streaming(int B, int prec, int mode);
/* Bytes per nonzero of the column index of the current matrix: 1 with
 * pattern codes (below), 2 when every |ja - row| fits in 16 bits, 4
 * otherwise; 0 without a matrix. */
int evsl_amd_spmm_index_bytes(void);
/* Index mode of the current matrix's fused filter launches for (B, prec,
 * mode) (the IX of the kernel instance): 0 int32 columns, 1 16-bit column
 * offsets, 2 and 3 pattern codes whose table offsets fit 16 bits (2) or
 * not (3) -- pattern codes are read by launches whose vector rows hold at
 * most 8 bytes (one or two FP32 columns, one FP64 column); -1 without a
 * matrix. */
int evsl_amd_spmm_index_mode(int B, int prec, int mode);
/* Pattern codes (no reference counterpart): a matrix whose off-diagonal
 * entries take at most 255 distinct (column offset, value) pairs -- stencil
 * and lattice operators -- is stored for the filter kernels as one byte
 * per nonzero naming its pair (the diagonal by a code of its own, its value
 * from a per-row array) beside the CSR arrays; the kernels then stream 1
 * byte per nonzero instead of an index and a value, with the same products
 * in the same order, so results do not change.
 *   evsl_amd_set_pattern_codes: 1 (the default) uses them for matrices set
 *     from now on, 0 does not.
 *   evsl_amd_pattern_codes: table entries in use for the current matrix,
 *     0 when it is not pattern-coded.
 * Launches whose vector rows hold more than 8 bytes read the index and
 * value arrays instead (evsl_amd_spmm_index_mode): there the table lookup
 * ahead of the gathers costs more than the bytes it saves. */
int evsl_amd_set_pattern_codes(int on);
int evsl_amd_pattern_codes(void);

/* Measurement hook: mean time of a device copy of `bytes` (16-byte
 * accesses) over reps launches -- the achievable-bandwidth reference
 * (2 * bytes move per copy). */
int evsl_amd_bench_copy(size_t bytes, int reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
