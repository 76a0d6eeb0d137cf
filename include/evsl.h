/*
 * evsl.h -- drop-in C interface of the MI355X filtered-Lanczos / KPM library.
 *
 * A program written against EVSL 1.1.1 (INC/evsl.h, INC/struct.h) compiles
 * against this header and links against libevsl_amd.so unchanged.  Struct
 * layouts are the CPU layouts of INC/struct.h:14-98 (no CUDA descriptor
 * fields); prototypes keep the argument lists and meanings of INC/evsl.h.
 * Every "device" pointer in this interface is HIP device memory on the
 * current device (allocated by evsl_Malloc_device / hipMalloc).
 *
 * Error behaviour differs from the reference in one deliberate way: no entry
 * point calls exit() or assert(); failures are returned as codes, and device
 * failures map to EVSL_ERR_DEVICE (< -100).
 */
#ifndef EVSL_AMD_EVSL_H
#define EVSL_AMD_EVSL_H

#include <stdio.h>
#include <stddef.h>
#include <stdint.h>

#define EVSL_PI 3.14159265358979323846
/* tolerance multiplier for "a >= b" style comparisons (INC/evsl.h:28) */
#define EVSL_DBL_EPS_MULT 10
#define evsl_max(a, b) ((a) > (b) ? (a) : (b))
#define evsl_min(a, b) ((a) < (b) ? (a) : (b))

/* error codes added by this library (the reference asserts instead) */
#define EVSL_ERR_DEVICE   (-101) /* a HIP call failed                      */
#define EVSL_ERR_NOMAT    (-102) /* no operator set (SetAMatrix/SetAMatvec) */
#define EVSL_ERR_ALLOC    (-103) /* device allocation failed               */
#define EVSL_ERR_ARG      (-104) /* invalid argument                       */
#define EVSL_ERR_EIG      (-105) /* tridiagonal eigensolver did not converge */
#define EVSL_ERR_UNSUPP   (-106) /* generalized problem on the device path  */

/* ---- data types: layouts identical to INC/struct.h ---------------------- */

/* COO, 0-based (INC/struct.h:14-21) */
typedef struct _cooMat {
  int nrows, ncols, nnz, *ir, *jc;
  double *vv;
} cooMat;

/* CSR, 0-based, nnz == ia[nrows] (INC/struct.h:29-42, CPU layout) */
typedef struct _csrMat {
  int owndata, nrows, ncols, nnz, *ia, *ja;
  double *a;
} csrMat;

/* polynomial filter parameters (INC/struct.h:68-98) */
typedef struct _polparams {
  int max_deg;        /* in: max degree                                   */
  int min_deg;        /* in: min degree                                   */
  int damping;        /* in: 0 none, 1 Jackson, 2 Lanczos sigma            */
  double thresh_ext;  /* in: acceptance threshold, end intervals           */
  double thresh_int;  /* in: acceptance threshold, interior intervals      */
  double intvtol;     /* in: cut-off deciding end vs middle interval       */
  int type;           /* out: 0 middle, 1 left end, 2 right end            */
  double *mu;         /* out: Chebyshev coefficients mu[0..deg] (host)     */
  double cc;          /* out: centre of [lmin,lmax]                        */
  double dd;          /* out: half-width of [lmin,lmax]                    */
  double gam;         /* out: centre of the expanded delta function        */
  double bar;         /* out: p(theta) >= bar marks a wanted Ritz value    */
  int deg;            /* in/out: forced degree if > 0, else found degree   */
} polparams;

typedef void (*MVFunc)(double *x, double *y, void *data);
typedef void (*SolFuncR)(double *b, double *x, void *data);
typedef void (*SolFuncC)(int n, double *br, double *bz, double *xr, double *xz, void *data);

/* timing statistics (INC/struct.h:223-253); same field order */
typedef struct _evslstat {
  double t_setBsv, t_setASigBsv, t_iter;
  double t_polAv, t_reorth, t_siorth, t_ritz, t_eig;
  double t_mvA, t_mvB, t_svB, t_svLT, t_svASigB, t_blas, t_ratAv;
  size_t n_mvA, n_mvB, n_svB, n_svLT, n_svASigB, n_polAv, n_ratAv;
  size_t alloced, alloced_total, alloced_max;
} evslStat;

extern evslStat evslstat;

/* ---- memory helpers (INC/evsl.h:48-135, 166-177) ------------------------ */
#define evsl_Malloc(count, type) ((type *)_evsl_Malloc((size_t)(sizeof(type) * (count))))
#define evsl_Calloc(count, type) ((type *)_evsl_Calloc((size_t)(count), (size_t)sizeof(type)))
#define evsl_Realloc(ptr, count, type) \
  ((type *)_evsl_Realloc((void *)(ptr), (size_t)(sizeof(type) * (count))))
#define evsl_Free(ptr) (_evsl_Free((void *)(ptr)), (ptr) = NULL)
#define evsl_Malloc_device(count, type) \
  ((type *)_evsl_Malloc_device((size_t)(sizeof(type) * (count))))
#define evsl_Calloc_device(count, type) \
  ((type *)_evsl_Calloc_device((size_t)(count), (size_t)sizeof(type)))
#define evsl_Realloc_device(ptr, old_count, old_type, new_count, new_type) \
  ((new_type *)_evsl_Realloc_device((void *)(ptr), (size_t)(sizeof(old_type) * (old_count)), \
                                    (size_t)(sizeof(new_type) * (new_count))))
#define evsl_Free_device(ptr) (_evsl_Free_device((void *)(ptr)), (ptr) = NULL)

#ifdef __cplusplus
extern "C" {
#endif

/* memory (SRC/evsl_memory.c:9-149) */
void *_evsl_Malloc(size_t nbytes);
void *_evsl_Calloc(size_t count, size_t nbytes);
void *_evsl_Realloc(void *ptr, size_t nbytes);
void _evsl_Free(void *ptr);
void *_evsl_Malloc_device(size_t nbytes);
void *_evsl_Calloc_device(size_t count, size_t nbytes);
void *_evsl_Realloc_device(void *old_ptr, size_t old_nbytes, size_t new_nbytes);
void _evsl_Free_device(void *ptr);
void evsl_memcpy_device_to_host(void *dst, void *src, size_t nbytes);
void evsl_memcpy_device_to_device(void *dst, void *src, size_t nbytes);
void evsl_memcpy_host_to_device(void *dst, void *src, size_t nbytes);
void evsl_memset_device(void *ptr, int value, size_t nbytes);

/* lifecycle and operator (INC/evsl.h:240-262; SRC/evsl.c) */
int EVSLStart(void);
int EVSLFinish(void);
int SetAMatrix(csrMat *A);             /* host CSR; uploaded to the device   */
int SetAMatvec(int n, MVFunc func, void *data); /* func gets device x, y   */
int SetBMatrix(csrMat *B);             /* EVSL_ERR_UNSUPP (generalized)      */
int SetBMatvec(int n, MVFunc func, void *data);
int SetBSol(SolFuncR func, void *data);
int SetLTSol(SolFuncR func, void *data);
int SetStdEig(void);
int SetGenEig(void);                   /* EVSL_ERR_UNSUPP                    */
void SetDiagScal(double *ds);
/* device-resident CSR, as the reference's CUDA build (INC/evsl.h:228-229,262) */
void evsl_create_csr_gpu(csrMat *Acpu, csrMat *Agpu);
void evsl_free_csr_gpu(csrMat *csr);
int SetAMatrix_device_csr(csrMat *A);

/* filters (INC/evsl.h:148-150; SRC/chebpoly.c) */
void set_pol_def(polparams *pol);
int find_pol(double *intv, polparams *pol);
void free_pol(polparams *pol);

/* slicing and DOS (INC/evsl.h:187, 271-275) */
int kpmdos(int Mdeg, int damping, int nvec, double *intv, double *mu, double *ecnt);
int spslicer(double *sli, double *mu, int Mdeg, double *intv, int n_int, int npts);
int LanDos(const int nvec, int msteps, int npts, double *xdos, double *ydos,
           double *neig, const double *const intv);
void spslicer2(double *xi, double *yi, int n_int, int npts, double *sli);
/* exact smoothed DOS of known eigenvalues (INC/evsl.h:180, SRC/exDOS.c:17-69;
 * the reference LanDos driver's comparison curve), host only */
int exDOS(double *vals, int n, int npts, double *x, double *y, double *intv);

/* solvers (INC/evsl.h:142-145, 183) */
int ChebLanNr(double *intv, int maxit, double tol, double *vinit, polparams *pol,
              int *nevOut, double **lamo, double **Wo, double **reso, FILE *fstats);
int ChebLanTr(int lanm, int nev, double *intv, int maxit, double tol, double *vinit,
              polparams *pol, int *nev2, double **vals, double **W, double **resW,
              FILE *fstats);
int LanBounds(int msteps, double *v, double *lmin, double *lmax);
/* SRC/chebsi.c:39-336 (INC/evsl.h:154): filtered subspace iteration on a
 * block of nev vectors vinit (n x nev, host or device) */
int ChebSI(int nev, double *intv, int maxit, double tol, double *vinit, polparams *pol,
           int *nevo, double **lamo, double **Yo, double **reso, FILE *fstats);
/* SRC/lanTrbounds.c:35-421 (INC/evsl.h:193): thick-restart Lanczos bounds;
 * bndtype <= 1 simple, otherwise Kato-Temple; vinit host or device */
int LanTrbounds(int lanm, int maxit, double tol, double *vinit, int bndtype,
                double *lammin, double *lammax, FILE *fstats);

/* sparse helpers (INC/evsl.h:211-237) */
int cooMat_to_csrMat(int cooidx, cooMat *coo, csrMat *csr);
void free_csr(csrMat *csr);
void free_coo(cooMat *coo);
void csr_copy(csrMat *A, csrMat *B, int allocB);
int arrays_copyto_csrMat(int nrow, int ncol, int *ia, int *ja, double *a, csrMat *A);
void matvec_csr(double *x, double *y, void *data); /* host CSR, host x,y */

/* vectors, RNG, timing, stats (INC/evsl.h:278-305) */
void rand_double(int n, double *v);
void rand_double_device(int n, double *v);
void randn_double(int n, double *v);
void randn_double_device(int n, double *v);
void sort_double(int n, double *v, int *ind);
void linspace(double a, double b, int num, double *arr);
double evsl_timer(void);
void StatsPrint(FILE *fstats);
void StatsReset(void);

#ifdef __cplusplus
}
#endif

#endif /* EVSL_AMD_EVSL_H */
