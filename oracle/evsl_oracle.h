/*
 * evsl_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of the EVSL 1.1.1 filtered-Lanczos / KPM hot path,
 * used to check the library's results by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py.  It is never linked into, loaded by or called from the
 * product library (evsl_amd/libevsl_amd.so).
 *
 * Each ora_* function restates the reference function named beside it
 * (paths relative to EVSL_1.1.1/).  Only the standard eigenproblem is
 * restated (the generalized branches are out of scope).
 */
#ifndef EVSL_ORACLE_H
#define EVSL_ORACLE_H

#include <stdio.h>
#include "../include/evsl.h" /* polparams / MVFunc layouts only */

#ifdef __cplusplus
extern "C" {
#endif

/* operator: CSR (SRC/spmat.c:257-298 dcsrmv) or a host callback */
void ora_set_csr(int n, const int *ia, const int *ja, const double *a);
void ora_set_matvec(int n, MVFunc func, void *data);
void ora_matvec(const double *x, double *y);

/* vectors and RNG (SRC/vect.c:14-20, 29-49, 82-92) */
void ora_rand_double(int n, double *v);
void ora_randn_double(int n, double *v);
void ora_linspace(double a, double b, int num, double *arr);

/* filters (SRC/chebpoly.c) */
int ora_dampcf(int m, int damping, double *jac);
int ora_chebxPltd(int m, const double *mu, int npts, const double *xi, double *yi);
void ora_set_pol_def(polparams *pol);
int ora_find_pol(const double *intv, polparams *pol);
void ora_free_pol(polparams *pol);
int ora_ChebAv(const polparams *pol, const double *v, double *y, double *w);

/* KPM DOS and slicing (SRC/spslice.c) */
int ora_kpmdos(int Mdeg, int damping, int nvec, const double *intv, double *mu,
               double *ecnt);
void ora_intChx(int Mdeg, const double *mu, int npts, const double *xi, double *yi);
int ora_spslicer(double *sli, const double *mu, int Mdeg, const double *intv,
                 int n_int, int npts);

/* dense kernels (SRC/misc_la.c) */
void ora_CGS_DGKS(int n, int k, int i_max, const double *Q, double *v, double *nrmv,
                  double *w);
int ora_tridiag_eig(int n, const double *diag, const double *sdiag, double *eval,
                    double *evec);
int ora_sym_eig(int n, const double *A, int lda, double *Q, int ldq, double *lam);

/* solvers */
int ora_ChebLanNr(const double *intv, int maxit, double tol, const double *vinit,
                  const polparams *pol, int *nevOut, double **lamo, double **Wo,
                  double **reso);
int ora_ChebLanTr(int lanm, int nev, const double *intv, int maxit, double tol,
                  const double *vinit, const polparams *pol, int *nev2,
                  double **vals, double **W, double **resW);
int ora_LanBounds(int msteps, const double *v, double *lmin, double *lmax);
int ora_ChebSI(int nev, const double *intv, int maxit, double tol, const double *vinit,
               const polparams *pol, int *nevo, double **lamo, double **Yo, double **reso);
int ora_LanTrbounds(int lanm, int maxit, double tol, const double *vinit, int bndtype,
                    double *lammin, double *lammax);
int ora_LanDos(int nvec, int msteps, int npts, double *xdos, double *ydos,
               double *neig, const double *intv);
void ora_simpson(const double *xi, double *yi, int npts);
void ora_spslicer2(const double *xi, double *yi, int n_int, int npts, double *sli);

/* misc */
void ora_free(void *p);
double ora_timer(void);
/* number of Lanczos steps taken by the last ora_ChebLanNr call */
int ora_last_lanczos_steps(void);

#ifdef __cplusplus
}
#endif
#endif
