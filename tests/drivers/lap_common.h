/* Shared helpers of the C drivers in this folder: the 5/7-point Laplacian in
 * COO (the matrix of TESTS/COMMON/lapl.c:13-58: diagonal 4 in 2-D, 6 in 3-D,
 * -1 to each grid neighbour) and its closed-form eigenvalues in [a, b]
 * (4 sin^2 terms per dimension, as TESTS/COMMON/lapl.c:79-109), sorted. */
#ifndef LAP_COMMON_H
#define LAP_COMMON_H
#include <math.h>
#include <stdlib.h>

#include "evsl.h"

static int lap_coo(int nx, int ny, int nz, cooMat *A) {
  const int n = nx * ny * nz;
  const int cap = 7 * n;
  A->nrows = A->ncols = n;
  A->ir = (int *)malloc(cap * sizeof(int));
  A->jc = (int *)malloc(cap * sizeof(int));
  A->vv = (double *)malloc(cap * sizeof(double));
  if (!A->ir || !A->jc || !A->vv) return 1;
  int nnz = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int r = i + nx * (j + ny * k);
        const int nb[6] = {i > 0 ? r - 1 : -1, i < nx - 1 ? r + 1 : -1,
                           j > 0 ? r - nx : -1, j < ny - 1 ? r + nx : -1,
                           k > 0 ? r - nx * ny : -1, k < nz - 1 ? r + nx * ny : -1};
        A->ir[nnz] = r;
        A->jc[nnz] = r;
        A->vv[nnz++] = nz > 1 ? 6.0 : 4.0;
        for (int q = 0; q < 6; q++)
          if (nb[q] >= 0) {
            A->ir[nnz] = r;
            A->jc[nnz] = nb[q];
            A->vv[nnz++] = -1.0;
          }
      }
  A->nnz = nnz;
  return 0;
}

static int cmp_double(const void *x, const void *y) {
  const double a = *(const double *)x, b = *(const double *)y;
  return a < b ? -1 : a > b;
}

/* closed-form eigenvalues in [a - 10 eps, b + 10 eps], sorted; malloc'ed */
static double *lap_eigs(int nx, int ny, int nz, double a, double b, int *count) {
  const double eps = EVSL_DBL_EPS_MULT * 2.220446049250313e-16;
  double *ev = (double *)malloc((size_t)nx * ny * nz * sizeof(double));
  int m = 0;
  for (int k = 1; k <= nz; k++) {
    const double tz = nz > 1 ? sin(k * 0.5 * EVSL_PI / (nz + 1.0)) : 0.0;
    for (int j = 1; j <= ny; j++) {
      const double ty = sin(j * 0.5 * EVSL_PI / (ny + 1.0));
      for (int i = 1; i <= nx; i++) {
        const double tx = sin(i * 0.5 * EVSL_PI / (nx + 1.0));
        const double t = 4.0 * (tx * tx + ty * ty + tz * tz);
        if (t >= a - eps && t <= b + eps) ev[m++] = t;
      }
    }
  }
  qsort(ev, m, sizeof(double), cmp_double);
  *count = m;
  return ev;
}
#endif
