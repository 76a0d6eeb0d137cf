/* A driver with the call sequence of the reference's LanDos test
 * (TESTS/Landos/LanDos.c:44-115), compiled against include/evsl.h and linked
 * against libevsl_amd.so: EVSLStart -> SetAMatrix -> LanTrbounds -> LanDos,
 * and exDOS on the closed-form eigenvalues for the comparison curve.  Checks
 * that the bounds enclose the closed-form spectrum to 1e-8, that the Lanczos
 * DOS stays within 5% of the exact curve's peak and its count within 2%.
 *   usage: landos_drv num      (a num^3 Laplacian)
 * Prints the figures and "PASS" or exits non-zero. */
#include <stdio.h>
#include <stdlib.h>

#include "evsl.h"
#include "lap_common.h"

int main(int argc, char *argv[]) {
  const int num = argc > 1 ? atoi(argv[1]) : 20;
  const int msteps = 40, npts = 200, nvec = 100;
  double intv[4], neig, lmin = 0.0, lmax = 0.0;
  cooMat Acoo;
  csrMat Acsr;
  srand(2);
  if (EVSLStart()) return 4;
  if (lap_coo(num, num, num, &Acoo) || cooMat_to_csrMat(0, &Acoo, &Acsr)) return 3;
  const int n = Acoo.nrows;
  double *xHist = evsl_Malloc(npts, double), *yHist = evsl_Malloc(npts, double);
  double *xdos = evsl_Malloc(npts, double), *ydos = evsl_Malloc(npts, double);
  SetStdEig();
  if (SetAMatrix(&Acsr)) return 5;
  double *vinit = evsl_Malloc_device(n, double);
  rand_double_device(n, vinit);
  int ierr = LanTrbounds(50, 200, 1e-8, vinit, 1, &lmin, &lmax, NULL);
  if (ierr) {
    printf("LanTrbounds error %d\n", ierr);
    return 1;
  }
  intv[0] = lmin;
  intv[1] = lmax;
  intv[2] = lmin;
  intv[3] = lmax;
  int numev;
  double *ev = lap_eigs(num, num, num, lmin - 1.0, lmax + 1.0, &numev);
  const double scale = fabs(ev[0]) + fabs(ev[numev - 1]);
  const int bounds_ok = lmin <= ev[0] + 1e-12 * scale && lmax >= ev[numev - 1] - 1e-12 * scale &&
                        ev[0] - lmin < 2e-8 * scale && lmax - ev[numev - 1] < 2e-8 * scale;
  ierr = LanDos(nvec, msteps, npts, xdos, ydos, &neig, intv);
  if (ierr) {
    printf("LanDos error %d\n", ierr);
    return 1;
  }
  /* the exact curve over the eigenvalues inside [lmin, lmax] */
  int m = 0;
  for (int i = 0; i < numev; i++)
    if (ev[i] >= lmin && ev[i] <= lmax) ev[m++] = ev[i];
  ierr = exDOS(ev, m, npts, xHist, yHist, intv);
  if (ierr) {
    printf("exDOS returned %d\n", ierr);
    return 1;
  }
  double peak = 0.0, dmax = 0.0;
  for (int i = 0; i < npts; i++) {
    peak = evsl_max(peak, yHist[i]);
    dmax = evsl_max(dmax, fabs(ydos[i] - yHist[i]));
  }
  const int dos_ok = dmax < 0.05 * peak && fabs(neig - n) < 0.02 * n;
  printf("bounds [%.12f, %.12f] exact [%.12f, %.12f]; LanDos neig %.2f of %d, "
         "max|y - y_exact| %.3e of peak %.3e\n",
         lmin, lmax, ev[0], ev[m - 1], neig, n, dmax, peak);
  free(ev);
  evsl_Free(xHist);
  evsl_Free(yHist);
  evsl_Free(xdos);
  evsl_Free(ydos);
  evsl_Free_device(vinit);
  free_coo(&Acoo);
  free_csr(&Acsr);
  EVSLFinish();
  if (!(bounds_ok && dos_ok)) return 1;
  printf("PASS\n");
  return 0;
}
