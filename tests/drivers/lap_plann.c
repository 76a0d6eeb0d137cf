/* A driver with the call sequence of the reference's LapPLanN
 * (TESTS/PLanN/LapPLanN.c:102-265), compiled against include/evsl.h and
 * linked against libevsl_amd.so: EVSLStart -> SetAMatrix -> kpmdos ->
 * spslicer -> per slice set_pol_def / find_pol / ChebLanNr -> StatsPrint ->
 * EVSLFinish.  Every slice's eigenvalues are checked against the closed-form
 * Laplacian spectrum (count exact, values to 1e-10, residuals below tol).
 *   usage: lap_plann nx ny nz a b nslices
 * Prints one line per slice and "PASS <total>" or exits non-zero. */
#include <stdio.h>
#include <stdlib.h>

#include "evsl.h"
#include "lap_common.h"

int main(int argc, char *argv[]) {
  if (argc != 7) {
    fprintf(stderr, "usage: %s nx ny nz a b nslices\n", argv[0]);
    return 2;
  }
  const int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]);
  double a = atof(argv[4]), b = atof(argv[5]);
  const int nslices = atoi(argv[6]);
  const double lmin = 0.0, lmax = nz == 1 ? 8.0 : 12.0, tol = 1e-8;
  double xintv[4] = {a, b, lmin, lmax};
  const int n = nx * ny * nz, Mdeg = 300, nvec = 60;
  cooMat Acoo;
  csrMat Acsr;
  double ecount;
  int ierr, bad = 0, total = 0;

  if (lap_coo(nx, ny, nz, &Acoo) || cooMat_to_csrMat(0, &Acoo, &Acsr)) return 3;
  if (EVSLStart()) return 4;
  if (SetAMatrix(&Acsr)) return 5;
  srand(1);
  double *mu = evsl_Malloc(Mdeg + 1, double);
  ierr = kpmdos(Mdeg, 1, nvec, xintv, mu, &ecount);
  if (ierr) {
    printf("kpmdos error %d\n", ierr);
    return 1;
  }
  const int npts = (int)(10 * ecount);
  double *sli = evsl_Malloc(nslices + 1, double);
  ierr = spslicer(sli, mu, Mdeg, xintv, nslices, npts);
  if (ierr) {
    printf("spslicer error %d\n", ierr);
    return 1;
  }
  const int ev_int = (int)(1 + ecount / ((double)nslices));
  double *vinit = evsl_Malloc_device(n, double);
  rand_double_device(n, vinit);
  for (int sl = 0; sl < nslices; sl++) {
    polparams pol;
    int nev2, nev_ex;
    double *lam, *Y, *res;
    StatsReset();
    a = sli[sl];
    b = sli[sl + 1];
    const int nev = ev_int + 2;
    int mlan = evsl_max(5 * nev, 300);
    mlan = evsl_min(mlan, n);
    xintv[0] = a;
    xintv[1] = b;
    xintv[2] = lmin;
    xintv[3] = lmax;
    set_pol_def(&pol);
    pol.damping = 2;
    pol.thresh_int = 0.8;
    pol.thresh_ext = 0.2;
    pol.max_deg = 3000;
    if (find_pol(xintv, &pol)) return 6;
    ierr = ChebLanNr(xintv, mlan, tol, vinit, &pol, &nev2, &lam, &Y, &res, NULL);
    if (ierr) {
      printf("ChebLanNr error %d\n", ierr);
      return 1;
    }
    int *ind = evsl_Malloc(nev2 > 0 ? nev2 : 1, int);
    sort_double(nev2, lam, ind);
    double *lam_ex = lap_eigs(nx, ny, nz, a, b, &nev_ex);
    double err = 0.0, rmax = 0.0;
    for (int i = 0; i < nev2; i++) {
      if (i < nev_ex) err = evsl_max(err, fabs(lam[i] - lam_ex[i]));
      rmax = evsl_max(rmax, res[ind[i]]);
    }
    const int ok = nev2 == nev_ex && err <= 1e-10 && rmax < tol;
    printf("slice %d [%.6f, %.6f] deg %d: found %d exact %d max|dlam| %.2e max res %.2e %s\n",
           sl, a, b, pol.deg, nev2, nev_ex, err, rmax, ok ? "ok" : "MISMATCH");
    bad += !ok;
    total += nev2;
    if (lam) evsl_Free(lam);
    if (Y) evsl_Free_device(Y);
    if (res) evsl_Free(res);
    free_pol(&pol);
    evsl_Free(ind);
    free(lam_ex);
    StatsPrint(NULL);
  }
  evsl_Free_device(vinit);
  evsl_Free(sli);
  evsl_Free(mu);
  free_coo(&Acoo);
  free_csr(&Acsr);
  EVSLFinish();
  if (bad) return 1;
  printf("PASS %d\n", total);
  return 0;
}
