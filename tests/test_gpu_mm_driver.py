"""GPU: the reference's Matrix Market driver (TESTS/PLanN/MMPLanN.c) on its
own test matrix MATRICES/A.mtx (tests/golden/A.mtx, matfile line
"A.mtx stiff1 MM1 0.1 0.3 5"): read_coo_MM + cooMat_to_csrMat ->
LanTrbounds(50, 200, 1e-8) -> LanDos(nvec 30, msteps 80, npts 100*n_int) ->
spslicer2 -> per slice find_pol (damping 2, thresholds .8/.2) and ChebLanNr
(mlan = max(5 nev, 300), tol 1e-8).  Bars: bounds and DOS equal to the
oracle's on the same libc rand() streams; every slice's count and
eigenvalues equal to LAPACK's (tests/golden/A_eigs.npy) to 1e-10."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_mm_driver_pipeline(ev, oracle, libc):
    err, A = ev.read_mm(os.path.join(GOLDEN, "A.mtx"))
    assert err == 0
    w = np.load(os.path.join(GOLDEN, "A_eigs.npy"))
    n = A.n
    assert n == len(w)
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    libc.srand(1)
    v = oracle.rand_double(n)
    eb, lmin, lmax = ev.LanTrbounds(50, 200, 1e-8, v, 1)
    ob, lmin_o, lmax_o = oracle.LanTrbounds(50, 200, 1e-8, v, 1)
    assert eb == ob == 0, ev.last_error()
    assert lmin == pytest.approx(lmin_o, rel=1e-9, abs=1e-12)
    assert lmax == pytest.approx(lmax_o, rel=1e-9)
    assert lmin <= w[0] + 1e-12 and lmax >= w[-1] - 1e-12
    a, b, n_int = 0.1, 0.3, 5
    xintv = np.array([a, b, lmin, lmax])
    npts = 100 * n_int
    libc.srand(2)
    ed, xdos, ydos, ecount = ev.LanDos(30, 80, npts, xintv)
    libc.srand(2)
    eo, xo, yo, eco = oracle.LanDos(30, 80, npts, xintv)
    assert ed == eo == 0
    np.testing.assert_allclose(ydos, yo, rtol=1e-8, atol=1e-12)
    assert ecount == pytest.approx(eco, rel=1e-8)
    exact_total = int(np.count_nonzero((w >= a) & (w <= b)))
    assert abs(ecount - exact_total) < 0.15 * exact_total  # stochastic estimate
    sli = ev.spslicer2(xdos, ydos, n_int)
    assert np.all(np.diff(sli) > 0) and sli[0] == a and sli[-1] == pytest.approx(b)
    nev = int(1.2 * int(1 + ecount / n_int))
    total = 0
    for s in range(n_int):
        iv = np.array([sli[s], sli[s + 1], lmin, lmax])
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext = 2, 0.8, 0.2
        assert ev.find_pol(iv, pol) == 0
        mlan = min(max(5 * nev, 300), n)
        e, lam, Y, res = ev.ChebLanNr(iv, mlan, 1e-8, v, pol)
        ev.free_pol(pol)
        assert e == 0, ev.last_error()
        exact = w[(w >= iv[0]) & (w <= iv[1])]
        print(f"  slice {s}: [{iv[0]:.6f}, {iv[1]:.6f}] deg {pol.deg} found {len(lam)} "
              f"exact {len(exact)}", flush=True)
        assert len(lam) == len(exact)
        np.testing.assert_allclose(np.sort(lam), exact, rtol=1e-10, atol=1e-13)
        assert np.all(res < 1e-8)
        R = A.to_scipy() @ Y - Y * lam
        np.testing.assert_allclose(np.linalg.norm(R, axis=0), res, rtol=1e-5, atol=1e-13)
        total += len(lam)
    assert total == exact_total
