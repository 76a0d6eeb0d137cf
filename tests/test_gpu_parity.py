"""GPU parity: the HIP path, called through the C ABI, against the CPU oracle
on the same matrices, intervals, filters and libc rand() streams.

Bars: identical per-slice eigenvalue counts,
FP64 eigenvalues within 1e-10 relative, residuals below tol, KPM moments
within 1e-12 (absolute; mu_0 = 1/pi), slice boundaries identical."""
import ctypes as C
import time

import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import VEC_GAUSSIAN, VEC_RADEMACHER
from evsl_amd.types import CsrMat, c_double_p

pytestmark = pytest.mark.gpu


def slice_pol(lib, intv):
    pol = lib.default_pol()
    pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
    assert lib.find_pol(intv, pol) == 0
    return pol


def both_solve(ev, oracle, A, intv, tol, v, maxit):
    pg, po = slice_pol(ev, intv), slice_pol(oracle, intv)
    assert pg.deg == po.deg
    t0 = time.time()
    eg, lg, Yg, rg = ev.ChebLanNr(intv, maxit, tol, v, pg)
    t1 = time.time()
    eo, lo, Yo, ro = oracle.ChebLanNr(intv, maxit, tol, v, po)
    t2 = time.time()
    ev.free_pol(pg)
    oracle.free_pol(po)
    assert eg == 0, ev.last_error()
    assert eo == 0
    print(f"  slice {intv[:2]}: deg {pg.deg}, gpu {len(lg)} pairs {t1 - t0:.2f}s, "
          f"oracle {len(lo)} pairs {t2 - t1:.2f}s", flush=True)
    return (lg, Yg, rg), (lo, Yo, ro)


def check_pairs(A, lam, Y, res, intv, tol):
    assert np.all(res < tol)
    assert np.all((lam >= intv[0] - 1e-15) & (lam <= intv[1] + 1e-15))
    R = A.to_scipy() @ Y - Y * lam
    np.testing.assert_allclose(np.linalg.norm(R, axis=0), res, rtol=1e-5, atol=1e-13)
    assert np.abs(Y.T @ Y - np.eye(Y.shape[1])).max() < 1e-7


@pytest.mark.parametrize("case", ["lap2d", "randsym"])
def test_kpmdos_parity(ev, oracle, libc, case):
    if case == "lap2d":  # the small plain configuration: 100x100, Mdeg 300, 60 vectors
        A = M.laplacian(100, 100, 1)
        intv = np.array([0.4, 0.8, 0.0, 8.0])
        Mdeg, nvec = 300, 60
    else:
        A = M.random_symmetric(5000, 8, 32, seed=2)
        lo, hi = M.gershgorin(A)
        intv = np.array([-1.0, 1.5, lo, hi])
        Mdeg, nvec = 200, 13
    assert ev.set_matrix(A) == 0, ev.last_error()
    oracle.set_csr(A)
    libc.srand(77)
    eg, mug, ecg = ev.kpmdos(Mdeg, 1, nvec, intv)
    libc.srand(77)
    eo, muo, eco = oracle.kpmdos(Mdeg, 1, nvec, intv)
    assert eg == 0 and eo == 0
    assert mug[0] == pytest.approx(1 / np.pi, abs=1e-16)
    assert np.max(np.abs(mug - muo)) <= 1e-12, np.max(np.abs(mug - muo))
    assert abs(ecg - eco) <= 1e-10 * abs(eco)
    # slicing on the GPU moments gives the oracle's boundaries
    npts = int(10 * ecg)
    e1, sg = ev.spslicer(mug, intv, 4, npts)
    e2, so = oracle.spslicer(muo, intv, 4, npts)
    assert e1 == e2 == 0 and np.array_equal(sg, so)


def test_kpmdos_device_random_vectors(ev):
    A = M.laplacian(60, 60, 1)
    assert ev.set_matrix(A) == 0
    intv = np.array([0.4, 0.8, 0.0, 8.0])
    exact = len(M.laplacian_eigs(60, 60, 1, 0.4, 0.8))
    for kind in (VEC_RADEMACHER, VEC_GAUSSIAN):
        e, mu, ec = ev.kpmdos(200, 1, 40, intv, vec_kind=kind, seed=5)
        e2, mu2, ec2 = ev.kpmdos(200, 1, 40, intv, vec_kind=kind, seed=5)
        e3, mu3, _ = ev.kpmdos(200, 1, 40, intv, vec_kind=kind, seed=6)
        assert e == e2 == e3 == 0
        assert np.array_equal(mu, mu2) and ec == ec2       # deterministic
        assert not np.array_equal(mu, mu3)                 # seeded
        assert mu[0] == pytest.approx(1 / np.pi, abs=1e-16)
        assert abs(ec - exact) < 0.15 * exact


def test_plain_small_config_parity(ev, oracle, libc):
    """2-D 100x100 Laplacian: LanBounds (80 steps) -> kpmdos (300, Jackson, 60)
    -> spslicer [0.4, 0.8] into 2 -> find_pol + ChebLanNr tol 1e-8 per slice."""
    A = M.laplacian(100, 100, 1)
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    libc.srand(1)
    v = oracle.rand_double(A.n)
    eb, lmin_g, lmax_g = ev.LanBounds(80, v)
    ob, lmin_o, lmax_o = oracle.LanBounds(80, v)
    assert eb == ob == 0
    assert lmin_g == pytest.approx(lmin_o, rel=1e-10, abs=1e-12)
    assert lmax_g == pytest.approx(lmax_o, rel=1e-10)
    intv = np.array([0.4, 0.8, lmin_o, lmax_o])
    libc.srand(2)
    _, mug, ecg = ev.kpmdos(300, 1, 60, intv)
    libc.srand(2)
    _, muo, eco = oracle.kpmdos(300, 1, 60, intv)
    assert np.max(np.abs(mug - muo)) <= 1e-12
    npts = int(10 * eco)
    _, sli = oracle.spslicer(muo, intv, 2, npts)
    _, slig = ev.spslicer(mug, intv, 2, npts)
    assert np.array_equal(sli, slig)
    ev_int = int(1 + eco / 2)
    mlan = min(max(5 * (ev_int + 2), 300), A.n)
    libc.srand(3)
    v0 = oracle.rand_double(A.n)
    total = 0
    for s in range(2):
        iv = np.array([sli[s], sli[s + 1], lmin_o, lmax_o])
        (lg, Yg, rg), (lo, Yo, ro) = both_solve(ev, oracle, A, iv, 1e-8, v0, mlan)
        exact = M.laplacian_eigs(100, 100, 1, iv[0], iv[1])
        assert len(lg) == len(lo), (len(lg), len(lo))
        np.testing.assert_allclose(np.sort(lg), np.sort(lo), rtol=1e-10)
        check_pairs(A, lg, Yg, rg, iv, 1e-8)
        assert len(lg) == len(exact)
        total += len(lg)
    assert total == len(M.laplacian_eigs(100, 100, 1, 0.4, 0.8))


@pytest.mark.parametrize("kind", ["perturbed3d", "anderson", "randsym"])
def test_cheblannr_parity_dense(ev, oracle, libc, kind):
    if kind == "perturbed3d":
        A = M.perturb_diag(M.laplacian(14, 14, 14), 1e-3, seed=3)
        intv = np.array([0.9, 1.2, 0.0, 12.001])
    elif kind == "anderson":
        A = M.anderson(12, 16.5, seed=4)
        lo, hi = M.gershgorin(A)
        intv = np.array([-0.3, 0.3, lo, hi])
    else:
        A = M.random_symmetric(3000, 8, 32, seed=9)
        lo, hi = M.gershgorin(A)
        intv = np.array([0.0, 0.25, lo, hi])
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    libc.srand(5)
    v = oracle.rand_double(A.n)
    (lg, Yg, rg), (lo_, Yo, ro) = both_solve(ev, oracle, A, intv, 1e-8, v, min(A.n, 800))
    dense = np.linalg.eigvalsh(A.to_scipy().toarray())
    exact = dense[(dense >= intv[0]) & (dense <= intv[1])]
    assert len(lg) == len(lo_) == len(exact), (len(lg), len(lo_), len(exact))
    np.testing.assert_allclose(np.sort(lg), np.sort(lo_), rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(np.sort(lg), exact, rtol=1e-10, atol=1e-12)
    check_pairs(A, lg, Yg, rg, intv, 1e-8)


def test_end_interval_slices(ev, oracle, libc):
    A = M.laplacian(40, 40, 1)
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    libc.srand(6)
    v = oracle.rand_double(A.n)
    for intv in (np.array([0.0, 0.15, 0.0, 8.0]), np.array([7.85, 8.0, 0.0, 8.0])):
        (lg, Yg, rg), (lo, _, _) = both_solve(ev, oracle, A, intv, 1e-8, v, 600)
        exact = M.laplacian_eigs(40, 40, 1, intv[0], intv[1])
        assert len(lg) == len(lo) == len(exact)
        np.testing.assert_allclose(np.sort(lg), exact, rtol=1e-10)
        check_pairs(A, lg, Yg, rg, intv, 1e-8)


def test_landos_parity(ev, oracle, libc):
    A = M.laplacian(30, 30, 1)
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    intv = np.array([0.5, 3.0, 0.0, 8.0])
    libc.srand(12)
    eg, xg, yg, ng = ev.LanDos(20, 40, 200, intv)
    libc.srand(12)
    eo, xo, yo, no = oracle.LanDos(20, 40, 200, intv)
    assert eg == eo == 0
    assert np.array_equal(xg, xo)
    np.testing.assert_allclose(yg, yo, rtol=1e-9, atol=1e-12)
    assert ng == pytest.approx(no, rel=1e-9)


def test_user_matvec_and_device_csr_paths(ev, oracle, libc):
    A = M.laplacian(25, 25, 1)
    S = A.to_scipy()
    intv = np.array([0.5, 0.9, 0.0, 8.0])
    libc.srand(21)
    v = oracle.rand_double(A.n)
    # 1) host CSR
    assert ev.set_matrix(A) == 0
    pol = slice_pol(ev, intv)
    e1, l1, Y1, r1 = ev.ChebLanNr(intv, 400, 1e-8, v, pol)
    # 2) user matvec on device vectors (host round trip inside the callback)
    n = A.n
    buf = np.empty(n)

    def mv(x, y, _):
        ev.lib.evsl_memcpy_device_to_host(buf.ctypes.data, C.cast(x, C.c_void_p), 8 * n)
        out = np.ascontiguousarray(S @ buf)
        ev.lib.evsl_memcpy_host_to_device(C.cast(y, C.c_void_p), out.ctypes.data, 8 * n)

    assert ev.set_matvec(n, mv) == 0
    e2, l2, Y2, r2 = ev.ChebLanNr(intv, 400, 1e-8, v, pol)
    # 3) device CSR made by evsl_create_csr_gpu
    gpu = CsrMat()
    ev.lib.evsl_create_csr_gpu(C.byref(A.csr), C.byref(gpu))
    assert ev.lib.SetAMatrix_device_csr(C.byref(gpu)) == 0
    e3, l3, Y3, r3 = ev.ChebLanNr(intv, 400, 1e-8, v, pol)
    # 4) start vector passed as a device pointer
    vd = ev.to_device(v)
    e4, l4, Y4, r4 = ev.ChebLanNr(intv, 400, 1e-8, vd, pol, n=n)
    ev.free_device(vd)
    ev.lib.evsl_free_csr_gpu(C.byref(gpu))
    ev.free_pol(pol)
    exact = M.laplacian_eigs(25, 25, 1, 0.5, 0.9)
    for e, l in ((e1, l1), (e2, l2), (e3, l3), (e4, l4)):
        assert e == 0 and len(l) == len(exact)
        np.testing.assert_allclose(np.sort(l), exact, rtol=1e-10)
    np.testing.assert_array_equal(l3, l1)
    np.testing.assert_array_equal(l4, l1)


def test_error_paths(ev):
    A = M.laplacian(8, 8, 1)
    assert ev.set_matrix(A) == 0
    pol = slice_pol(ev, np.array([0.4, 0.8, 0.0, 8.0]))
    v = np.ones(A.n)
    err, lam, Y, res = ev.ChebLanNr(np.array([0.8, 0.4, 0.0, 8.0]), 30, 1e-8, v, pol)
    assert err == 0 and len(lam) == 0
    err, lam, Y, res = ev.ChebLanNr(np.array([9.0, 10.0, 0.0, 8.0]), 30, 1e-8, v, pol)
    assert err == 0 and len(lam) == 0
    gam = pol.gam
    pol.gam = 1.5
    assert ev.ChebLanNr(np.array([0.4, 0.8, 0.0, 8.0]), 30, 1e-8, v, pol)[0] == -1
    pol.gam = gam
    ev.free_pol(pol)
    assert ev.kpmdos(10, 1, 2, np.array([0.8, 0.4, 0.0, 8.0]))[0] == -1
    assert ev.lib.SetGenEig() < 0


def test_full_size_slice_properties(ev):
    """3-D 100^3 Laplacian + U[0,1e-3] diagonal (n = 1e6): one slice of the
    [0.2, 0.3] workload, checked by size-independent properties (residuals,
    interval, orthonormality on a column sample) and against the closed-form
    count of the unperturbed Laplacian shifted by at most 1e-3."""
    A = M.perturb_diag(M.laplacian(100, 100, 100), 1e-3, seed=2024)
    assert ev.set_matrix(A) == 0
    intv = np.array([0.25, 0.2625, 0.0, 12.001])
    pol = slice_pol(ev, intv)
    rng = np.random.default_rng(0)
    v = rng.uniform(-1, 1, A.n)
    t0 = time.time()
    err, lam, Y, res = ev.ChebLanNr(intv, 785, 1e-8, v, pol)
    dt = time.time() - t0
    print(f"  n=1e6 slice: deg {pol.deg}, {len(lam)} pairs in {dt:.1f}s, stats {ev.stats()}",
          flush=True)
    ev.free_pol(pol)
    assert err == 0, ev.last_error()
    assert np.all(res < 1e-8)
    assert np.all((lam >= intv[0]) & (lam <= intv[1]))
    lo = len(M.laplacian_eigs(100, 100, 100, intv[0] - 1e-3, intv[1]))
    hi = len(M.laplacian_eigs(100, 100, 100, intv[0], intv[1] - 1e-3))
    assert min(lo, hi) - 2 <= len(lam) <= max(lo, hi) + 2
    cols = rng.choice(Y.shape[1], size=min(40, Y.shape[1]), replace=False)
    Ys = Y[:, cols]
    assert np.abs(Ys.T @ Ys - np.eye(len(cols))).max() < 1e-7
    R = A.to_scipy() @ Ys - Ys * lam[cols]
    assert np.max(np.linalg.norm(R, axis=0)) < 1e-8


def both_solve_tr(ev, oracle, intv, lanm, nev, maxit, tol, v):
    pg, po = slice_pol(ev, intv), slice_pol(oracle, intv)
    assert pg.deg == po.deg
    t0 = time.time()
    eg, lg, Yg, rg = ev.ChebLanTr(lanm, nev, intv, maxit, tol, v, pg)
    t1 = time.time()
    eo, lo, Yo, ro = oracle.ChebLanTr(lanm, nev, intv, maxit, tol, v, po)
    t2 = time.time()
    ev.free_pol(pg)
    oracle.free_pol(po)
    assert eg == 0, ev.last_error()
    assert eo == 0
    print(f"  ChebLanTr {intv[:2]} lanm {lanm} nev {nev}: deg {pg.deg}, gpu {len(lg)} pairs "
          f"{t1 - t0:.2f}s, oracle {len(lo)} pairs {t2 - t1:.2f}s", flush=True)
    return (lg, Yg, rg), (lo, Yo, ro)


@pytest.mark.parametrize("case", ["driver", "restarts"])
def test_cheblantr_parity(ev, oracle, libc, case):
    """Thick-restart Lanczos (SRC/cheblanTr.c) against the oracle and the
    closed-form Laplacian spectrum.  'driver' uses the reference driver's
    sizes (TESTS/PLanR/LapPLanR.c:155-159: nev = ev_int + 2,
    mlan = max(4 nev, 300), max_its = 3 mlan); 'restarts' uses a small basis
    and nev so that several restarts, locking across restarts and growth of
    the locked block all happen."""
    A = M.laplacian(40, 40, 1)
    intv = np.array([0.4, 0.8, 0.0, 8.0])
    exact = M.laplacian_eigs(40, 40, 1, intv[0], intv[1])
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    libc.srand(8)
    v = oracle.rand_double(A.n)
    if case == "driver":
        nev = len(exact) + 2
        lanm = min(max(4 * nev, 300), A.n)
    else:
        nev = 4
        lanm = 60
    maxit = 3 * lanm if case == "driver" else 4000
    (lg, Yg, rg), (lo, _, ro) = both_solve_tr(ev, oracle, intv, lanm, nev, maxit, 1e-8, v)
    assert len(lg) == len(lo) == len(exact), (len(lg), len(lo), len(exact))
    np.testing.assert_allclose(np.sort(lg), np.sort(lo), rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(np.sort(lg), exact, rtol=1e-10)
    check_pairs(A, lg, Yg, rg, intv, 1e-8)


def test_cheblantr_dense_perturbed(ev, oracle, libc):
    A = M.perturb_diag(M.laplacian(12, 12, 12), 1e-3, seed=5)
    intv = np.array([1.0, 1.3, 0.0, 12.001])
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    libc.srand(9)
    v = oracle.rand_double(A.n)
    dense = np.linalg.eigvalsh(A.to_scipy().toarray())
    exact = dense[(dense >= intv[0]) & (dense <= intv[1])]
    nev = len(exact) + 2
    lanm = max(4 * nev, 300)
    (lg, Yg, rg), (lo, _, _) = both_solve_tr(ev, oracle, intv, lanm, nev, 3 * lanm, 1e-8, v)
    assert len(lg) == len(lo) == len(exact)
    np.testing.assert_allclose(np.sort(lg), exact, rtol=1e-10, atol=1e-12)
    check_pairs(A, lg, Yg, rg, intv, 1e-8)


@pytest.mark.parametrize("bndtype", [1, 2])
def test_lantrbounds_parity(ev, oracle, libc, bndtype):
    """Thick-restart Lanczos bounds (SRC/lanTrbounds.c) on the device against
    the oracle with the same libc rand() start vector: equal to 1e-10
    relative, and enclosing the dense spectrum."""
    for A in (M.laplacian(40, 40, 1), M.anderson(10, 16.5, seed=4),
              M.random_symmetric(2000, 8, 32, seed=9)):
        assert ev.set_matrix(A) == 0
        oracle.set_csr(A)
        libc.srand(3)
        v = oracle.rand_double(A.n)
        eg, lg, hg = ev.LanTrbounds(50, 200, 1e-8, v, bndtype)
        eo, lo, ho = oracle.LanTrbounds(50, 200, 1e-8, v, bndtype)
        assert eg == 0 and eo == 0, ev.last_error()
        assert lg == pytest.approx(lo, rel=1e-10, abs=1e-12)
        assert hg == pytest.approx(ho, rel=1e-10, abs=1e-12)
        d = np.linalg.eigvalsh(A.to_scipy().toarray())
        scale = abs(d[0]) + abs(d[-1])
        assert lg <= d[0] + 1e-12 * scale and hg >= d[-1] - 1e-12 * scale


@pytest.mark.parametrize("case", ["lap2d", "perturbed3d"])
def test_chebsi_parity(ev, oracle, libc, case):
    """Filtered subspace iteration (SRC/chebsi.c) on the device against the
    oracle from the same libc rand() start block, with the PSI driver's
    settings (TESTS/PSI/LapPSI.c:150-177): identical counts, eigenvalues to
    1e-10 of each other and of the exact spectrum, residuals below tol."""
    if case == "lap2d":
        A, intv = M.laplacian(60, 60, 1), np.array([0.4, 0.8, 0.0, 8.0])
    else:
        A, intv = M.perturb_diag(M.laplacian(12, 12, 12), 1e-3, seed=2), \
            np.array([1.0, 1.3, 0.0, 12.001])
    assert ev.set_matrix(A) == 0
    oracle.set_csr(A)
    d = np.linalg.eigvalsh(A.to_scipy().toarray())
    exact = d[(d >= intv[0]) & (d <= intv[1])]
    nev = int(1.2 * (1 + len(exact)))
    pg, po = ev.default_pol(), oracle.default_pol()
    for p in (pg, po):
        p.damping, p.thresh_int = 0, 0.5
    assert ev.find_pol(intv, pg) == 0 and oracle.find_pol(intv, po) == 0
    libc.srand(4)
    V0 = oracle.rand_double(A.n * nev).reshape(nev, A.n).T
    t0 = time.time()
    eg, lg, Yg, rg = ev.ChebSI(nev, intv, 1000, 1e-8, V0, pg)
    t1 = time.time()
    eo, lo, _, ro = oracle.ChebSI(nev, intv, 1000, 1e-8, V0, po)
    t2 = time.time()
    ev.free_pol(pg)
    oracle.free_pol(po)
    assert eg == 0 and eo == 0, ev.last_error()
    print(f"  ChebSI {case}: nev {nev} deg {pg.deg}, gpu {len(lg)} pairs {t1 - t0:.2f}s, "
          f"oracle {len(lo)} pairs {t2 - t1:.2f}s", flush=True)
    assert len(lg) == len(lo) == len(exact)
    np.testing.assert_allclose(np.sort(lg), np.sort(lo), rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(np.sort(lg), exact, rtol=1e-10, atol=1e-12)
    check_pairs(A, lg, Yg, rg, intv, 1e-8)


def test_solvers_through_user_matvec(ev, oracle, libc):
    """ChebLanTr, ChebSI, the batched ChebLanNr and LanTrbounds through a
    user MVFunc (SetAMatvec, device vectors in the callback) give the
    closed-form spectrum, and the same results as through SetAMatrix."""
    A = M.laplacian(24, 24, 1)
    S = A.to_scipy()
    n = A.n
    intv = np.array([0.5, 0.9, 0.0, 8.0])
    exact = M.laplacian_eigs(24, 24, 1, 0.5, 0.9)
    libc.srand(31)
    v = oracle.rand_double(n)
    libc.srand(32)
    V0 = oracle.rand_double(n * (len(exact) + 4)).reshape(len(exact) + 4, n).T
    buf = np.empty(n)

    def mv(x, y, _):
        ev.lib.evsl_memcpy_device_to_host(buf.ctypes.data, C.cast(x, C.c_void_p), 8 * n)
        out = np.ascontiguousarray(S @ buf)
        ev.lib.evsl_memcpy_host_to_device(C.cast(y, C.c_void_p), out.ctypes.data, 8 * n)

    results = []
    for mode in ("csr", "matvec"):
        if mode == "csr":
            assert ev.set_matrix(A) == 0
        else:
            assert ev.set_matvec(n, mv) == 0
        pol = slice_pol(ev, intv)
        etr, ltr, _, _ = ev.ChebLanTr(300, len(exact) + 2, intv, 900, 1e-8, v, pol)
        psi = ev.default_pol()
        psi.damping, psi.thresh_int = 0, 0.5
        assert ev.find_pol(intv, psi) == 0
        esi, lsi, _, _ = ev.ChebSI(len(exact) + 4, intv, 1000, 1e-8, V0, psi)
        eb, lo, hi = ev.LanTrbounds(50, 200, 1e-8, v, 1)
        ebt, outs = ev.ChebLanNr_batch([intv, intv], 400, 1e-8, v, [pol, pol])
        ev.free_pol(pol)
        ev.free_pol(psi)
        assert etr == 0 and esi == 0 and eb == 0 and ebt == 0, ev.last_error()
        for lam in (ltr, lsi, outs[0][0], outs[1][0]):
            assert len(lam) == len(exact)
            np.testing.assert_allclose(np.sort(lam), exact, rtol=1e-10)
        results.append((np.sort(ltr), np.sort(lsi), lo, hi))
    (t1, s1, lo1, hi1), (t2, s2, lo2, hi2) = results
    np.testing.assert_allclose(t1, t2, rtol=1e-10)
    np.testing.assert_allclose(s1, s2, rtol=1e-10)
    assert lo1 == pytest.approx(lo2, rel=1e-9, abs=1e-12) and hi1 == pytest.approx(hi2, rel=1e-9)


def test_solver_error_returns(ev):
    """Reference error conventions for the added solvers: an empty interval
    returns 0 with no pairs (ChebLanTr, ChebSI), |gam| > 1 returns -1, and a
    Krylov dimension too small for the restart set returns EVSL_ERR_ARG
    (the reference exits)."""
    A = M.laplacian(20, 20, 1)
    assert ev.set_matrix(A) == 0
    v = np.random.default_rng(3).uniform(-1, 1, A.n)
    pol = slice_pol(ev, np.array([0.5, 0.9, 0.0, 8.0]))
    err, lam, _, _ = ev.ChebLanTr(100, 5, np.array([0.9, 0.5, 0.0, 8.0]), 300, 1e-8, v, pol)
    assert err == 0 and len(lam) == 0
    err, lam, _, _ = ev.ChebSI(5, np.array([0.9, 0.5, 0.0, 8.0]), 100, 1e-8,
                               np.ones((A.n, 5)), pol)
    assert err == 0 and len(lam) == 0
    gam = pol.gam
    pol.gam = 2.0
    assert ev.ChebLanTr(100, 5, np.array([0.5, 0.9, 0.0, 8.0]), 300, 1e-8, v, pol)[0] == -1
    assert ev.ChebSI(5, np.array([0.5, 0.9, 0.0, 8.0]), 100, 1e-8, np.ones((A.n, 5)), pol)[0] == -1
    pol.gam = gam
    assert ev.LanTrbounds(4, 50, 1e-8, v, 1)[0] == -104
    ev.free_pol(pol)


def test_landos_matches_exact_dos_on_device(ev, libc):
    """Device LanDos on the reference LanDos test problem (3-D 20^3 Laplacian,
    msteps 40, npts 200, nvec 100) against the exact smoothed DOS of the
    closed-form spectrum (SRC/exDOS.c)."""
    from test_oracle_kpm import exact_smoothed_dos
    A = M.laplacian(20, 20, 20)
    assert ev.set_matrix(A) == 0
    intv = np.array([0.0, 12.0, 0.0, 12.0])
    libc.srand(2)
    err, x, y, neig = ev.LanDos(100, 40, 200, intv)
    assert err == 0
    lam = M.laplacian_eigs(20, 20, 20, -1.0, 13.0)
    yex = exact_smoothed_dos(lam, x, intv, 40)
    assert np.abs(y - yex).max() < 0.04 * yex.max()
    assert np.abs(y - yex).sum() < 0.02 * yex.sum()
    assert abs(neig - len(lam)) < 0.01 * len(lam)
