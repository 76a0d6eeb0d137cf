"""Pins of the oracle's KPM DOS and slicers (SRC/spslice.c, SRC/spslice2.c,
SRC/simpson.c, SRC/landos.c) by dense eigendecompositions, quadrature and
exact integration."""
import numpy as np
import pytest
from scipy.integrate import quad

from evsl_amd import matrices as M
from tests.conftest import libc_rand_double


def dense(A):
    return A.to_scipy().toarray()


def kpm_dense(A, vecs, Mdeg, intv, damping_jac):
    """Moments and count from the eigendecomposition of A for the given
    (unnormalised) vectors."""
    lam, Q = np.linalg.eigh(dense(A))
    ctr, wid = (intv[3] + intv[2]) / 2, (intv[3] - intv[2]) / 2
    lt = (lam - ctr) / wid
    th = np.arccos(np.clip(lt, -1, 1))
    mu = np.zeros(Mdeg + 1)
    for v in vecs:
        v = v / np.linalg.norm(v)
        c2 = (Q.T @ v) ** 2
        for k in range(Mdeg + 1):
            mu[k] += (1.0 if k == 0 else 2.0) * damping_jac[k] * np.sum(c2 * np.cos(k * th))
    return mu / (len(vecs) * np.pi)


@pytest.mark.parametrize("kind", ["lap", "rand"])
def test_kpmdos_matches_dense(oracle, libc, kind):
    if kind == "lap":
        A = M.laplacian(12, 11, 1)
        intv = np.array([1.0, 3.0, 0.0, 8.0])
    else:
        A = M.random_symmetric(150, 6, 12, seed=5)
        lo, hi = M.gershgorin(A)
        intv = np.array([-0.5, 0.7, lo, hi])
    oracle.set_csr(A)
    Mdeg, nvec = 40, 5
    libc.srand(1234)
    err, mu, ecnt = oracle.kpmdos(Mdeg, 1, nvec, intv)
    assert err == 0
    libc.srand(1234)
    vecs = [libc_rand_double(libc, A.n) for _ in range(nvec)]
    jac = oracle.dampcf(Mdeg, 1)
    jac[0] = 1.0
    ref = kpm_dense(A, vecs, Mdeg, intv, jac)
    assert mu[0] * np.pi == pytest.approx(1.0, abs=1e-15)
    np.testing.assert_allclose(mu, ref, rtol=0, atol=1e-13)
    # eigenvalue count = n * integral of the KPM density over [a, b], by quadrature in theta
    ctr, wid = (intv[3] + intv[2]) / 2, (intv[3] - intv[2]) / 2
    b1 = np.arccos(max(-1 + 2.2e-16, (max(intv[0], intv[2]) - ctr) / wid))
    b2 = np.arccos(min(1 - 2.2e-16, (min(intv[1], intv[3]) - ctr) / wid))
    k = np.arange(Mdeg + 1)
    f = lambda th: np.sum(mu * np.cos(k * th))
    val, _ = quad(f, b2, b1, limit=200, epsabs=1e-13)
    assert ecnt == pytest.approx(A.n * val, rel=1e-10, abs=1e-9)


def test_kpmdos_count_statistics(oracle, libc):
    A = M.laplacian(30, 30, 1)
    oracle.set_csr(A)
    intv = np.array([0.4, 0.8, 0.0, 8.0])
    libc.srand(7)
    err, mu, ecnt = oracle.kpmdos(300, 1, 60, intv)
    exact = len(M.laplacian_eigs(30, 30, 1, 0.4, 0.8))
    assert err == 0 and abs(ecnt - exact) < 0.2 * exact


def test_kpmdos_invalid(oracle):
    oracle.set_csr(M.laplacian(5, 5, 1))
    assert oracle.kpmdos(10, 1, 2, np.array([1.0, 0.5, 0.0, 8.0]))[0] == -1


def test_intChx_matches_quadrature(oracle):
    rng = np.random.default_rng(3)
    mu = rng.standard_normal(31) / np.arange(1, 32)
    xi = np.linspace(-0.7, 0.9, 23)
    yi = oracle.intChx(mu, xi)
    k = np.arange(len(mu))
    f = lambda th: np.sum(mu * np.cos(k * th))
    th0 = np.arccos(xi[0])
    ref = [quad(f, np.arccos(x), th0, epsabs=1e-14)[0] for x in xi]
    np.testing.assert_allclose(yi, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n_int", [2, 5, 8])
def test_spslicer_invariants(oracle, libc, n_int):
    A = M.laplacian(40, 40, 1)
    oracle.set_csr(A)
    intv = np.array([0.3, 1.9, 0.0, 8.0])
    libc.srand(2)
    err, mu, ecnt = oracle.kpmdos(200, 1, 20, intv)
    npts = int(10 * ecnt)
    err, sli = oracle.spslicer(mu, intv, n_int, npts)
    assert err == 0
    assert sli[0] == intv[0] and sli[-1] == intv[1]
    assert np.all(np.diff(sli) > 0)
    ctr, wid = 4.0, 4.0
    xi = np.linspace((intv[0] - ctr) / wid, (intv[1] - ctr) / wid, npts)
    grid = ctr + wid * xi
    for s in sli[1:-1]:
        assert np.min(np.abs(grid - s)) == 0.0          # boundaries are grid points
    yi = oracle.intChx(mu, xi)
    cell = np.max(np.diff(yi))
    cuts = [0.0] + [yi[np.argmin(np.abs(grid - s))] for s in sli[1:-1]] + [yi[-1]]
    shares = np.diff(cuts)
    assert np.all(np.abs(shares - yi[-1] / n_int) <= 3 * cell + 1e-12)


def test_spslicer_single_and_invalid(oracle):
    mu = np.r_[1 / np.pi, np.zeros(10)]
    err, sli = oracle.spslicer(mu, np.array([0.1, 0.2, 0.0, 1.0]), 1, 50)
    assert err == 0 and list(sli) == [0.1, 0.2]
    assert oracle.spslicer(mu, np.array([0.2, 0.1, 0.0, 1.0]), 3, 50)[0] == -1


def test_simpson_exact_for_cubics(oracle):
    for npts in (9, 10):
        x = np.linspace(-1.0, 2.0, npts)
        f = x ** 3 - 2 * x + 1
        F = lambda t: t ** 4 / 4 - t ** 2 + t
        y = oracle.simpson(x, f)
        exact = F(x) - F(x[0])
        even = np.arange(0, npts, 2) if npts % 2 else np.arange(0, npts - 1, 2)
        np.testing.assert_allclose(y[even], exact[even], rtol=0, atol=1e-12)
        h = x[1] - x[0]
        # trapezoid points: local error bound h^3/12 * max|f''| (= 12 on [-1,2])
        assert np.max(np.abs(y - exact)) <= h ** 3 * 12 / 12 * 2 + 1e-12


def test_spslicer2_invariants(oracle):
    x = np.linspace(0.0, 4.0, 201)
    y = 1.0 + np.sin(x) ** 2
    sli = oracle.spslicer2(x, y, 4)
    assert sli[0] == 0.0 and sli[-1] == 4.0 and np.all(np.diff(sli) > 0)
    cum = oracle.simpson(x, y)
    want = cum[-1] / 4
    for a, b in zip(sli[:-2], sli[1:-1]):
        ia, ib = np.searchsorted(x, a), np.searchsorted(x, b)
        assert cum[ib] - cum[ia] >= want - 1e-12


def test_landos_count(oracle, libc):
    A = M.laplacian(20, 20, 1)
    oracle.set_csr(A)
    intv = np.array([0.5, 2.5, 0.0, 8.0])
    libc.srand(9)
    err, x, y, neig = oracle.LanDos(30, 40, 200, intv)
    exact = len(M.laplacian_eigs(20, 20, 1, 0.5, 2.5))
    assert err == 0
    np.testing.assert_allclose(x, np.linspace(0.5, 2.5, 200), rtol=0, atol=1e-15)
    assert np.all(y >= 0)
    assert abs(neig - exact) < 0.15 * exact


def exact_smoothed_dos(vals, x, intv, msteps):
    """The reference's exact DOS (SRC/exDOS.c, used by TESTS/Landos/LanDos.c):
    Gaussians of width sigma = H / sqrt(8 ln 1.25), H = (lmax - lmin)/(M - 1),
    M = min(n, msteps, 30), summed at the grid points and normalised to unit
    mass (1 / (n sqrt(2 pi sigma^2)))."""
    n = len(vals)
    m = min(n, msteps, 30)
    h = (intv[3] - intv[2]) / (m - 1)
    sigma = h / np.sqrt(8 * np.log(1.25))
    s2 = 2 * sigma * sigma
    width = sigma * np.sqrt(-2 * np.log(1e-8))
    y = np.zeros_like(x)
    for t in vals:
        if t < intv[0] or t > intv[1]:
            continue
        sel = np.abs(x - t) <= width
        y[sel] += np.exp(-(x[sel] - t) ** 2 / s2)
    return y / (n * np.sqrt(s2 * np.pi))


def test_landos_matches_exact_dos(oracle, libc):
    """The reference LanDos test problem (TESTS/Landos/LanDos.c: 3-D 20^3
    Laplacian, msteps 40, npts 200, nvec 100, the whole spectrum [0, 12]):
    the Lanczos DOS agrees with the exact smoothed DOS of the closed-form
    eigenvalues to a few percent of its peak, and the count to 1%."""
    A = M.laplacian(20, 20, 20)
    oracle.set_csr(A)
    intv = np.array([0.0, 12.0, 0.0, 12.0])
    libc.srand(2)
    err, x, y, neig = oracle.LanDos(100, 40, 200, intv)
    assert err == 0
    ev = M.laplacian_eigs(20, 20, 20, -1.0, 13.0)
    yex = exact_smoothed_dos(ev, x, intv, 40)
    assert np.abs(y - yex).max() < 0.04 * yex.max()
    assert np.abs(y - yex).sum() < 0.02 * yex.sum()
    assert abs(neig - len(ev)) < 0.01 * len(ev)
