"""Pins of the oracle's dense kernels and filtered Lanczos solvers
(SRC/misc_la.c, SRC/chebpoly.c ChebAv, SRC/cheblanNr.c, SRC/cheblanTr.c,
SRC/lanbounds.c) by LAPACK, dense eigendecompositions and closed-form
Laplacian spectra."""
import numpy as np
import pytest
import scipy.linalg as sl
from numpy.polynomial import chebyshev as npc

from evsl_amd import matrices as M


def test_tridiag_eig_matches_lapack(oracle):
    rng = np.random.default_rng(1)
    cases = [(rng.standard_normal(80), rng.standard_normal(79)),
             (np.abs(np.arange(-10, 11, dtype=float)), np.ones(20)),       # Wilkinson W21+
             (np.full(30, 2.0), np.full(29, -1.0)),
             (np.r_[np.ones(10), 1 + 1e-13 * np.arange(10)], np.r_[1e-9 * np.ones(9), 0.0, 1e-9 * np.ones(9)])]
    for d, e in cases:
        info, ev, Z = oracle.tridiag_eig(d, e, vectors=True)
        assert info == 0
        ref = sl.eigh_tridiagonal(d, e, eigvals_only=True)
        np.testing.assert_allclose(ev, ref, rtol=0, atol=1e-13 * max(1, np.abs(ref).max()))
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.abs(T @ Z - Z * ev).max() < 1e-12 * max(1, np.abs(ref).max())
        assert np.abs(Z.T @ Z - np.eye(len(d))).max() < 1e-12
        info2, ev2 = oracle.tridiag_eig(d, e)
        np.testing.assert_allclose(ev2, ev, rtol=0, atol=1e-14 * max(1, np.abs(ref).max()))


def test_sym_eig_matches_lapack(oracle):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((60, 60))
    A = X + X.T
    info, lam, Q = oracle.sym_eig(A)
    assert info == 0
    np.testing.assert_allclose(lam, np.linalg.eigvalsh(A), rtol=0, atol=1e-12)
    assert np.abs(A @ Q - Q * lam).max() < 1e-11
    assert np.abs(Q.T @ Q - np.eye(60)).max() < 1e-12


def test_cgs_dgks_projection(oracle):
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((500, 30)))
    v = rng.standard_normal(500)
    out, nrm = oracle.CGS_DGKS(Q, v)
    ref = v - Q @ (Q.T @ v)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-13)
    assert np.abs(Q.T @ out).max() < 1e-14 * np.linalg.norm(v)
    assert nrm == pytest.approx(np.linalg.norm(out), rel=1e-14)


def test_chebav_matches_dense(oracle):
    A = M.random_symmetric(80, 4, 10, seed=11)
    lo, hi = M.gershgorin(A)
    oracle.set_csr(A)
    pol = oracle.default_pol()
    intv = np.array([-0.3, 0.4, lo, hi])
    assert oracle.find_pol(intv, pol) == 0
    mu = pol.coefficients()
    lam, Q = np.linalg.eigh(A.to_scipy().toarray())
    v = np.random.default_rng(4).standard_normal(A.n)
    y = oracle.ChebAv(pol, v)
    ref = Q @ (npc.chebval((lam - pol.cc) / pol.dd, mu) * (Q.T @ v))
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-11 * np.linalg.norm(v))
    oracle.free_pol(pol)


def run_slice(oracle, A, intv, tol, v, maxit=None):
    pol = oracle.default_pol()
    pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
    assert oracle.find_pol(intv, pol) == 0
    if maxit is None:
        maxit = min(A.n, 600)
    out = oracle.ChebLanNr(intv, maxit, tol, v, pol)
    oracle.free_pol(pol)
    return out


def test_cheblannr_laplacian_closed_form(oracle, libc):
    A = M.laplacian(30, 30, 1)
    oracle.set_csr(A)
    libc.srand(1)
    v = oracle.rand_double(A.n)
    intv = np.array([0.4, 0.8, 0.0, 8.0])
    err, lam, Y, res = run_slice(oracle, A, intv, 1e-8, v)
    exact = M.laplacian_eigs(30, 30, 1, 0.4, 0.8)
    assert err == 0 and len(lam) == len(exact)
    np.testing.assert_allclose(np.sort(lam), exact, rtol=1e-10, atol=0)
    assert np.all(res < 1e-8)
    assert np.abs(Y.T @ Y - np.eye(len(lam))).max() < 1e-8


@pytest.mark.parametrize("kind", ["perturbed3d", "anderson"])
def test_cheblannr_matches_dense(oracle, libc, kind):
    if kind == "perturbed3d":
        A = M.perturb_diag(M.laplacian(10, 10, 10), 1e-3, seed=3)
        lo, hi = 0.0, 12.0
        intv = np.array([1.0, 1.3, lo, hi])
    else:
        A = M.anderson(8, 16.5, seed=4)
        lo, hi = M.gershgorin(A)
        intv = np.array([-0.5, 0.5, lo, hi])
    oracle.set_csr(A)
    libc.srand(5)
    v = oracle.rand_double(A.n)
    err, lam, Y, res = run_slice(oracle, A, intv, 1e-8, v)
    ev = np.linalg.eigvalsh(A.to_scipy().toarray())
    exact = ev[(ev >= intv[0]) & (ev <= intv[1])]
    assert err == 0 and len(lam) == len(exact)
    np.testing.assert_allclose(np.sort(lam), exact, rtol=1e-10, atol=1e-12)
    assert np.all(res < 1e-8)
    R = A.to_scipy() @ Y - Y * lam
    np.testing.assert_allclose(np.linalg.norm(R, axis=0), res, rtol=1e-6, atol=1e-14)


def test_cheblannr_error_paths(oracle):
    A = M.laplacian(6, 6, 1)
    oracle.set_csr(A)
    pol = oracle.default_pol()
    assert oracle.find_pol(np.array([0.4, 0.8, 0.0, 8.0]), pol) == 0
    v = np.ones(A.n)
    err, lam, Y, res = oracle.ChebLanNr(np.array([0.8, 0.4, 0.0, 8.0]), 30, 1e-8, v, pol)
    assert err == 0 and len(lam) == 0                 # invalid interval: nothing, code 0
    pol.gam = 1.5
    assert oracle.ChebLanNr(np.array([0.4, 0.8, 0.0, 8.0]), 30, 1e-8, v, pol)[0] == -1
    oracle.free_pol(pol)


def test_cheblanTr_laplacian_closed_form(oracle, libc):
    A = M.laplacian(25, 25, 1)
    oracle.set_csr(A)
    libc.srand(8)
    v = oracle.rand_double(A.n)
    intv = np.array([0.5, 0.9, 0.0, 8.0])
    exact = M.laplacian_eigs(25, 25, 1, 0.5, 0.9)
    pol = oracle.default_pol()
    assert oracle.find_pol(intv, pol) == 0
    nev = len(exact) + 2
    err, lam, Y, res = oracle.ChebLanTr(max(5 * nev, 100), nev, intv, 2000, 1e-8, v, pol)
    assert err == 0 and len(lam) == len(exact)
    np.testing.assert_allclose(np.sort(lam), exact, rtol=1e-10)
    assert np.all(res < 1e-8)
    oracle.free_pol(pol)


def test_lanbounds_encloses_spectrum(oracle, libc):
    A = M.laplacian(40, 40, 1)
    oracle.set_csr(A)
    ex = M.laplacian_eigs(40, 40, 1)
    libc.srand(3)
    v = oracle.rand_double(A.n)
    err, lmin, lmax = oracle.LanBounds(80, v)
    assert err == 0
    assert abs(lmin - ex[0]) < 0.01 * (ex[-1] - ex[0])
    assert abs(lmax - ex[-1]) < 0.01 * (ex[-1] - ex[0])
    # with a full Krylov space the Ritz values are the spectrum
    B = M.random_symmetric(40, 3, 6, seed=1)
    oracle.set_csr(B)
    err, lo, hi = oracle.LanBounds(40, np.random.default_rng(0).standard_normal(40))
    evb = np.linalg.eigvalsh(B.to_scipy().toarray())
    assert lo <= evb[0] + 1e-9 and hi >= evb[-1] - 1e-9


@pytest.mark.parametrize("bndtype", [1, 2])
def test_lantrbounds_encloses_and_converges(oracle, libc, bndtype):
    """Thick-restart bounds (SRC/lanTrbounds.c): the returned interval
    encloses the spectrum (up to rounding of the converged Ritz values) and
    its ends lie within the stopping tolerance tol*(|lmin|+|lmax|) of the
    extreme eigenvalues, for the closed-form 2-D/3-D Laplacian spectra and
    dense spectra of an Anderson and a random matrix; the driver call of
    TESTS/PLanN/MMPLanN.c:132 (lanm 50, maxit 200, tol 1e-8)."""
    cases = [
        (M.laplacian(40, 40, 1), M.laplacian_eigs(40, 40, 1)),
        (M.laplacian(12, 12, 12), M.laplacian_eigs(12, 12, 12)),
        (M.anderson(10, 16.5, seed=4), None),
        (M.random_symmetric(2000, 8, 32, seed=9), None),
    ]
    tol = 1e-8
    for A, ex in cases:
        if ex is None:
            ex = np.linalg.eigvalsh(A.to_scipy().toarray())
        oracle.set_csr(A)
        libc.srand(3)
        v = oracle.rand_double(A.n)
        err, lo, hi = oracle.LanTrbounds(50, 200, tol, v, bndtype)
        assert err == 0
        scale = abs(ex[0]) + abs(ex[-1])
        assert lo <= ex[0] + 1e-12 * scale and hi >= ex[-1] - 1e-12 * scale
        assert ex[0] - lo < 2 * tol * scale and hi - ex[-1] < 2 * tol * scale


def test_chebsi_closed_form_and_dense(oracle, libc):
    """Filtered subspace iteration (SRC/chebsi.c) with the PSI driver's
    filter settings (TESTS/PSI/LapPSI.c:150-168: damping 0, thresh_int 0.5,
    nev = 1.2 (1 + count)): every eigenvalue in the interval, against the
    closed-form 2-D Laplacian spectrum and a dense spectrum."""
    cases = [(M.laplacian(30, 30, 1), np.array([0.4, 0.8, 0.0, 8.0]), None),
             (M.perturb_diag(M.laplacian(9, 9, 9), 1e-3, seed=2), np.array([1.0, 1.4, 0.0, 12.001]),
              None)]
    for A, intv, _ in cases:
        oracle.set_csr(A)
        d = np.linalg.eigvalsh(A.to_scipy().toarray())
        exact = d[(d >= intv[0]) & (d <= intv[1])]
        nev = int(1.2 * (1 + len(exact)))
        pol = oracle.default_pol()
        pol.damping, pol.thresh_int = 0, 0.5
        assert oracle.find_pol(intv, pol) == 0
        libc.srand(4)
        V0 = oracle.rand_double(A.n * nev).reshape(nev, A.n).T
        err, lam, Y, res = oracle.ChebSI(nev, intv, 1000, 1e-8, V0, pol)
        oracle.free_pol(pol)
        assert err == 0
        assert len(lam) == len(exact)
        np.testing.assert_allclose(np.sort(lam), exact, rtol=1e-10, atol=1e-12)
        assert np.all(res < 1e-8)
        R = A.to_scipy() @ Y - Y * lam
        np.testing.assert_allclose(np.linalg.norm(R, axis=0), res, rtol=1e-5, atol=1e-13)
