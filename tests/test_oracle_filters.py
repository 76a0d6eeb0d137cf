"""Pins of the oracle's filter construction (SRC/chebpoly.c) by closed forms
and by an independent re-derivation of the degree search."""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as npc
from scipy.optimize import brentq


def jackson_textbook(m):
    """Jackson kernel g_k with N = m + 1 (Weisse et al., Rev. Mod. Phys. 78, 275, eq. 71)."""
    N = m + 1
    k = np.arange(1, m + 1)
    q = np.pi / (N + 1)
    return ((N - k + 1) * np.cos(q * k) + np.sin(q * k) / np.tan(q)) / (N + 1)


@pytest.mark.parametrize("m", [5, 40, 300])
def test_dampcf_closed_forms(oracle, m):
    j0 = oracle.dampcf(m, 0)
    assert j0[0] == 0.5 and np.all(j0[1:] == 1.0)
    j1 = oracle.dampcf(m, 1)
    assert j1[0] == 0.5
    np.testing.assert_allclose(j1[1:], jackson_textbook(m), rtol=0, atol=1e-13)
    j2 = oracle.dampcf(m, 2)
    k = np.arange(1, m + 1)
    np.testing.assert_allclose(j2[1:], np.sinc(k / (m + 1.0)), rtol=0, atol=1e-14)


def test_chebxPltd_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    for m in (0, 1, 7, 120):
        mu = rng.standard_normal(m + 1)
        x = np.cos(rng.uniform(0, np.pi, 50))
        np.testing.assert_allclose(oracle.chebxPltd(mu, x), npc.chebval(x, mu),
                                   rtol=0, atol=1e-12 * max(1, np.abs(mu).sum()))


def _T(k, x):
    x = np.asarray(x, dtype=float)
    out = np.empty_like(x)
    inside = np.abs(x) <= 1
    out[inside] = np.cos(k * np.arccos(x[inside]))
    xo = x[~inside]
    out[~inside] = np.sign(xo) ** k * np.cosh(k * np.arccosh(np.abs(xo)))
    return out


@pytest.mark.parametrize("side", ["left", "right"])
def test_find_pol_end_interval_is_scaled_chebyshev(oracle, side):
    lmin, lmax = 0.0, 8.0
    intv = np.array([lmin, 0.6, lmin, lmax]) if side == "left" else np.array([7.1, lmax, lmin, lmax])
    pol = oracle.default_pol()
    assert oracle.find_pol(intv, pol) == 0
    assert pol.type == (1 if side == "left" else 2)
    cc, dd = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    assert pol.cc == cc and pol.dd == dd
    if side == "left":
        bIn = (intv[1] - cc) / dd
        x, gam = bIn, -1.0
        a, b = x + 0.1 * np.sqrt((bIn + 1.0) * 0.5), 1.0
    else:
        aIn = (intv[0] - cc) / dd
        x, gam = aIn, 1.0
        a, b = -1.0, x - 0.1 * np.sqrt((1.0 - aIn) * 0.5)
    e, c = (b - a) / 2, (b + a) / 2
    m = pol.deg
    mu = pol.coefficients()
    p = lambda t: _T(m, (np.asarray(t) - c) / e) / _T(m, np.array([(gam - c) / e]))[0]
    ts = np.linspace(-1, 1, 41)
    np.testing.assert_allclose(npc.chebval(ts, mu), p(ts), rtol=1e-9, atol=1e-11)
    assert pol.gam == gam
    assert abs(pol.bar - p(np.array([x]))[0]) < 1e-10
    # minimal degree: thresh_ext crossed exactly at deg
    q = lambda k: _T(k, np.array([(x - c) / e]))[0] / _T(k, np.array([(gam - c) / e]))[0]
    assert q(m) < pol.thresh_ext <= q(m - 1)


def _balanced_filter(m, tha, thb, damping):
    """Independent re-derivation of one degree of the middle-interval search:
    damped delta expansion centred at the angle that equalises p(a), p(b)."""
    k = np.arange(m + 1)
    if damping == 2:
        jac = np.ones(m + 1)
        jac[1:] = np.sinc(k[1:] / (m + 1.0))
    else:
        raise NotImplementedError
    jac[0] = 0.5
    v = np.cos(k * thb) - np.cos(k * tha)
    f = lambda th: np.sum(v * jac * np.cos(k * th))
    fa, fb = f(tha), f(thb)
    if fa > 0 or fb < 0:
        return None
    thc = brentq(f, thb, tha, xtol=1e-15, rtol=1e-15) if fa * fb < 0 else 0.5 * (tha + thb)
    mu = np.cos(k * thc) * jac
    gam = np.cos(thc)
    t = npc.chebval(gam, mu)
    return mu / t, gam, npc.chebval(np.cos(tha), mu) / t, npc.chebval(np.cos(thb), mu) / t


@pytest.mark.parametrize("ab", [(0.4, 0.8), (2.0, 2.3), (0.2, 0.3)])
def test_find_pol_middle_interval(oracle, ab):
    lmin, lmax = 0.0, 8.0 if ab[1] < 1 else 12.0
    intv = np.array([ab[0], ab[1], lmin, lmax])
    pol = oracle.default_pol()
    assert oracle.find_pol(intv, pol) == 0
    assert pol.type == 0
    m = pol.deg
    mu = pol.coefficients()
    cc, dd = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    at, bt = (ab[0] - cc) / dd, (ab[1] - cc) / dd
    # scaled so that p(gam) = 1, shape = damped delta expansion at gam
    assert abs(npc.chebval(pol.gam, mu) - 1.0) < 1e-12
    k = np.arange(m + 1)
    jac = np.ones(m + 1)
    jac[1:] = np.sinc(k[1:] / (m + 1.0))
    jac[0] = 0.5
    shape = np.cos(k * np.arccos(pol.gam)) * jac
    np.testing.assert_allclose(mu, shape / npc.chebval(pol.gam, shape), rtol=0, atol=1e-12)
    pa, pb = npc.chebval(at, mu), npc.chebval(bt, mu)
    assert abs(pa - pb) < 1e-8                      # balanced
    assert abs(pol.bar - min(pa, pb)) < 1e-12       # bar = min(p(a), p(b))
    assert max(pa, pb) <= pol.thresh_int + 1e-12    # accepted
    # no smaller degree of the search passes the same test
    tha, thb = np.arccos(at), np.arccos(bt)
    min_deg = max(int(2 + 0.5 / (bt - at)), 2)
    for mm in range(min_deg, m):
        r = _balanced_filter(mm, tha, thb, 2)
        if r is None:
            continue
        _, _, qa, qb = r
        assert not (qa <= pol.thresh_int and qb <= pol.thresh_int), mm
    r = _balanced_filter(m, tha, thb, 2)
    np.testing.assert_allclose(r[0], mu, rtol=0, atol=1e-9)
    oracle.free_pol(pol)


def test_find_pol_invalid_interval(oracle):
    pol = oracle.default_pol()
    assert oracle.find_pol(np.array([0.8, 0.4, 0.0, 8.0]), pol) == -1
    assert oracle.find_pol(np.array([9.0, 10.0, 0.0, 8.0]), pol) == -1
