"""The allocation-free filter entry point (evsl_amd_cheb_av_cols_noalloc)
against evsl_amd_cheb_av_cols, through the C ABI: the same columns bit for
bit, on two calls in a row (the second on a smaller matrix with fewer
columns), in FP64 and in the mixed form; and its refusals (host pointers, a
mixed call before the matrix's FP32 copy exists, a column larger than the
scratch)."""
import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import PREC_FP64, PREC_MIXED

pytestmark = pytest.mark.gpu

ERR_ARG = -104


def make_pols(ev, ivs, lo, hi):
    out = []
    for a, b in ivs:
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
        assert ev.find_pol(np.array([a, b, lo, hi]), pol) == 0
        out.append(pol)
    return out


def run_noalloc(ev, pols, X, prec):
    n, m = X.shape
    Xd = ev.to_device(np.asfortranarray(X).ravel(order="F"))
    Yd = ev.to_device(np.zeros(n * m))
    try:
        err = ev.cheb_av_cols_noalloc(pols, Xd, Yd, prec)
        Y = ev.from_device(Yd, n * m).reshape((n, m), order="F")
    finally:
        ev.free_device(Xd)
        ev.free_device(Yd)
    return err, Y


def test_noalloc_equals_cheb_av_cols_on_two_calls(ev):
    cases = [
        # (matrix, intervals, columns): the second call is the smaller one
        (M.perturb_diag(M.laplacian(34, 34, 34), 1e-3, seed=3),
         [(0.4, 0.7), (1.5, 1.9), (5.0, 6.0)], 11),
        (M.laplacian(100, 100, 1), [(0.5, 0.9)], 3),
    ]
    for call, (A, ivs, ncol) in enumerate(cases):
        lo, hi = M.gershgorin(A)
        assert ev.set_matrix(A) == 0, ev.last_error()
        pg = make_pols(ev, ivs, lo, hi)
        pols = [pg[j % len(pg)] for j in range(ncol)]
        X = np.random.default_rng(10 + call).uniform(-1.0, 1.0, (A.n, ncol))
        # a mixed call before any mixed call on this matrix: no FP32 copy yet
        err, _ = run_noalloc(ev, pols, X, PREC_MIXED)
        assert err == ERR_ARG, err
        for prec in (PREC_FP64, PREC_MIXED):
            err, Yref = ev.cheb_av_cols(pols, X, prec)
            assert err == 0, ev.last_error()
            err, Y = run_noalloc(ev, pols, X, prec)
            assert err == 0, ev.last_error()
            print(f"  call {call + 1}: n {A.n}, {ncol} columns, degrees "
                  f"{sorted({p.deg for p in pols})}, prec {prec}: max |diff| "
                  f"{np.abs(Y - Yref).max():.1e}", flush=True)
            assert np.array_equal(Y, Yref), (call, prec)
            assert np.linalg.norm(Yref) > 0
        # host pointers are refused
        P = (type(pols[0]) * ncol)(*pols)
        Xh = np.asfortranarray(X)
        Yh = np.zeros_like(Xh)
        err = ev.lib.evsl_amd_cheb_av_cols_noalloc(ncol, P, Xh.ctypes.data, Yh.ctypes.data,
                                                   PREC_FP64)
        assert err == ERR_ARG, err
        for p in pg:
            ev.free_pol(p)


def test_noalloc_refuses_a_column_beyond_the_scratch(ev):
    # FP64: one column needs n x 40 bytes; 1.7 million rows exceed 64 MiB
    A = M.laplacian(1_700_000, 1, 1)
    lo, hi = M.gershgorin(A)
    assert ev.set_matrix(A) == 0, ev.last_error()
    (pol,) = make_pols(ev, [(3.0, 3.2)], lo, hi)
    X = np.ones((A.n, 1))
    err, _ = run_noalloc(ev, [pol], X, PREC_FP64)
    assert err == ERR_ARG, err
    assert "scratch" in ev.last_error()
    ev.free_pol(pol)
