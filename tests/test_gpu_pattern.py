"""GPU: pattern codes (evsl_amd_set_pattern_codes, evsl_amd_pattern_codes).

A matrix whose off-diagonal entries take at most 255 distinct (column
offset, value) pairs is stored for the filter kernels as one byte per
nonzero; the kernels decode it while staging a row block, so every product
and every sum is the one the CSR-indexed kernels form.

Bars:
* the matrices of the stencil and lattice families are pattern-coded, with
  16-bit table offsets (3-D Laplacian, a small periodic Anderson lattice, a
  2-D Laplacian with rows that store no diagonal entry) or 32-bit ones (a
  periodic Anderson lattice whose wrap-around neighbours lie more than
  32767 rows away); a random symmetric matrix is not;
* with pattern codes on and off, the blocked ChebAv (the block form of
  SRC/chebpoly.c:519-598) gives the same bits at every width in FP64 and in
  the mixed form (shifted FP32 copy), and so do the FP64 KPM moments
  (SRC/spslice.c:29-148) and a mixed batched solve (eigenvalues, residuals,
  eigenvectors)."""
import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import PREC_FP64, PREC_MIXED, VEC_RADEMACHER

pytestmark = pytest.mark.gpu


def make_pol(ev, a, b, lo, hi, max_deg=200):
    pol = ev.default_pol()
    pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, max_deg
    assert ev.find_pol(np.array([a, b, lo, hi]), pol) == 0
    return pol


def case_matrix(case):
    if case == "lap2d_gaps":
        # a 2-D Laplacian whose every third row stores no diagonal entry: the
        # FP32 path then runs on the unshifted values and their tables
        import scipy.sparse as sp
        from evsl_amd.types import HostCsr
        S = M.laplacian(60, 60, 1).to_scipy().tolil()
        for r in range(0, S.shape[0], 3):
            S[r, r] = 0.0
        S = sp.csr_matrix(S)
        S.eliminate_zeros()
        return HostCsr.from_scipy(S), 2
    if case == "lap3d":
        return M.perturb_diag(M.laplacian(30, 30, 30), 1e-3, seed=3), 2
    if case == "anderson_short":  # wrap-around offsets 24 * 23 * 24 < 32768
        return M.anderson(24, 16.5, seed=9), 2
    if case == "anderson_long":  # wrap-around offsets 40 * 39 * 40 > 32767
        return M.anderson(40, 16.5, seed=9), 3
    return M.random_symmetric(20000, 8, 32, seed=11), None


@pytest.mark.parametrize("case", ["lap3d", "lap2d_gaps", "anderson_short", "anderson_long",
                                  "randsym"])
def test_pattern_codes_same_bits(ev, case):
    A, mode = case_matrix(case)
    lo, hi = M.gershgorin(A)
    a, b = lo + 0.45 * (hi - lo), lo + 0.5 * (hi - lo)
    X = np.random.default_rng(5).uniform(-1.0, 1.0, (A.n, 32))
    out = {}
    try:
        for on in (0, 1):
            ev.set_pattern_codes(on)
            assert ev.set_matrix(A) == 0, ev.last_error()
            if on and mode is not None:
                assert 0 < ev.pattern_codes() <= 255
                assert ev.lib.evsl_amd_spmm_index_bytes() == 1
                # pattern codes at 1-2 FP32 and 1 FP64 column, indices beyond
                assert ev.lib.evsl_amd_spmm_index_mode(1, PREC_FP64, 1) == mode
                assert ev.lib.evsl_amd_spmm_index_mode(2, PREC_MIXED, 1) == mode
                assert ev.lib.evsl_amd_spmm_index_mode(2, PREC_FP64, 1) in (0, 1)
                assert ev.lib.evsl_amd_spmm_index_mode(4, PREC_MIXED, 1) in (0, 1)
            else:
                assert ev.pattern_codes() == 0
                assert ev.lib.evsl_amd_spmm_index_mode(1, PREC_FP64, 1) in (0, 1)
            pol = make_pol(ev, a, b, lo, hi)
            for prec in (PREC_FP64, PREC_MIXED):
                for B in (1, 2, 3, 4, 7, 8, 16, 32):
                    err, Y = ev.cheb_av_cols([pol] * B, X[:, :B], prec)
                    assert err == 0, ev.last_error()
                    out[(on, prec, B)] = Y
            err, mu, ecnt = ev.kpmdos(60, 1, 8, np.array([a, b, lo, hi]),
                                      vec_kind=VEC_RADEMACHER, seed=4)
            assert err == 0
            out[(on, "kpm")] = (mu, ecnt)
            ev.free_pol(pol)
    finally:
        ev.set_pattern_codes(1)
    for key in [k for k in out if k[0] == 0]:
        k1 = (1,) + key[1:]
        if key[1] == "kpm":
            assert np.array_equal(out[key][0], out[k1][0]) and out[key][1] == out[k1][1], case
        else:
            assert np.array_equal(out[key], out[k1]), (case, key)


def test_pattern_codes_same_solve(ev):
    """A mixed batched solve of two slices of the perturbed 3-D Laplacian
    (FP32 filter on the shifted pattern tables, FP64 refinement) gives the
    same bits with pattern codes on and off."""
    A = M.perturb_diag(M.laplacian(24, 24, 24), 1e-3, seed=8)
    lo, hi = M.gershgorin(A)
    v0 = np.random.default_rng(2).uniform(-1.0, 1.0, A.n)
    res = {}
    try:
        for on in (0, 1):
            ev.set_pattern_codes(on)
            assert ev.set_matrix(A) == 0
            assert (ev.pattern_codes() > 0) == bool(on)
            ivs = [np.array([0.40, 0.46, lo, hi]), np.array([0.46, 0.52, lo, hi])]
            pols = [make_pol(ev, iv[0], iv[1], lo, hi, max_deg=3000) for iv in ivs]
            ev.set_precision(PREC_MIXED)
            err, outs = ev.ChebLanNr_batch(ivs, 400, 1e-8, v0, pols)
            assert err == 0, ev.last_error()
            res[on] = outs
            for p in pols:
                ev.free_pol(p)
    finally:
        ev.set_precision(PREC_FP64)
        ev.set_pattern_codes(1)
    for (l0, y0, r0), (l1, y1, r1) in zip(res[0], res[1]):
        assert len(l0) > 0
        assert np.array_equal(l0, l1) and np.array_equal(r0, r1) and np.array_equal(y0, y1)
