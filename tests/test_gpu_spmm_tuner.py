"""GPU: the per-matrix layout tuner of the fused filter kernel
(evsl_amd_spmm_layout, evsl_amd_set_spmm_layout, evsl_amd_set_spmm_tuning).

Bars:
* every candidate layout of every tuned key gives the fixed policy's bits:
  the ChebAv step through evsl_amd_cheb_av_cols (the block form of ChebAv,
  SRC/chebpoly.c:519-598) at 2, 4, 8, 16 and 32 columns in FP64 and in the mixed
  form, with 16-bit and 32-bit column indices; the residual step through the
  mixed batched solver's refinement with every residual key forced to each
  candidate in turn;
* the tuner settles on one of its candidates within a few filter calls, and
  the calls made while it measures equal calls with tuning off, bit for bit;
* the C-ABI takes exactly the candidate layouts and rejects other widths,
  modes and layouts."""
import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import PREC_FP64, PREC_MIXED

pytestmark = pytest.mark.gpu

ERR_ARG = -104
WIDTHS = (2, 4, 8, 16, 32)
# candidates (10*V + R) per precision and width, as the kernel library lists them
CANDIDATES = {
    PREC_FP64: {2: [11, 12, 21], 4: [11, 12, 14, 21, 22], 8: [11, 12, 21, 22, 24],
                16: [11, 12, 21, 22, 24], 32: [11, 12, 21, 22, 24]},
    PREC_MIXED: {2: [11, 12, 21], 4: [11, 12, 14, 21, 22, 41], 8: [11, 21, 22, 24, 41, 42],
                 16: [11, 21, 22, 41, 42, 44], 32: [11, 21, 22, 41, 42, 44]},
}


def make_pols(ev, ivs, lo, hi, max_deg=3000):
    out = []
    for a, b in ivs:
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, max_deg
        assert ev.find_pol(np.array([a, b, lo, hi]), pol) == 0
        out.append(pol)
    return out


def reset_keys(ev):
    for prec in (PREC_FP64, PREC_MIXED):
        for B in WIDTHS:
            for mode in (1, 3):
                assert ev.set_spmm_layout(B, prec, mode, -1) == 0


def test_tuner_candidates_and_argument_errors(ev):
    A = M.laplacian(20, 20, 20)
    assert ev.set_matrix(A) == 0, ev.last_error()
    for prec in (PREC_FP64, PREC_MIXED):
        for B in WIDTHS:
            for mode in (1, 3):
                ok = [10 * v + r for v in (1, 2, 4, 8) for r in (1, 2, 4, 8)
                      if ev.set_spmm_layout(B, prec, mode, 10 * v + r) == 0]
                assert ok == CANDIDATES[prec][B], (prec, B, mode, ok)
                assert ev.set_spmm_layout(B, prec, mode, 0) == 0
                lay, dec = ev.spmm_layout(B, prec, mode)
                assert dec and lay in ok  # the fixed policy is one of the candidates
    for B in (1, 3, 6, 7):
        assert ev.set_spmm_layout(B, PREC_FP64, 1, 11) == ERR_ARG
    assert ev.set_spmm_layout(8, PREC_FP64, 2, 11) == ERR_ARG  # KPM keeps its layout
    assert ev.set_spmm_layout(8, PREC_FP64, 1, -2) == ERR_ARG
    assert ev.spmm_layout(8, PREC_FP64, 2)[1]
    assert ev.spmm_layout(8, PREC_FP64, 4)[0] == -1
    reset_keys(ev)


@pytest.mark.parametrize("case", ["lap3d_i16", "lap3d_pattern", "anderson_periodic",
                                  "randsym_i32"])
def test_every_layout_same_bits_chebav(ev, case):
    if case in ("lap3d_i16", "lap3d_pattern"):
        A = M.perturb_diag(M.laplacian(24, 24, 24), 1e-3, seed=2)
        lo, hi = 0.0, 12.001
        ivs = [(0.4, 0.7), (1.5, 1.9), (5.0, 6.0)]
    elif case == "anderson_periodic":  # wrap-around neighbours far from the row
        A = M.anderson(18, 16.5, seed=6)
        lo, hi = M.gershgorin(A)
        ivs = [(-0.5, 0.5), (2.0, 3.0), (-6.0, -4.0)]
    else:
        A = M.random_symmetric(20000, 8, 32, seed=12)
        lo, hi = M.gershgorin(A)
        ivs = [(-0.3, 0.3), (1.0, 2.0), (4.0, 6.0)]
    ev.set_pattern_codes(case != "lap3d_i16")
    try:
        assert ev.set_matrix(A) == 0, ev.last_error()
    finally:
        ev.set_pattern_codes(1)
    # the stencil and lattice matrices run pattern codes unless switched off
    assert (ev.pattern_codes() > 0) == (case in ("lap3d_pattern", "anderson_periodic"))
    pg = make_pols(ev, ivs, lo, hi)
    X = np.random.default_rng(3).uniform(-1.0, 1.0, (A.n, 32))
    # one polynomial for all columns: every launch of a call runs at width B
    for prec in (PREC_FP64, PREC_MIXED):
        ev.set_spmm_tuning(False)
        ref = {}
        for B in WIDTHS:
            err, ref[B] = ev.cheb_av_cols([pg[B % 3]] * B, X[:, :B], prec)
            assert err == 0, ev.last_error()
        ev.set_spmm_tuning(True)
        for B in WIDTHS:
            for lay in CANDIDATES[prec][B]:
                assert ev.set_spmm_layout(B, prec, 1, lay) == 0
                assert ev.spmm_layout(B, prec, 1) == (lay, True)
                err, Y = ev.cheb_av_cols([pg[B % 3]] * B, X[:, :B], prec)
                assert err == 0, ev.last_error()
                assert np.array_equal(Y, ref[B]), (case, prec, B, lay,
                                                   np.abs(Y - ref[B]).max())
            assert ev.set_spmm_layout(B, prec, 1, -1) == 0
    for p in pg:
        ev.free_pol(p)


def test_every_layout_same_bits_residual_step(ev):
    """The mixed batched solver's refinement runs the residual step at 32
    columns and at the padded width of the last chunk; each residual key is
    forced to its i-th candidate in turn."""
    A = M.laplacian(100, 100, 1)
    lo, hi = 0.0, 8.0
    assert ev.set_matrix(A) == 0, ev.last_error()
    ivs = [np.array([0.4, 0.6, lo, hi]), np.array([0.6, 0.8, lo, hi])]
    pols = make_pols(ev, [(0.4, 0.6), (0.6, 0.8)], lo, hi)
    v0 = np.random.default_rng(21).uniform(-1, 1, A.n)
    ev.set_precision(PREC_MIXED)
    # both slices on the FP32 filter (and so through the refinement) whatever
    # their degree
    assert ev.lib.evsl_amd_set_mixed_degree_range(2, 20000) == 0
    try:
        ev.set_spmm_tuning(False)
        err, ref = ev.ChebLanNr_batch(ivs, 900, 1e-8, v0, pols)
        assert err == 0, ev.last_error()
        assert all(len(r[0]) > 32 for r in ref)  # more than one refinement chunk
        ev.set_spmm_tuning(True)
        for i in range(6):
            for B in WIDTHS:
                cand = CANDIDATES[PREC_MIXED][B]
                assert ev.set_spmm_layout(B, PREC_MIXED, 3, cand[i % len(cand)]) == 0
            err, got = ev.ChebLanNr_batch(ivs, 900, 1e-8, v0, pols)
            assert err == 0, ev.last_error()
            for (l0, Y0, r0), (l1, Y1, r1) in zip(ref, got):
                assert np.array_equal(l0, l1) and np.array_equal(r0, r1), i
                assert np.array_equal(Y0, Y1), i
    finally:
        ev.set_precision(PREC_FP64)
        ev.lib.evsl_amd_set_mixed_degree_range(32, 20000)
        ev.set_spmm_tuning(True)
        reset_keys(ev)
    for p in pols:
        ev.free_pol(p)


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_tuner_settles_and_measuring_calls_keep_bits(ev, prec):
    A = M.random_symmetric(40000, 8, 32, seed=13)
    lo, hi = M.gershgorin(A)
    assert ev.set_matrix(A) == 0, ev.last_error()
    (pol,) = make_pols(ev, [(-0.3, 0.3)], lo, hi)
    assert pol.deg >= 20
    X = np.random.default_rng(4).uniform(-1.0, 1.0, (A.n, 8))
    ev.set_spmm_tuning(False)
    err, ref = ev.cheb_av_cols([pol] * 8, X, prec)
    assert err == 0
    ev.set_spmm_tuning(True)
    assert ev.spmm_layout(8, prec, 1)[1] is False  # a new matrix starts measuring
    calls = 0
    while not ev.spmm_layout(8, prec, 1)[1]:
        assert calls < 30, "the tuner did not settle"
        err, Y = ev.cheb_av_cols([pol] * 8, X, prec)
        assert err == 0, ev.last_error()
        assert np.array_equal(Y, ref), calls
        calls += 1
    lay, _ = ev.spmm_layout(8, prec, 1)
    print(f"  settled after {calls} calls of {pol.deg} launches: layout {lay}", flush=True)
    assert lay in CANDIDATES[prec][8]
    err, Y = ev.cheb_av_cols([pol] * 8, X, prec)
    assert err == 0 and np.array_equal(Y, ref)
    ev.free_pol(pol)


def arrow_matrix(n, dense_rows, seed):
    """A random symmetric sparse matrix plus `dense_rows` rows (and columns)
    that touch every column: rows far longer than one row block's 2048
    nonzeros, which the filter kernel splits over the lanes of a block."""
    import scipy.sparse as sp
    from evsl_amd.types import HostCsr
    base = M.random_symmetric(n, 4, 8, seed=seed).to_scipy()
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(dense_rows), n)
    cols = np.tile(np.arange(n), dense_rows)
    vals = rng.standard_normal(rows.size) * 0.05
    D = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    A = (base + D + D.T).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return HostCsr.from_scipy(A)


@pytest.mark.parametrize("n", [6000, 40000])
def test_long_rows_same_bits_every_layout_and_width(ev, n):
    """Rows of n nonzeros (longer than a row block's 2048; 16-bit column
    offsets at n = 6000, 32-bit indices at 40000): every width and every
    candidate layout gives each column the bits of its one-column run."""
    A = arrow_matrix(n, 3 if n < 10000 else 2, seed=5)
    assert np.diff(A.ia).max() > 2 * 2048
    assert ev.set_matrix(A) == 0, ev.last_error()
    assert ev.lib.evsl_amd_spmm_index_bytes() == (2 if n < 32768 else 4)
    lo, hi = M.gershgorin(A)
    # the degree is capped at 79 (find_pol would reach 1325 and 2999 here):
    # the bits are compared launch for launch, and a long row's block walks
    # its stripes one gather at a time, ~1 ms per launch at n = 40000
    (pol,) = make_pols(ev, [(-0.2, 0.4)], lo, hi, max_deg=80)
    assert pol.deg == 79, pol.deg
    X = np.random.default_rng(8).uniform(-1.0, 1.0, (A.n, 32))
    for prec in (PREC_FP64, PREC_MIXED):
        ev.set_spmm_tuning(False)
        one = np.empty_like(X)
        for c in range(32):
            err, Yc = ev.cheb_av_cols([pol], X[:, c:c + 1], prec)
            assert err == 0, ev.last_error()
            one[:, c] = Yc[:, 0]
        for B in (1, 3, 6, 7):
            err, Y = ev.cheb_av_cols([pol] * B, X[:, :B], prec)
            assert err == 0 and np.array_equal(Y, one[:, :B]), (prec, B)
        ev.set_spmm_tuning(True)
        for B in WIDTHS:
            for lay in CANDIDATES[prec][B]:
                assert ev.set_spmm_layout(B, prec, 1, lay) == 0
                err, Y = ev.cheb_av_cols([pol] * B, X[:, :B], prec)
                assert err == 0, ev.last_error()
                assert np.array_equal(Y, one[:, :B]), (prec, B, lay, np.abs(Y - one[:, :B]).max())
            assert ev.set_spmm_layout(B, prec, 1, -1) == 0
    ev.free_pol(pol)
