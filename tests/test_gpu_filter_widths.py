"""GPU: the blocked filter (evsl_amd_cheb_av_cols, the block form of ChebAv,
SRC/chebpoly.c:519-598) at every block width the fused filter kernel is
instantiated for -- 1..8, 16 and 32 columns, the padded widths between them
and a run of two chunks -- with the 16-bit and the 32-bit column index.

Bars:
* per-column results equal bit for bit at every width, in FP64 and in the
  mixed form: a column's arithmetic in the kernel (its row sums, recurrence
  and y updates) does not depend on the block width or the lane layout, which
  the degree-phased narrowing of the batched solver relies on;
* FP64 against the oracle's ChebAv, and mixed (FP32 filter on the shifted
  matrix) against FP64: each column within 8 * eps * sum_k k^2 |mu_k| * ||x||
  (eps of the filter's arithmetic), a first-order bound of the three-term
  recurrence -- a rounding of relative size c * eps at step j reaches T_k
  through U_{k-j}, of norm <= k - j + 1 on [-1, 1], and summing over j <= k
  gives c * eps * k^2 / 2 per term of the series."""
import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import PREC_FP64, PREC_MIXED

pytestmark = pytest.mark.gpu

WIDTHS = [1, 2, 3, 4, 5, 6, 7, 8, 11, 16, 23, 32, 40]


def make_case(name):
    if name in ("lap3d_i16", "lap3d_pattern"):
        # every |ja - row| <= 256: 16-bit column offsets, or with pattern
        # codes one byte per nonzero (7 (offset, value) pairs)
        A = M.perturb_diag(M.laplacian(16, 16, 16), 1e-3, seed=9)
        lo, hi = 0.0, 12.001
        ivs = [(0.4, 0.7), (1.5, 1.9), (5.0, 6.0)]
        ibytes = 2 if name == "lap3d_i16" else 1
    else:  # columns anywhere in [0, n): 32-bit column indices
        A = M.random_symmetric(40000, 8, 32, seed=6)
        lo, hi = M.gershgorin(A)
        ivs = [(-0.3, 0.3), (1.0, 2.0), (4.0, 6.0)]
        ibytes = 4
    return A, lo, hi, ivs, ibytes


def make_pols(lib, ivs, lo, hi):
    out = []
    for a, b in ivs:
        pol = lib.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
        assert lib.find_pol(np.array([a, b, lo, hi]), pol) == 0
        out.append(pol)
    return out


def mu_of(pol):
    return np.array([pol.mu[k] for k in range(pol.deg + 1)])


@pytest.mark.parametrize("case", ["lap3d_i16", "lap3d_pattern", "randsym_i32"])
def test_filter_every_width(ev, oracle, case):
    A, lo, hi, ivs, ibytes = make_case(case)
    ev.set_pattern_codes(case != "lap3d_i16")
    try:
        assert ev.set_matrix(A) == 0, ev.last_error()
    finally:
        ev.set_pattern_codes(1)
    assert ev.lib.evsl_amd_spmm_index_bytes() == ibytes
    oracle.set_csr(A)
    pg, po = make_pols(ev, ivs, lo, hi), make_pols(oracle, ivs, lo, hi)
    degs = [p.deg for p in pg]
    assert degs == [p.deg for p in po] and len(set(degs)) == 3, degs
    assert min(degs) >= 3
    print(f"  {case}: n {A.n}, degrees {degs}", flush=True)
    nmax = max(WIDTHS)
    X = np.random.default_rng(21).uniform(-1.0, 1.0, (A.n, nmax))
    # column j uses filter j % 3 (mixed degrees in every block of >= 2 columns)
    which = [j % 3 for j in range(nmax)]
    ref = np.column_stack([oracle.ChebAv(po[which[j]], X[:, j]) for j in range(nmax)])
    xn = np.linalg.norm(X, axis=0)
    k2 = np.array([(np.arange(pg[which[j]].deg + 1) ** 2 * np.abs(mu_of(pg[which[j]]))).sum()
                   for j in range(nmax)])
    single = {}
    for prec in (PREC_FP64, PREC_MIXED):
        # each column on its own (width 1): the width-invariance reference
        Y1 = np.empty((A.n, nmax))
        for j in range(nmax):
            err, y = ev.cheb_av_cols([pg[which[j]]], X[:, j:j + 1], prec)
            assert err == 0, ev.last_error()
            Y1[:, j] = y[:, 0]
        single[prec] = Y1
        for w in WIDTHS:
            err, Y = ev.cheb_av_cols([pg[which[j]] for j in range(w)], X[:, :w], prec)
            assert err == 0, (w, prec, ev.last_error())
            assert np.array_equal(Y, Y1[:, :w]), (case, prec, w, np.abs(Y - Y1[:, :w]).max())
    e64 = np.linalg.norm(single[PREC_FP64] - ref, axis=0)
    b64 = 8 * np.finfo(np.float64).eps * k2 * xn
    emx = np.linalg.norm(single[PREC_MIXED] - single[PREC_FP64], axis=0)
    bmx = 8 * np.finfo(np.float32).eps * k2 * xn
    print(f"  fp64 vs oracle: max {(e64 / b64).max():.2e} of its bound, "
          f"{(e64 / np.linalg.norm(ref, axis=0)).max():.2e} relative; mixed vs fp64: "
          f"max {(emx / bmx).max():.2e} of its bound", flush=True)
    assert np.all(e64 <= b64)
    assert np.all(emx <= bmx)
    # the filters pass their interval and damp the rest: a non-trivial output
    assert np.all(np.linalg.norm(ref, axis=0) > 1e-4 * xn)
    for p in pg:
        ev.free_pol(p)
    for p in po:
        oracle.free_pol(p)


def test_filter_cols_rejects_bad_arguments(ev):
    A = M.laplacian(20, 20, 1)
    assert ev.set_matrix(A) == 0
    p1, p2 = make_pols(ev, [(0.4, 0.8)], 0.0, 8.0)[0], make_pols(ev, [(0.4, 0.8)], 0.0, 9.0)[0]
    X = np.ones((A.n, 2))
    err, _ = ev.cheb_av_cols([p1, p2], X)  # different cc, dd
    assert err == -104
    err, _ = ev.cheb_av_cols([p1], X[:, :1], prec=7)
    assert err == -104
    ev.free_pol(p1)
    ev.free_pol(p2)


def test_filter_nontemporal_instances_bitwise(ev):
    """On the 88^3 Laplacian one launch's working set passes 320 MB at 32
    FP32 columns and at 16 and 32 FP64 columns, so those widths run the
    fused kernel's non-temporal instance; its columns must equal the
    one-column runs (plain instance) bit for bit."""
    A = M.perturb_diag(M.laplacian(88, 88, 88), 1e-3, seed=4)
    assert ev.set_matrix(A) == 0, ev.last_error()
    lib = ev.lib
    assert lib.evsl_amd_spmm_streaming(32, PREC_MIXED, 1) == 1
    assert lib.evsl_amd_spmm_streaming(8, PREC_MIXED, 1) == 0
    assert lib.evsl_amd_spmm_streaming(16, PREC_FP64, 1) == 1
    assert lib.evsl_amd_spmm_streaming(4, PREC_FP64, 1) == 0
    pg = make_pols(ev, [(0.4, 0.7), (1.5, 1.9), (5.0, 6.0)], 0.0, 12.001)
    ncol = 32
    X = np.random.default_rng(5).uniform(-1.0, 1.0, (A.n, ncol))
    which = [j % 3 for j in range(ncol)]
    for prec in (PREC_FP64, PREC_MIXED):
        Y1 = np.empty((A.n, ncol))
        for j in range(ncol):
            err, y = ev.cheb_av_cols([pg[which[j]]], X[:, j:j + 1], prec)
            assert err == 0, ev.last_error()
            Y1[:, j] = y[:, 0]
        for w in (16, 32):
            err, Y = ev.cheb_av_cols([pg[which[j]] for j in range(w)], X[:, :w], prec)
            assert err == 0, (w, prec, ev.last_error())
            assert np.array_equal(Y, Y1[:, :w]), (prec, w, np.abs(Y - Y1[:, :w]).max())
    for p in pg:
        ev.free_pol(p)
