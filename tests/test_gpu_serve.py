"""GPU: the request loop over device-resident matrices (evsl_amd_server_*):
slice requests (ChebLanNr on an interval) and DOS requests (kpmdos_ex)
arriving while others run, admitted by projected memory, advanced together.

Bars:
* every request's result is the same alone and in the mix, bit for bit, in
  FP64 and in the mixed form (a slice's arithmetic does not depend on the
  slices sharing its filter passes, a DOS request's on the grouped launch);
* FP64 slices equal the drop-in ChebLanNr (SRC/cheblanNr.c) on the same
  matrix, interval, filter and start vector, bit for bit; FP64 DOS requests
  equal kpmdos_ex with the same seed, bit for bit;
* served slices, FP64 and mixed, match the oracle's ChebLanNr directly
  (counts equal, eigenvalues within 1e-10 relative);
* admission keeps the projected memory within the budget, and the requests
  that wait are served in arrival order."""
import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import PREC_FP64, PREC_MIXED, VEC_RADEMACHER, Server

pytestmark = pytest.mark.gpu

MAXIT, TOL = 300, 1e-8


def pool():
    mats = [M.laplacian(60, 60, 1),
            M.perturb_diag(M.laplacian(14, 14, 14), 1e-3, seed=8),
            M.anderson(10, 16.5, seed=9)]
    out = []
    for A in mats:
        lo, hi = M.gershgorin(A)
        out.append((A, lo, hi))
    return out


def workload(P):
    """(kind, matrix, a, b, extra) in arrival order: slices and DOS requests."""
    reqs = [("slice", 0, 0.3, 0.5, 11), ("dos", 1, 0.5, 1.0, 21), ("slice", 1, 0.6, 0.9, 12),
            ("slice", 2, -0.4, 0.4, 13), ("dos", 0, 0.3, 0.9, 22), ("slice", 0, 0.5, 0.7, 14),
            ("slice", 1, 1.0, 1.3, 15), ("dos", 2, -1.0, 1.0, 23)]
    return reqs


def start_vector(P, m, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, P[m][0].n)


def run(ev, P, reqs, prec, budget=1 << 34, stagger=True):
    """Submits reqs (half, three steps, the rest), steps until idle; returns
    per-request results and the server's stats."""
    srv = Server(ev, budget, prec)
    mids = [srv.add_matrix(A, lo, hi) for A, lo, hi in P]
    assert min(mids) >= 0, ev.last_error()
    ids = []
    half = len(reqs) // 2 if stagger else len(reqs)

    def submit(r):
        kind, m, a, b, x = r
        if kind == "slice":
            rid = srv.submit_slice(mids[m], a, b, MAXIT, TOL, start_vector(P, m, x))
        else:
            rid = srv.submit_dos(mids[m], a, b, 120, 1, 40, VEC_RADEMACHER, x)
        assert rid >= 0, ev.last_error()
        ids.append(rid)

    for r in reqs[:half]:
        submit(r)
    peak_running = 0
    for _ in range(3 if stagger else 0):
        assert srv.step() >= 0, ev.last_error()
        peak_running = max(peak_running, srv.stats()["running"])
    for r in reqs[half:]:
        submit(r)
    while True:
        left = srv.step()
        assert left >= 0, ev.last_error()
        peak_running = max(peak_running, srv.stats()["running"])
        if left == 0:
            break
    out = []
    for r, rid in zip(reqs, ids):
        st, ts, t0, td = srv.status(rid)
        assert st == 2 and ts <= t0 <= td
        if r[0] == "slice":
            err, lam, Y, res = srv.slice_result(rid, P[r[1]][0].n)
            assert err == 0
            out.append((lam, res))
        else:
            err, mu, ec = srv.dos_result(rid)
            assert err == 0
            out.append((mu, ec))
    stats = srv.stats()
    stats["peak_running"] = peak_running
    srv.close()
    return out, stats


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_server_results_same_alone_and_in_the_mix(ev, prec):
    P = pool()
    reqs = workload(P)
    mix, _ = run(ev, P, reqs, prec)
    for i, r in enumerate(reqs):
        alone, _ = run(ev, P, [r], prec, stagger=False)
        a, b = mix[i], alone[0]
        assert np.array_equal(a[0], b[0]), (prec, r)
        assert np.array_equal(a[1], b[1]), (prec, r)
        if r[0] == "slice":
            assert len(a[0]) > 0, r
            assert np.all(a[1] < TOL)


def test_server_fp64_equals_drop_in_calls(ev):
    P = pool()
    reqs = workload(P)
    mix, _ = run(ev, P, reqs, PREC_FP64)
    ev.set_precision(PREC_FP64)
    for (kind, m, a, b, x), got in zip(reqs, mix):
        A, lo, hi = P[m]
        assert ev.set_matrix(A) == 0
        iv = np.array([a, b, lo, hi])
        if kind == "slice":
            pol = ev.default_pol()
            pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
            assert ev.find_pol(iv, pol) == 0
            e, lam, Y, res = ev.ChebLanNr(iv, MAXIT, TOL, start_vector(P, m, x), pol)
            ev.free_pol(pol)
            assert e == 0
            assert np.array_equal(lam, got[0]) and np.array_equal(res, got[1])
        else:
            e, mu, ec = ev.kpmdos(120, 1, 40, iv, vec_kind=VEC_RADEMACHER, seed=x)
            assert e == 0
            assert np.array_equal(mu, got[0]) and ec == got[1]


def test_server_admission_by_memory(ev):
    P = pool()
    A = P[0][0]
    per_slice = A.n * (MAXIT + 2) * 8
    reqs = [("slice", 0, 0.3 + 0.1 * i, 0.4 + 0.1 * i, 30 + i) for i in range(5)]
    out, stats = run(ev, P, reqs, PREC_FP64, budget=2 * per_slice, stagger=False)
    assert stats["mem_peak"] <= 2 * per_slice
    assert stats["peak_running"] <= 2
    assert stats["mem_in_use"] == 0 and stats["running"] == 0 and stats["queued"] == 0
    for lam, res in out:
        assert len(lam) > 0 and np.all(res < TOL)


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_server_grouped_launches_equal_per_matrix_filters(ev, prec):
    """The request loop filters the slices of all its matrices in grouped
    launches (one launch per degree and class of column blocks: precision,
    width, 16- or 32-bit column offsets).  The same stream with grouping off
    (one matrix at a time, the matrix's own blocked filter) gives the same
    bits: five slices share the 2-D Laplacian (a 5-column block, filters of
    different degrees), and a random symmetric matrix of 40,000 rows takes
    the 32-bit-index class."""
    import scipy.sparse.linalg as sla
    P = pool()
    R = M.random_symmetric(40000, 8, 16, seed=3)
    lo, hi = M.gershgorin(R)
    top = np.sort(sla.eigsh(R.to_scipy(), k=24, which="LA", return_eigenvectors=False))
    P.append((R, lo, hi))
    reqs = [("slice", 0, 0.3 + 0.08 * i, 0.38 + 0.08 * i, 40 + i) for i in range(5)]
    reqs += [("slice", 1, 0.6, 0.9, 12), ("slice", 3, 0.5 * (top[2] + top[3]),
                                          0.5 * (top[-3] + top[-2]), 16),
             ("dos", 3, lo, hi, 24)]
    try:
        ev.lib.evsl_amd_set_grouped_filter(1)
        grouped, _ = run(ev, P, reqs, prec)
        ev.lib.evsl_amd_set_grouped_filter(0)
        single, _ = run(ev, P, reqs, prec)
    finally:
        ev.lib.evsl_amd_set_grouped_filter(1)
    for r, a, b in zip(reqs, grouped, single):
        assert np.array_equal(a[0], b[0]), (prec, r)
        assert np.array_equal(a[1], b[1]), (prec, r)
        if r[0] == "slice":
            assert len(a[0]) > 0, r
            assert np.all(a[1] < TOL)


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_server_slices_match_the_oracle(ev, oracle, prec):
    """Every slice the request loop serves in the mix, against the oracle's
    ChebLanNr (the C restatement of SRC/cheblanNr.c) on the same matrix,
    interval, filter and start vector: the same number of eigenpairs,
    eigenvalues within 1e-10 relative, residuals below tol."""
    P = pool()
    reqs = workload(P)
    mix, _ = run(ev, P, reqs, prec)
    for (kind, m, a, b, x), got in zip(reqs, mix):
        if kind != "slice":
            continue
        A, lo, hi = P[m]
        oracle.set_csr(A)
        iv = np.array([a, b, lo, hi])
        pol = oracle.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
        assert oracle.find_pol(iv, pol) == 0
        e, lam, _, res = oracle.ChebLanNr(iv, MAXIT, TOL, start_vector(P, m, x), pol)
        oracle.free_pol(pol)
        assert e == 0
        gl = np.sort(got[0])
        print(f"  slice m{m} [{a}, {b}]: served {len(gl)}, oracle {len(lam)}", flush=True)
        assert len(gl) == len(lam) > 0, (prec, kind, m, a, b)
        np.testing.assert_allclose(gl, np.sort(lam), rtol=1e-10, atol=1e-13)
        assert np.all(got[1] < TOL)
