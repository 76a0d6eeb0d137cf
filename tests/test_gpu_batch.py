"""GPU: the batched slice solver (ChebLanNr_batch) against the per-slice
drop-in ChebLanNr, in FP64 (identical results) and in the mixed-precision
form (FP32 filter, FP64 Lanczos, FP64 filtered Rayleigh-Ritz refinement).

Bars: FP64 batch == plain bit for bit; mixed: identical per-slice counts,
eigenvalues within 1e-10 relative of the FP64 plain run, residuals < tol."""
import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import PREC_FP64, PREC_MIXED, VEC_RADEMACHER

pytestmark = pytest.mark.gpu


def setup_slices(ev, A, a, b, lo, hi, nslices, nvec=40, mdeg=300):
    assert ev.set_matrix(A) == 0, ev.last_error()
    intv = np.array([a, b, lo, hi])
    err, mu, ecnt = ev.kpmdos(mdeg, 1, nvec, intv, vec_kind=VEC_RADEMACHER, seed=7)
    assert err == 0
    err, sli = ev.spslicer(mu, intv, nslices, int(10 * ecnt))
    assert err == 0
    mlan = min(max(5 * (int(1 + ecnt / nslices) + 2), 300), A.n)
    pols, ivs = [], []
    for s in range(nslices):
        iv = np.array([sli[s], sli[s + 1], lo, hi])
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
        assert ev.find_pol(iv, pol) == 0
        pols.append(pol)
        ivs.append(iv)
    return ivs, pols, mlan


CASES = {
    "lap2d": lambda: (M.laplacian(100, 100, 1), 0.4, 0.8, 0.0, 8.0, 2),
    # 16 slices of different degrees: a 16-column block (padded series) until
    # slices finish, then the degree-phased narrowing below 8 columns
    "lap2d_16slices": lambda: (M.laplacian(100, 100, 1), 0.3, 1.5, 0.0, 8.0, 16),
    "lap3d": lambda: (M.perturb_diag(M.laplacian(30, 30, 30), 1e-3, seed=5), 0.5, 0.6, None, None, 4),
    "anderson": lambda: (M.anderson(24, 16.5, seed=4), -0.1, 0.1, None, None, 4),
    "randsym": lambda: (M.random_symmetric(20000, 8, 32, seed=3), -0.2, 0.2, None, None, 3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_batch_fp64_equals_plain_and_mixed_matches(ev, case):
    A, a, b, lo, hi, ns = CASES[case]()
    if lo is None:
        lo, hi = M.gershgorin(A)
    ivs, pols, mlan = setup_slices(ev, A, a, b, lo, hi, ns)
    v0 = np.random.default_rng(11).uniform(-1, 1, A.n)
    ev.set_precision(PREC_FP64)
    plain = [ev.ChebLanNr(ivs[s], mlan, 1e-8, v0, pols[s]) for s in range(ns)]
    for p in plain:
        assert p[0] == 0, ev.last_error()
    err, b64 = ev.ChebLanNr_batch(ivs, mlan, 1e-8, v0, pols)
    assert err == 0, ev.last_error()
    for (e, lp, Yp, rp), (lb, Yb, rb) in zip(plain, b64):
        assert np.array_equal(lp, lb) and np.array_equal(rp, rb)
    ev.set_precision(PREC_MIXED)
    err, mx = ev.ChebLanNr_batch(ivs, mlan, 1e-8, v0, pols)
    # the drop-in ChebLanNr in MIXED mode is the one-slice batch: per-column
    # results do not depend on the block width, so slice 0 is reproduced
    e1, l1, _, r1 = ev.ChebLanNr(ivs[0], mlan, 1e-8, v0, pols[0])
    ev.set_precision(PREC_FP64)
    assert err == 0, ev.last_error()
    assert e1 == 0, ev.last_error()
    assert len(l1) == len(mx[0][0])
    np.testing.assert_allclose(l1, mx[0][0], rtol=1e-13, atol=1e-15)
    S = A.to_scipy()
    for s, ((e, lp, _, _), (lm, Ym, rm)) in enumerate(zip(plain, mx)):
        assert len(lm) == len(lp), (s, len(lm), len(lp))
        if len(lm) == 0:
            continue
        np.testing.assert_allclose(np.sort(lm), np.sort(lp), rtol=1e-10, atol=1e-13)
        assert np.all(rm < 1e-8)
        R = S @ Ym - Ym * lm
        np.testing.assert_allclose(np.linalg.norm(R, axis=0), rm, rtol=1e-5, atol=1e-13)
        assert np.abs(Ym.T @ Ym - np.eye(len(lm))).max() < 1e-8
        assert np.all((lm >= ivs[s][0] - 1e-15) & (lm <= ivs[s][1] + 1e-15))
    for p in pols:
        ev.free_pol(p)


def test_batch_rejects_mixed_bounds(ev):
    A = M.laplacian(20, 20, 1)
    assert ev.set_matrix(A) == 0
    p1, p2 = ev.default_pol(), ev.default_pol()
    assert ev.find_pol(np.array([0.4, 0.8, 0.0, 8.0]), p1) == 0
    assert ev.find_pol(np.array([0.8, 1.2, 0.0, 9.0]), p2) == 0
    err, _ = ev.ChebLanNr_batch([np.array([0.4, 0.8, 0.0, 8.0]), np.array([0.8, 1.2, 0.0, 9.0])],
                                300, 1e-8, np.ones(A.n), [p1, p2])
    assert err == -104  # EVSL_ERR_ARG
    ev.free_pol(p1)
    ev.free_pol(p2)


def test_mixed_low_degree_end_interval_matches_fp64(ev):
    """An end interval of a random symmetric matrix gets a degree-4 filter,
    too weak for the FP32-filtered basis plus refinement to reach tol.  With
    the default minimum degree the slice runs in FP64 under the mixed form;
    with the minimum lowered to 1 it runs mixed, its refinement stalls, and it
    is redone from its start vector in FP64.  Either way the result equals
    the FP64 run bit for bit."""
    A = M.random_symmetric(39817, 8, 32, seed=487532661)
    lo, hi = M.gershgorin(A)
    assert ev.set_matrix(A) == 0
    iv = np.array([-35.942844323629906, -10.02213023817233, lo, hi])
    pol = ev.default_pol()
    pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
    assert ev.find_pol(iv, pol) == 0
    assert pol.deg < 32, pol.deg
    v0 = np.random.default_rng(3).uniform(-1, 1, A.n)
    ev.set_precision(PREC_FP64)
    e64, (r64,) = ev.ChebLanNr_batch([iv], 300, 1e-8, v0, [pol])
    assert e64 == 0 and len(r64[0]) > 0
    ev.set_precision(PREC_MIXED)
    try:
        for mindeg in (32, 1):
            assert ev.lib.evsl_amd_set_mixed_degree_range(mindeg, 20000) == 0
            ev.reset_stats()
            em, (rm,) = ev.ChebLanNr_batch([iv], 300, 1e-8, v0, [pol])
            assert em == 0, ev.last_error()
            assert np.array_equal(rm[0], r64[0]) and np.array_equal(rm[2], r64[2]), mindeg
            st = ev.stats()
            if mindeg == 1:
                assert st["n_fp64_reruns"] == 1 and st["n_fp64_slices"] == 0  # redone in FP64
            else:
                assert st["n_fp64_slices"] == 1 and st["n_fp64_reruns"] == 0  # FP64 from the start
    finally:
        ev.lib.evsl_amd_set_mixed_degree_range(32, 20000)
        ev.set_precision(PREC_FP64)
        ev.free_pol(pol)


def test_mixed_capped_filter_runs_fp64(ev):
    """A filter that ran out of degrees (find_pol's thresholds unmet, deg =
    max_deg - 1) runs in FP64 under the mixed form even inside the mixed
    degree range: the result equals the FP64 run bit for bit and no
    refinement column is spent.  The same interval with room for the degree
    it needs runs the FP32 filter (refinement columns > 0)."""
    A = M.perturb_diag(M.laplacian(60, 60, 1), 1e-3, seed=17)
    lo, hi = M.gershgorin(A)
    assert ev.set_matrix(A) == 0
    iv = np.array([0.40, 0.43, lo, hi])
    v0 = np.random.default_rng(5).uniform(-1, 1, A.n)
    pols = {}
    for cap in (40, 3000):
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, cap
        assert ev.find_pol(iv, pol) == 0
        pols[cap] = pol
    assert pols[40].deg == 39 and 32 <= pols[40].deg <= 3000
    assert 32 <= pols[3000].deg < 2999, pols[3000].deg
    try:
        for cap, pol in pols.items():
            ev.set_precision(PREC_FP64)
            e64, (r64,) = ev.ChebLanNr_batch([iv], 400, 1e-8, v0, [pol])
            assert e64 == 0, ev.last_error()
            ev.set_precision(PREC_MIXED)
            ev.reset_stats()
            em, (rm,) = ev.ChebLanNr_batch([iv], 400, 1e-8, v0, [pol])
            assert em == 0, ev.last_error()
            cols = ev.stats()["n_refine_cols"]
            print(f"  max_deg {cap}: deg {pol.deg}, FP64 {len(r64[0])} pairs, mixed "
                  f"{len(rm[0])}, refinement columns {cols}", flush=True)
            if cap == 40:
                assert cols == 0
                assert np.array_equal(rm[0], r64[0]) and np.array_equal(rm[2], r64[2])
                assert np.array_equal(rm[1], r64[1])
            else:
                assert cols > 0
                assert len(rm[0]) == len(r64[0])
    finally:
        ev.set_precision(PREC_FP64)
        for pol in pols.values():
            ev.free_pol(pol)


def test_degenerate_spectrum_mixed_pairs_are_true_eigenpairs(ev):
    """Unperturbed 3-D Laplacian: eigenvalues of multiplicity up to 6.  A
    single-start-vector Lanczos in FP64 finds one copy of each multiple
    eigenvalue; the FP32 filter's rounding lets the mixed form find further
    copies.  Every pair it returns must be a true one: each eigenvalue within
    1e-10 of a closed-form eigenvalue, no value more often than its
    multiplicity, orthonormal vectors, residuals below tol."""
    nx = 22
    A = M.laplacian(nx, nx, nx)
    a, b = 0.9, 1.2
    exact = M.laplacian_eigs(nx, nx, nx, a, b)
    ivs, pols, mlan = setup_slices(ev, A, a, b, 0.0, 12.0, 2)
    v0 = np.random.default_rng(5).uniform(-1, 1, A.n)
    ev.set_precision(PREC_FP64)
    e64, r64 = ev.ChebLanNr_batch(ivs, mlan, 1e-8, v0, pols)
    ev.set_precision(PREC_MIXED)
    emx, rmx = ev.ChebLanNr_batch(ivs, mlan, 1e-8, v0, pols)
    ev.set_precision(PREC_FP64)
    assert e64 == 0 and emx == 0, ev.last_error()
    S = A.to_scipy()
    uniq, mult = np.unique(np.round(exact, 10), return_counts=True)
    for (l64, _, _), (lm, Ym, rm) in zip(r64, rmx):
        print(f"  slice: FP64 {len(l64)} pairs, mixed {len(lm)}", flush=True)
        idx = np.searchsorted(uniq, lm)
        idx = np.clip(idx, 0, len(uniq) - 1)
        near = np.where(np.abs(uniq[idx] - lm) < np.abs(uniq[np.maximum(idx - 1, 0)] - lm),
                        idx, np.maximum(idx - 1, 0))
        assert np.all(np.abs(uniq[near] - lm) < 1e-10 * np.maximum(1.0, np.abs(lm)))
        used = np.bincount(near, minlength=len(uniq))
        assert np.all(used <= mult)
        assert np.all(rm < 1e-8)
        assert np.abs(Ym.T @ Ym - np.eye(len(lm))).max() < 1e-8
        R = S @ Ym - Ym * lm
        assert np.all(np.linalg.norm(R, axis=0) < 1e-8)
    for p in pols:
        ev.free_pol(p)


def test_mixed_fp32_filter_at_high_degree_anderson(ev):
    """The precision form where the filter degree is in the thousands: two
    slices at the band centre of a 3-D Anderson model (18^3 periodic, W =
    16.5) get filters of degree > 3000, which the mixed form runs in FP32
    (no slice falls back to FP64).  Per-slice counts equal the dense
    spectrum's and the FP64 path's, eigenvalues agree with the dense ones to
    1e-10 and residuals stay below tol."""
    A = M.anderson(18, 16.5, seed=77)
    lo, hi = M.gershgorin(A)
    assert ev.set_matrix(A) == 0
    dense = np.linalg.eigvalsh(A.to_scipy().toarray())
    edges = [-0.012, 0.0, 0.012]
    ivs, pols = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        iv = np.array([a, b, lo, hi])
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 10000
        assert ev.find_pol(iv, pol) == 0
        assert pol.deg > 3000, pol.deg
        ivs.append(iv)
        pols.append(pol)
    want = [int(np.sum((dense >= a) & (dense <= b))) for a, b in zip(edges[:-1], edges[1:])]
    assert min(want) >= 3, want
    v0 = np.random.default_rng(8).uniform(-1, 1, A.n)
    tol = 1e-8
    try:
        ev.set_precision(PREC_FP64)
        e64, r64 = ev.ChebLanNr_batch(ivs, 300, tol, v0, pols)
        assert e64 == 0, ev.last_error()
        ev.set_precision(PREC_MIXED)
        ev.reset_stats()
        em, rm = ev.ChebLanNr_batch(ivs, 300, tol, v0, pols)
        assert em == 0, ev.last_error()
        st = ev.stats()
    finally:
        ev.set_precision(PREC_FP64)
    print(f"  degrees {[p.deg for p in pols]}, counts dense {want}, "
          f"fp64 {[len(r[0]) for r in r64]}, mixed {[len(r[0]) for r in rm]}; refinement "
          f"{st['n_refine_degcols']} column-degrees, {st['n_refine_extra']} extra passes",
          flush=True)
    assert st["n_fp64_slices"] == 0 and st["n_fp64_reruns"] == 0
    for (a, b), w, (l64, _, _), (lm, _, res) in zip(zip(edges[:-1], edges[1:]), want, r64, rm):
        ref = np.sort(dense[(dense >= a) & (dense <= b)])
        assert len(l64) == w and len(lm) == w, (a, b, w, len(l64), len(lm))
        assert np.max(np.abs(np.sort(lm) - ref)) <= 1e-10
        assert np.max(res) < tol
    for p in pols:
        ev.free_pol(p)


def test_mixed_stalled_refinement_retried_at_half_guard(ev):
    """A refinement that stops converging is run once more at half the guard
    threshold before the slice would be redone in FP64.  Four band-centre
    slices of a 3-D Anderson model (24^3 periodic, W = 16.5, filter degrees
    ~1800) with the guard raised to 0.99: one slice's refinement stalls,
    is retried at 0.495 and converges there -- no slice is redone in FP64,
    the per-slice counts equal the FP64 run's, eigenvalues agree with it to
    1e-10 and residuals stay below tol."""
    A = M.anderson(24, 16.5, seed=5)
    lo, hi = M.gershgorin(A)
    assert ev.set_matrix(A) == 0
    intv = np.array([-0.05, 0.05, lo, hi])
    ev.set_precision(PREC_FP64)
    err, mu, ecnt = ev.kpmdos(300, 1, 40, intv, vec_kind=VEC_RADEMACHER, seed=7)
    assert err == 0
    err, sli = ev.spslicer(mu, intv, 4, int(10 * ecnt) + 50)
    assert err == 0
    mlan = min(max(5 * (int(1 + ecnt / 4) + 2), 300), A.n)
    ivs, pols = [], []
    for k in range(4):
        iv = np.array([sli[k], sli[k + 1], lo, hi])
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 10000
        assert ev.find_pol(iv, pol) == 0
        ivs.append(iv)
        pols.append(pol)
    v0 = np.random.default_rng(3).uniform(-1, 1, A.n)
    tol = 1e-8
    try:
        e64, r64 = ev.ChebLanNr_batch(ivs, mlan, tol, v0, pols)
        assert e64 == 0, ev.last_error()
        ev.set_precision(PREC_MIXED)
        ev.set_refine_guard(0.99)
        ev.reset_stats()
        em, rm = ev.ChebLanNr_batch(ivs, mlan, tol, v0, pols)
        assert em == 0, ev.last_error()
        st = ev.stats()
    finally:
        ev.set_refine_guard(0.8)
        ev.set_precision(PREC_FP64)
        for p in pols:
            ev.free_pol(p)
    print(f"  degrees {[p.deg for p in pols]}, counts {[len(r[0]) for r in rm]}, "
          f"retries {st['n_refine_retries']}, reruns {st['n_fp64_reruns']}, "
          f"extra passes {st['n_refine_extra']}", flush=True)
    assert st["n_refine_retries"] >= 1 and st["n_fp64_reruns"] == 0
    for (l64, _, _), (lm, _, res) in zip(r64, rm):
        assert len(lm) == len(l64) > 0
        l64s, lms = np.sort(l64), np.sort(lm)
        assert np.all(np.abs(lms - l64s) <= 1e-10 * np.maximum(1.0, np.abs(l64s)))
        assert np.all(res < tol)


def random_trial(trial):
    """Trial `trial` of the seeded sweep: a matrix without repeated
    eigenvalues (perturbed 2-D or 3-D Laplacian, Anderson model with W in
    [4, 20], random symmetric), 2-4 consecutive slices of about 40-150
    eigenvalues each, cut by spslicer from the KPM density of the whole
    spectrum (the first trial of each kind whose run starts at slice 0 has
    an end interval)."""
    rng = np.random.default_rng(1000 + trial)
    kind = trial % 4
    seed = int(rng.integers(1 << 30))
    if kind == 0:
        nx, ny = (int(v) for v in rng.integers(70, 140, size=2))
        A = M.perturb_diag(M.laplacian(nx, ny, 1), 1e-3, seed=seed)
    elif kind == 1:
        m = int(rng.integers(16, 24))
        A = M.perturb_diag(M.laplacian(m, m + 1, m + 2), 1e-3, seed=seed)
    elif kind == 2:
        A = M.anderson(int(rng.integers(14, 22)), float(rng.uniform(4.0, 20.0)), seed=seed)
    else:
        A = M.random_symmetric(int(rng.integers(5000, 16000)), 8, 32, seed=seed)
    per = int(rng.integers(40, 150))
    ns = int(rng.integers(2, 5))
    nall = max(ns + 1, A.n // per)
    first = 0 if trial < 4 else int(rng.integers(0, nall - ns + 1))
    return A, nall, first, ns


@pytest.mark.parametrize("trial", range(12))
def test_mixed_counts_equal_fp64_randomized(ev, trial):
    """Seeded sweep of the mixed form against FP64 batched (the stopping
    test of the mixed form -- trace change below 1e-7 max(1, nconv), count of
    Ritz values above bar unchanged -- differs from the reference's 2e-12):
    identical per-slice counts, eigenvalues within 1e-10 relative, residuals
    below tol, on matrices whose eigenvalues are simple."""
    A, nall, first, ns = random_trial(trial)
    lo, hi = M.gershgorin(A)
    assert ev.set_matrix(A) == 0, ev.last_error()
    ev.set_precision(PREC_FP64)
    whole = np.array([lo, hi, lo, hi])
    err, mu, ecnt = ev.kpmdos(300, 1, 40, whole, vec_kind=VEC_RADEMACHER, seed=7)
    assert err == 0
    err, sli = ev.spslicer(mu, whole, nall, 10 * A.n)
    assert err == 0
    ivs = [np.array([sli[k], sli[k + 1], lo, hi]) for k in range(first, first + ns)]
    mlan = min(max(5 * (int(1 + A.n / nall) + 2), 300), A.n)
    pols = []
    for iv in ivs:
        pol = ev.default_pol()
        pol.damping, pol.thresh_int, pol.thresh_ext, pol.max_deg = 2, 0.8, 0.2, 3000
        assert ev.find_pol(iv, pol) == 0
        pols.append(pol)
    v0 = np.random.default_rng(50 + trial).uniform(-1, 1, A.n)
    tol = 1e-8
    try:
        e64, r64 = ev.ChebLanNr_batch(ivs, mlan, tol, v0, pols)
        assert e64 == 0, ev.last_error()
        ev.set_precision(PREC_MIXED)
        em, rm = ev.ChebLanNr_batch(ivs, mlan, tol, v0, pols)
        assert em == 0, ev.last_error()
    finally:
        ev.set_precision(PREC_FP64)
        for p in pols:
            ev.free_pol(p)
    print(f"  trial {trial}: n {A.n}, slices {first}..{first + ns - 1} of {nall}, degrees "
          f"{[p.deg for p in pols]}, counts {[len(r[0]) for r in r64]}", flush=True)
    assert sum(len(r[0]) for r in r64) > 0
    for (l64, _, _), (lm, _, res) in zip(r64, rm):
        assert len(lm) == len(l64)
        if len(lm):
            l64s, lms = np.sort(l64), np.sort(lm)
            assert np.all(np.abs(lms - l64s) <= 1e-10 * np.maximum(1.0, np.abs(l64s)))
            assert np.all(res < tol)


@pytest.mark.parametrize("dims", [(27, 29, 31), (30, 30, 30)])
def test_fp32_basis_copy_keeps_results(ev, dims):
    """Mixed form with the Gram-Schmidt passes reading the FP32 copy of the
    basis (evsl_amd_set_fp32_basis, the default) against the same solve with
    the passes on the FP64 basis, and against FP64 batched: identical
    per-slice counts, eigenvalues within 1e-10 relative, residuals below tol,
    and the passes' byte count (8 bytes per basis entry instead of 16)
    follows.  An odd n (27 x 29 x 31) takes the kernels' single-row tails."""
    A = M.perturb_diag(M.laplacian(*dims), 1e-3, seed=17)
    lo, hi = M.gershgorin(A)
    ivs, pols, mlan = setup_slices(ev, A, 0.45, 0.6, lo, hi, 4)
    v0 = np.random.default_rng(8).uniform(-1, 1, A.n)
    tol = 1e-8
    out, st = {}, {}
    try:
        ev.set_precision(PREC_FP64)
        e, out["fp64"] = ev.ChebLanNr_batch(ivs, mlan, tol, v0, pols)
        assert e == 0, ev.last_error()
        ev.set_precision(PREC_MIXED)
        for on in (0, 1):
            ev.set_fp32_basis(on)
            ev.reset_stats()
            e, out[on] = ev.ChebLanNr_batch(ivs, mlan, tol, v0, pols)
            assert e == 0, ev.last_error()
            st[on] = ev.stats()
    finally:
        ev.set_fp32_basis(1)
        ev.set_precision(PREC_FP64)
        for p in pols:
            ev.free_pol(p)
    n = A.n
    for on in (0, 1):
        assert st[on]["n_fp64_slices"] == 0 and st[on]["n_lanczos"] > 0
    # a pass against k + 1 vectors counts (16 or 8) n (k + 1) + 24 n bytes
    r = st[1]["reorth_bytes"] - 24.0 * n * st[1]["n_reorth_pass"]
    r0 = st[0]["reorth_bytes"] - 24.0 * n * st[0]["n_reorth_pass"]
    assert 0.4 < r / r0 < 0.6
    print(f"  n {n}: counts {[len(o[0]) for o in out[1]]}, steps {st[0]['n_lanczos']} / "
          f"{st[1]['n_lanczos']}, basis bytes ratio {r / r0:.3f}, reorth "
          f"{st[0]['t_reorth']:.3f} / {st[1]['t_reorth']:.3f} s", flush=True)
    for (l64, _, _), (l0, _, r0_), (l1, _, r1_) in zip(out["fp64"], out[0], out[1]):
        assert len(l64) == len(l0) == len(l1) > 0
        s64 = np.sort(l64)
        for lm, rm in ((l0, r0_), (l1, r1_)):
            assert np.all(np.abs(np.sort(lm) - s64) <= 1e-10 * np.maximum(1.0, np.abs(s64)))
            assert np.all(rm < tol)
