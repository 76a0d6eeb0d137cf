"""GPU: KPM densities of several matrices of uneven size in one batched run
(kpmdos_batch; each request is kpmdos, SRC/spslice.c:29-148).

Bars:
* FP64: every request's moments and count equal kpmdos_ex on its matrix with
  the same seed bit for bit (the grouped launch runs each row block as the
  one-matrix launch does, and sums the block partials in the same order);
* a request gives the same result alone and in the mix, FP64 and mixed;
* mixed (FP32 recurrence on (A - ctr I) 2^-s, FP64 moment sums): each moment
  within (2/pi) * 8 * eps_fp32 * k^2 of the FP64 one -- the first-order bound
  of the three-term recurrence's rounding (a rounding at step j reaches T_k
  through U_{k-j}, of norm <= k - j + 1), times the moment scale 2/pi for
  unit vectors."""
import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import PREC_FP64, PREC_MIXED, VEC_GAUSSIAN, VEC_RADEMACHER

pytestmark = pytest.mark.gpu

MDEG, NVEC = 120, 40  # 40 vectors: a 32-column block, then an 8-column one


def pool():
    """Matrices of uneven size from the benchmark families, with intervals."""
    mats = [
        (M.laplacian(40, 40, 1), (0.4, 0.8, 0.0, 8.0)),
        (M.perturb_diag(M.laplacian(20, 20, 20), 1e-3, seed=1), (0.5, 1.0, 0.0, 12.001)),
        (M.random_symmetric(6000, 8, 32, seed=2), None),
        (M.laplacian(100, 100, 1), (1.0, 2.0, 0.0, 8.0)),
        (M.anderson(12, 16.5, seed=3), (-0.5, 0.5, None, None)),
    ]
    out = []
    for A, iv in mats:
        lo, hi = M.gershgorin(A)
        if iv is None:
            iv = (-0.3 * hi, 0.3 * hi, lo, hi)
        elif iv[2] is None:
            iv = (iv[0], iv[1], lo, hi)
        out.append((A, np.array(iv, dtype=np.float64)))
    return out


def requests(P):
    """11 requests over the pool, repeats included (a matrix serves several)."""
    order = [0, 1, 2, 3, 4, 1, 3, 0, 2, 4, 3]
    mats = [P[i][0] for i in order]
    ivs = []
    for j, i in enumerate(order):
        iv = P[i][1].copy()
        if j >= len(P):  # repeats: a shifted interval
            w = iv[1] - iv[0]
            iv[0] += 0.25 * w
            iv[1] += 0.25 * w
        ivs.append(iv)
    seeds = [101 + 7 * j for j in range(len(order))]
    return mats, ivs, seeds


def test_kpm_batch_fp64_equals_single_matrix_runs(ev):
    mats, ivs, seeds = requests(pool())
    err, mu, ecnt = ev.kpmdos_batch(mats, ivs, MDEG, 1, NVEC, VEC_RADEMACHER, seeds, PREC_FP64)
    assert err == 0, ev.last_error()
    for r, (A, iv, sd) in enumerate(zip(mats, ivs, seeds)):
        assert ev.set_matrix(A) == 0
        e1, mu1, ec1 = ev.kpmdos(MDEG, 1, NVEC, iv, vec_kind=VEC_RADEMACHER, seed=sd)
        assert e1 == 0
        assert np.array_equal(mu[r], mu1), (r, np.abs(mu[r] - mu1).max())
        assert ecnt[r] == ec1, (r, ecnt[r], ec1)
        assert mu[r][0] == pytest.approx(1 / np.pi, abs=1e-16)


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_kpm_batch_same_alone_and_in_the_mix(ev, prec):
    mats, ivs, seeds = requests(pool())
    err, mu, ecnt = ev.kpmdos_batch(mats, ivs, MDEG, 1, NVEC, VEC_GAUSSIAN, seeds, prec)
    assert err == 0, ev.last_error()
    for r in (0, 2, 4, 7, 10):
        e1, mu1, ec1 = ev.kpmdos_batch([mats[r]], [ivs[r]], MDEG, 1, NVEC, VEC_GAUSSIAN,
                                       [seeds[r]], prec)
        assert e1 == 0, ev.last_error()
        assert np.array_equal(mu[r], mu1[0]) and ecnt[r] == ec1[0], r


def test_kpm_batch_mixed_within_recurrence_bound(ev):
    mats, ivs, seeds = requests(pool())
    e64, mu64, ec64 = ev.kpmdos_batch(mats, ivs, MDEG, 1, NVEC, VEC_RADEMACHER, seeds,
                                      PREC_FP64)
    e32, mu32, ec32 = ev.kpmdos_batch(mats, ivs, MDEG, 1, NVEC, VEC_RADEMACHER, seeds,
                                      PREC_MIXED)
    assert e64 == 0 and e32 == 0, ev.last_error()
    k = np.arange(MDEG + 1)
    bound = (2 / np.pi) * 8 * np.finfo(np.float32).eps * np.maximum(k, 1) ** 2
    ratio = np.abs(mu32 - mu64) / bound
    print(f"  mixed vs FP64 moments: max |dmu| {np.abs(mu32 - mu64).max():.2e}, "
          f"max {ratio.max():.2e} of the bound; counts {np.abs(ec32 - ec64).max():.2e}",
          flush=True)
    assert np.all(ratio <= 1.0)
    assert np.all(mu32[:, 0] == mu64[:, 0])


def test_kpm_batch_errors(ev):
    P = pool()
    A, iv = P[0]
    bad = iv.copy()
    bad[0], bad[1] = bad[1], bad[0]
    err, _, _ = ev.kpmdos_batch([A, A], [iv, bad], 20, 1, 4, VEC_RADEMACHER)
    assert err == -1
    err, _, _ = ev.kpmdos_batch([A], [iv], 20, 1, 4, vec_kind=0)  # libc stream: unsupported
    assert err == -106
    err, _, _ = ev.kpmdos_batch([A], [iv], 20, 1, 0, VEC_RADEMACHER)
    assert err == -104
