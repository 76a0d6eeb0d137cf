"""CPU: the premise of the mixed form's FP32 basis copy, in numpy.

The mixed form's Gram-Schmidt passes read an FP32 copy of the Lanczos basis
(products and sums in FP64).  After the three-term recurrence w is already
orthogonal to the last two basis vectors in FP64, so the projection
h = V^T w is small, and the copy's rounding (about 6e-8 per entry, with
signs unrelated to w) leaves components along the FP64 basis that shrink
like n^-1/2.  This test runs full-reorthogonalisation Lanczos (the
recurrence of SRC/cheblanNr.c:180-260, one CGS pass per step) twice with the
same operator -- a symmetric matrix (1476 eigenvalues spread over [-1, 1],
24 separated ones in [1.3, 3.6]) times a vector plus a seeded relative
perturbation of 1e-7, standing in for an FP32 filter -- once projecting
against the FP64 basis and once against its FP32 copy, and checks: the FP64
run stays orthonormal to rounding and the copy's run below the
semi-orthogonality level sqrt(eps) = 1.5e-8; the 24 separated Ritz values
converge in both runs; and the copy moves them by less than a tenth of
what the operator's noise already puts between them and the eigenvalues."""
import numpy as np
import scipy.linalg as sl


def lanczos(A, v0, m, noise_seed, copy32):
    rng = np.random.default_rng(noise_seed)
    n = A.shape[0]
    V = np.zeros((n, m + 1))
    V32 = np.zeros((n, m + 1), dtype=np.float32)
    V[:, 0] = v0 / np.linalg.norm(v0)
    V32[:, 0] = V[:, 0].astype(np.float32)
    al, be = np.zeros(m), np.zeros(m)
    beta = 0.0
    for k in range(m):
        w = A @ V[:, k]
        w = w + 1e-7 * np.linalg.norm(w) / np.sqrt(n) * rng.standard_normal(n)
        if k > 0:
            w -= beta * V[:, k - 1]
        al[k] = w @ V[:, k]
        w -= al[k] * V[:, k]
        B = V32[:, :k + 1].astype(np.float64) if copy32 else V[:, :k + 1]
        w -= B @ (B.T @ w)
        beta = np.linalg.norm(w)
        be[k] = beta
        V[:, k + 1] = w / beta
        V32[:, k + 1] = V[:, k + 1].astype(np.float32)
    return al, be, V


def test_fp32_copy_keeps_orthogonality_and_ritz_values():
    n, m = 1500, 160
    rng = np.random.default_rng(3)
    d = np.sort(np.concatenate([rng.uniform(-1.0, 1.0, n - 24), np.linspace(1.3, 3.6, 24)]))
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * d) @ Q.T
    A = 0.5 * (A + A.T)
    v0 = rng.uniform(-1.0, 1.0, n)
    runs = {}
    for copy32 in (False, True):
        al, be, V = lanczos(A, v0, m, 7, copy32)
        W = V[:, :m]
        orth = np.abs(W.T @ W - np.eye(m)).max()
        theta, S = sl.eigh_tridiagonal(al, be[:m - 1])
        res = np.abs(be[m - 1] * S[-1, :])  # Ritz residual estimates
        runs[copy32] = (orth, theta, res)
    assert runs[False][0] < 1e-13
    assert runs[True][0] < 1.5e-8
    t64, r64 = runs[False][1], runs[False][2]
    t32, r32 = runs[True][1], runs[True][2]
    conv = (r64 < 1e-9) & (r32 < 1e-9)
    assert conv.sum() == 24 and np.all(t64[conv] > 1.2)
    noise = max(np.min(np.abs(d - t)) for t in t64[conv])
    assert noise < 1e-7  # the operator's 1e-7 noise, seen in the Ritz values
    assert np.max(np.abs(t32[conv] - t64[conv])) < 0.1 * noise
    assert max(np.min(np.abs(d - t)) for t in t32[conv]) < 1e-7
