"""Seeded generators for the matrix families the path is measured on.

* ``laplacian(nx, ny, nz)`` -- 5/7-point Laplacian, diag 4 or 6, off-diag -1,
  the matrix TESTS/COMMON/lapl.c:13-65 generates, rows column-sorted as
  cooMat_to_csrMat leaves them (SRC/spmat.c:160-195).
* ``laplacian_eigs(nx, ny, nz, a, b)`` -- its closed-form spectrum in [a,b]
  (TESTS/COMMON/lapl.c:79-109).
* ``perturb_diag`` -- a seeded U[0, eps] diagonal shift (splits the
  Laplacian's degenerate eigenvalues).
* ``anderson(L, W)`` -- 3-D Anderson model on an L^3 periodic lattice,
  hopping -1, on-site potential U[-W/2, W/2].
* ``random_symmetric(n, nnz_row_lo, nnz_row_hi)`` -- N(0,1) values,
  symmetrised.

All return ``evsl_amd.types.HostCsr`` with int32 indices and float64 values.
"""
import numpy as np

from .types import HostCsr

EPS_MULT = 10


def _csr_from_coo(n, rows, cols, vals):
    import scipy.sparse as sp
    m = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    m.sum_duplicates()
    m.sort_indices()
    return HostCsr.from_scipy(m)


def laplacian(nx, ny, nz=1):
    n = nx * ny * nz
    idx = np.arange(n, dtype=np.int64)
    k = idx // (nx * ny)
    i = (idx - k * nx * ny) // nx
    j = idx - k * nx * ny - i * nx
    rows, cols = [idx], [idx]
    vals = [np.full(n, 6.0 if nz > 1 else 4.0)]
    for cond, off in ((k > 0, -nx * ny), (k < nz - 1, nx * ny), (i > 0, -nx),
                      (i < ny - 1, nx), (j > 0, -1), (j < nx - 1, 1)):
        r = idx[cond]
        rows.append(r)
        cols.append(r + off)
        vals.append(np.full(r.size, -1.0))
    return _csr_from_coo(n, np.concatenate(rows), np.concatenate(cols),
                         np.concatenate(vals))


def laplacian_eigs(nx, ny, nz=1, a=-np.inf, b=np.inf):
    """Sorted closed-form eigenvalues in [a - 10 eps, b + 10 eps]."""
    tx = np.sin(np.arange(1, nx + 1) * (0.5 * np.pi / (nx + 1.0)))
    ty = np.sin(np.arange(1, ny + 1) * (0.5 * np.pi / (ny + 1.0)))
    tz = np.sin(np.arange(1, nz + 1) * (0.5 * np.pi / (nz + 1.0)))
    if nz == 1:
        tz = np.zeros(1)
    ev = 4.0 * (tx[:, None, None] ** 2 + ty[None, :, None] ** 2 + tz[None, None, :] ** 2)
    ev = ev.ravel()
    eps = EPS_MULT * np.finfo(float).eps
    ev = ev[(ev >= a - eps) & (ev <= b + eps)]
    return np.sort(ev)


def perturb_diag(A, eps, seed):
    """A + diag(U[0, eps]) with a seeded generator (new HostCsr)."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.0, eps, size=A.n)
    rowid = np.repeat(np.arange(A.n, dtype=np.int64), np.diff(A.ia))
    diag = np.nonzero(A.ja == rowid)[0]
    assert diag.size == A.n, "every row needs a stored diagonal entry"
    a = A.a.copy()
    a[diag] += d[rowid[diag]]
    return HostCsr(A.n, A.ia, A.ja, a)


def anderson(L, W, seed):
    """3-D Anderson model, L^3 periodic lattice, hopping -1, potential U[-W/2,W/2]."""
    n = L ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % L
    y = (idx // L) % L
    z = idx // (L * L)
    rng = np.random.default_rng(seed)
    pot = rng.uniform(-W / 2.0, W / 2.0, size=n)
    rows, cols, vals = [idx], [idx], [pot]
    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        nb = ((x + dx) % L) + ((y + dy) % L) * L + ((z + dz) % L) * L * L
        rows.append(idx)
        cols.append(nb)
        vals.append(np.full(n, -1.0))
    return _csr_from_coo(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def random_symmetric(n, lo, hi, seed):
    """Random sparse symmetric matrix: about lo..hi nonzeros per row, N(0,1)."""
    rng = np.random.default_rng(seed)
    per = rng.integers(max(1, lo // 2), max(2, hi // 2) + 1, size=n)
    rows = np.repeat(np.arange(n, dtype=np.int64), per)
    cols = rng.integers(0, n, size=rows.size)
    vals = rng.standard_normal(rows.size)
    r = np.concatenate([rows, cols, np.arange(n)])
    c = np.concatenate([cols, rows, np.arange(n)])
    v = np.concatenate([vals, vals, rng.standard_normal(n)])
    return _csr_from_coo(n, r, c, v)


def gershgorin(A):
    """[lmin, lmax] enclosing the spectrum (Gershgorin discs)."""
    rowid = np.repeat(np.arange(A.n), np.diff(A.ia))
    diag = np.zeros(A.n)
    off = np.zeros(A.n)
    isd = A.ja == rowid
    np.add.at(diag, rowid[isd], A.a[isd])
    np.add.at(off, rowid[~isd], np.abs(A.a[~isd]))
    return float(np.min(diag - off)), float(np.max(diag + off))
