"""The C-ABI library loads, exports every entry point the headers declare, and
its host-only logic (filter design, slicers, sparse helpers) matches the
oracle.  No device work is done here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from evsl_amd import matrices as M
from evsl_amd.evsl import EXPORTS, LIB_PATH
from evsl_amd.types import CooMat, CsrMat, PolParams, dptr, iptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    return C.CDLL(LIB_PATH)


@pytest.fixture(scope="module")
def lib_ev(lib):
    """The ctypes mirror without a started context (host entry points only)."""
    from evsl_amd.evsl import EVSL
    return EVSL()


def declared_functions():
    names = set()
    for h in ("evsl.h", "evsl_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}()]*(\([^)]*\)[^;{}()]*)*)\)\s*;", text):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_exports_every_declared_symbol(lib):
    decl = declared_functions()
    assert len(decl) >= 60
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(EXPORTS) <= decl


def test_struct_sizes_match_reference_layout():
    # LP64 layouts of INC/struct.h (CPU build)
    assert C.sizeof(CsrMat) == 40
    assert C.sizeof(CooMat) == 40
    assert C.sizeof(PolParams) == 96


def test_find_pol_bitwise_equal_to_oracle(lib, oracle):
    lib.set_pol_def.argtypes = [C.POINTER(PolParams)]
    lib.find_pol.argtypes = [C.POINTER(C.c_double), C.POINTER(PolParams)]
    lib.free_pol.argtypes = [C.POINTER(PolParams)]
    cases = [(0.4, 0.8, 0, 8), (0.2, 0.3, 0, 12), (0.0, 0.6, 0, 8), (7.2, 8.0, 0, 8),
             (-0.05, 0.05, -14.25, 14.25), (1.0, 1.3, -0.01, 12.05)]
    for a, b, lo, hi in cases:
        intv = np.array([a, b, lo, hi], dtype=float)
        pg, po = PolParams(), oracle.default_pol()
        lib.set_pol_def(C.byref(pg))
        pg.max_deg = po.max_deg = 3000 if hi - lo < 20 else 10000
        assert lib.find_pol(dptr(intv), C.byref(pg)) == 0
        assert oracle.find_pol(intv, po) == 0
        assert (pg.type, pg.deg) == (po.type, po.deg), (a, b)
        assert (pg.gam, pg.bar, pg.cc, pg.dd) == (po.gam, po.bar, po.cc, po.dd)
        assert np.array_equal(pg.coefficients(), po.coefficients())
        lib.free_pol(C.byref(pg))
        oracle.free_pol(po)


def test_forced_degree_matches_oracle(lib, oracle):
    lib.find_pol.argtypes = [C.POINTER(C.c_double), C.POINTER(PolParams)]
    intv = np.array([0.4, 0.8, 0.0, 8.0])
    for d in (15, 40):
        pg, po = PolParams(), oracle.default_pol()
        lib.set_pol_def(C.byref(pg))
        pg.deg = po.deg = d
        assert lib.find_pol(dptr(intv), C.byref(pg)) == 0
        assert oracle.find_pol(intv, po) == 0
        assert pg.deg == po.deg == d - 1
        assert np.array_equal(pg.coefficients(), po.coefficients())


def test_spslicer_and_spslicer2_match_oracle(lib, oracle, libc):
    A = M.laplacian(30, 30, 1)
    oracle.set_csr(A)
    intv = np.array([0.4, 1.6, 0.0, 8.0])
    libc.srand(4)
    _, mu, ecnt = oracle.kpmdos(100, 1, 10, intv)
    lib.spslicer.argtypes = [C.POINTER(C.c_double)] * 2 + [C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int]
    for n_int in (1, 3, 6):
        sg = np.empty(n_int + 1)
        eg = lib.spslicer(dptr(sg), dptr(mu), 100, dptr(intv), n_int, int(10 * ecnt))
        eo, so = oracle.spslicer(mu, intv, n_int, int(10 * ecnt))
        assert eg == eo and np.array_equal(sg, so)
    x = np.linspace(0, 3, 101)
    y = np.exp(-x) + 0.2
    lib.spslicer2.argtypes = [C.POINTER(C.c_double)] * 2 + [C.c_int, C.c_int, C.POINTER(C.c_double)]
    sg = np.empty(6)
    yy = y.copy()
    lib.spslicer2(dptr(x), dptr(yy), 5, len(x), dptr(sg))
    assert np.array_equal(sg, oracle.spslicer2(x, y, 5))


def test_coo_to_csr_sorts_rows(lib):
    rng = np.random.default_rng(0)
    n = 50
    r = rng.integers(0, n, 300).astype(np.int32)
    c = rng.integers(0, n, 300).astype(np.int32)
    keep = np.unique(r.astype(np.int64) * n + c, return_index=True)[1]
    r, c = r[keep], c[keep]
    v = rng.standard_normal(len(r))
    coo = CooMat(n, n, len(r), iptr(r), iptr(c), dptr(v))
    csr = CsrMat()
    lib.cooMat_to_csrMat.argtypes = [C.c_int, C.POINTER(CooMat), C.POINTER(CsrMat)]
    assert lib.cooMat_to_csrMat(0, C.byref(coo), C.byref(csr)) == 0
    ia = np.ctypeslib.as_array(csr.ia, shape=(n + 1,)).copy()
    ja = np.ctypeslib.as_array(csr.ja, shape=(len(r),)).copy()
    a = np.ctypeslib.as_array(csr.a, shape=(len(r),)).copy()
    import scipy.sparse as sp
    ref = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    ref.sort_indices()
    assert np.array_equal(ia, ref.indptr) and np.array_equal(ja, ref.indices)
    assert np.array_equal(a, ref.data)
    lib.free_csr.argtypes = [C.POINTER(CsrMat)]
    lib.free_csr(C.byref(csr))


def test_host_rng_matches_reference_mapping(lib, libc):
    lib.rand_double.argtypes = [C.c_int, C.POINTER(C.c_double)]
    libc.srand(99)
    a = np.empty(1000)
    lib.rand_double(1000, dptr(a))
    libc.srand(99)
    from tests.conftest import libc_rand_double
    assert np.array_equal(a, libc_rand_double(libc, 1000))
    assert a.min() >= -1.0 and a.max() <= 1.0


def test_read_coo_mm_matches_scipy(lib_ev):
    """The drivers' Matrix Market reader (TESTS/COMMON/io.c:115-217) plus
    cooMat_to_csrMat on the reference's own MATRICES/A.mtx: the symmetric
    expansion gives scipy's matrix entry for entry."""
    import os
    import scipy.io
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "A.mtx")
    err, A = lib_ev.read_mm(path)
    assert err == 0
    ref = scipy.io.mmread(path).tocsr()
    ref.sort_indices()
    assert A.n == ref.shape[0]
    assert np.array_equal(A.ia, ref.indptr) and np.array_equal(A.ja, ref.indices)
    assert np.array_equal(A.a, ref.data)
    assert lib_ev.read_mm(path + ".missing")[0] < 0


def test_fails_loudly_without_a_gpu(lib_ev):
    """Without a usable HIP device every device entry point returns
    EVSL_ERR_DEVICE (-101) with a message; nothing falls back to the CPU.
    (On a machine with a GPU the context starts and the test is skipped.)"""
    from evsl_amd.types import CsrMat
    err = lib_ev.lib.EVSLStart()
    if err == 0:
        lib_ev.lib.EVSLFinish()
        pytest.skip("a HIP device is present")
    assert err == -101
    assert "device" in lib_ev.last_error().lower()
    A = M.laplacian(10, 10, 1)
    assert lib_ev.set_matrix(A) == -101
    intv = np.array([0.5, 1.0, 0.0, 8.0])
    assert lib_ev.kpmdos(20, 1, 2, intv)[0] == -101
    pol = lib_ev.default_pol()
    assert lib_ev.find_pol(intv, pol) == 0  # host-only filter design still works
    v = np.ones(A.n)
    assert lib_ev.ChebLanNr(intv, 50, 1e-8, v, pol)[0] == -101
    assert lib_ev.ChebLanTr(60, 4, intv, 100, 1e-8, v, pol)[0] == -101
    assert lib_ev.LanTrbounds(20, 50, 1e-8, v, 1)[0] == -101
    lib_ev.free_pol(pol)
