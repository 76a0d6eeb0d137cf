"""ctypes mirror of the drop-in C interface exported by libevsl_amd.so.

Every call goes through the C ABI declared in include/evsl.h and
include/evsl_amd.h; this module adds only argument marshalling (numpy in,
numpy out).  Loading fails loudly when the library has not been built --
there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

from .types import MVFUNC, PolParams, c_double_p, c_int_p, dptr, iptr, CsrMat, CooMat, HostCsr

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libevsl_amd.so")

PREC_FP64 = 0
PREC_MIXED = 1
VEC_RAND_DOUBLE = 0
VEC_RADEMACHER = 1
VEC_GAUSSIAN = 2

# C entry points declared in include/evsl.h and include/evsl_amd.h
EXPORTS = [
    "_evsl_Malloc", "_evsl_Calloc", "_evsl_Realloc", "_evsl_Free",
    "_evsl_Malloc_device", "_evsl_Calloc_device", "_evsl_Realloc_device", "_evsl_Free_device",
    "evsl_memcpy_device_to_host", "evsl_memcpy_device_to_device",
    "evsl_memcpy_host_to_device", "evsl_memset_device",
    "EVSLStart", "EVSLFinish", "SetAMatrix", "SetAMatvec", "SetBMatrix", "SetBMatvec",
    "SetBSol", "SetLTSol", "SetStdEig", "SetGenEig", "SetDiagScal",
    "evsl_create_csr_gpu", "evsl_free_csr_gpu", "SetAMatrix_device_csr",
    "set_pol_def", "find_pol", "free_pol",
    "kpmdos", "spslicer", "LanDos", "spslicer2",
    "ChebLanNr", "ChebLanTr", "LanBounds", "LanTrbounds", "ChebSI",
    "cooMat_to_csrMat", "free_csr", "free_coo", "csr_copy", "arrays_copyto_csrMat",
    "matvec_csr", "rand_double", "rand_double_device", "randn_double",
    "randn_double_device", "sort_double", "linspace", "evsl_timer", "StatsPrint",
    "StatsReset",
    "evsl_amd_set_filter_precision", "evsl_amd_get_filter_precision",
    "evsl_amd_set_refine_passes", "evsl_amd_set_refine_guard", "evsl_amd_set_refine_widen",
    "evsl_amd_get_stats", "evsl_amd_reset_stats",
    "evsl_amd_last_error", "evsl_amd_stats_size", "kpmdos_ex", "evsl_amd_bench_filter",
    "ChebLanNr_batch", "evsl_amd_read_coo_mm", "evsl_amd_bench_copy", "evsl_amd_bench_dense",
    "evsl_amd_spmm_vec", "evsl_amd_spmm_index_bytes", "evsl_amd_spmm_index_mode",
    "evsl_amd_set_pattern_codes", "evsl_amd_pattern_codes", "evsl_amd_cheb_av_cols",
    "evsl_amd_cheb_av_cols_noalloc", "evsl_amd_set_grouped_filter",
    "evsl_amd_set_mixed_capped_fp64", "evsl_amd_set_fp32_basis",
    "evsl_amd_spmm_streaming", "kpmdos_batch",
    "evsl_amd_server_create", "evsl_amd_server_destroy", "evsl_amd_server_add_matrix",
    "evsl_amd_server_submit_slice", "evsl_amd_server_submit_dos", "evsl_amd_server_step",
    "evsl_amd_server_status", "evsl_amd_server_slice_result", "evsl_amd_server_dos_result",
    "evsl_amd_server_stats", "evsl_amd_set_mixed_degree_range",
    "evsl_amd_spmm_layout", "evsl_amd_set_spmm_layout", "evsl_amd_set_spmm_tuning",
]


class AmdStats(C.Structure):
    _fields_ = [("t_filter", C.c_double), ("t_reorth", C.c_double), ("t_eig", C.c_double),
                ("t_ritz", C.c_double), ("n_filter_steps", C.c_size_t),
                ("n_filter_cols", C.c_size_t), ("n_lanczos", C.c_size_t),
                ("n_reorth_pass", C.c_size_t), ("filter_bytes", C.c_double),
                ("reorth_bytes", C.c_double), ("n_refine_cols", C.c_size_t),
                ("n_refine_extra", C.c_size_t), ("refine_bytes", C.c_double),
                ("n_refine_degcols", C.c_size_t), ("n_fp64_slices", C.c_size_t),
                ("n_fp64_reruns", C.c_size_t), ("n_refine_fp64_cols", C.c_size_t),
                ("n_refine_retries", C.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    return C.CDLL(path)


class EVSL:
    """The drop-in interface, one process-wide context (as the reference)."""

    def __init__(self, path=LIB_PATH):
        self.lib = L = load_library(path)
        D, I, V, VP = C.c_double, C.c_int, None, C.c_void_p
        dp, ip = c_double_p, c_int_p
        pp = C.POINTER(PolParams)
        ddp = C.POINTER(c_double_p)
        sig = {
            "_evsl_Malloc_device": (VP, [C.c_size_t]),
            "_evsl_Free_device": (V, [VP]),
            "_evsl_Free": (V, [VP]),
            "evsl_memcpy_device_to_host": (V, [VP, VP, C.c_size_t]),
            "evsl_memcpy_host_to_device": (V, [VP, VP, C.c_size_t]),
            "evsl_memcpy_device_to_device": (V, [VP, VP, C.c_size_t]),
            "evsl_memset_device": (V, [VP, I, C.c_size_t]),
            "EVSLStart": (I, []),
            "EVSLFinish": (I, []),
            "SetAMatrix": (I, [C.POINTER(CsrMat)]),
            "SetAMatvec": (I, [I, MVFUNC, VP]),
            "SetAMatrix_device_csr": (I, [C.POINTER(CsrMat)]),
            "evsl_create_csr_gpu": (V, [C.POINTER(CsrMat), C.POINTER(CsrMat)]),
            "evsl_free_csr_gpu": (V, [C.POINTER(CsrMat)]),
            "set_pol_def": (V, [pp]),
            "find_pol": (I, [dp, pp]),
            "free_pol": (V, [pp]),
            "kpmdos": (I, [I, I, I, dp, dp, dp]),
            "kpmdos_ex": (I, [I, I, I, dp, I, C.c_ulonglong, dp, dp]),
            "spslicer": (I, [dp, dp, I, dp, I, I]),
            "spslicer2": (V, [dp, dp, I, I, dp]),
            "LanDos": (I, [I, I, I, dp, dp, dp, dp]),
            "ChebLanNr": (I, [dp, I, D, dp, pp, ip, ddp, ddp, ddp, VP]),
            "ChebLanTr": (I, [I, I, dp, I, D, dp, pp, ip, ddp, ddp, ddp, VP]),
            "ChebSI": (I, [I, dp, I, D, dp, pp, ip, ddp, ddp, ddp, VP]),
            "LanBounds": (I, [I, dp, dp, dp]),
            "LanTrbounds": (I, [I, I, D, dp, I, dp, dp, C.c_void_p]),
            "evsl_amd_read_coo_mm": (I, [C.c_char_p, I, I, C.POINTER(CooMat)]),
            "cooMat_to_csrMat": (I, [I, C.POINTER(CooMat), C.POINTER(CsrMat)]),
            "free_coo": (None, [C.POINTER(CooMat)]),
            "free_csr": (None, [C.POINTER(CsrMat)]),
            "rand_double": (V, [I, dp]),
            "rand_double_device": (V, [I, dp]),
            "randn_double": (V, [I, dp]),
            "evsl_timer": (D, []),
            "StatsReset": (V, []),
            "evsl_amd_set_filter_precision": (I, [I]),
            "evsl_amd_get_filter_precision": (I, []),
            "evsl_amd_set_refine_passes": (I, [I]),
            "evsl_amd_set_refine_guard": (I, [D]),
            "evsl_amd_set_refine_widen": (I, [D]),
            "evsl_amd_set_mixed_degree_range": (I, [I, I]),
            "evsl_amd_get_stats": (I, [C.POINTER(AmdStats)]),
            "evsl_amd_reset_stats": (V, []),
            "evsl_amd_last_error": (C.c_char_p, []),
            "evsl_amd_stats_size": (C.c_size_t, []),
            "evsl_amd_bench_filter": (I, [I, I, I, I, D, D, dp, dp]),
            "evsl_amd_bench_copy": (I, [C.c_size_t, I, dp]),
            "evsl_amd_bench_dense": (I, [I, I, I, I, I, dp, dp, dp]),
            "evsl_amd_spmm_vec": (I, [I, I, I]),
            "evsl_amd_spmm_layout": (I, [I, I, I, C.POINTER(I)]),
            "evsl_amd_set_spmm_layout": (I, [I, I, I, I]),
            "evsl_amd_set_spmm_tuning": (I, [I]),
            "evsl_amd_spmm_index_bytes": (I, []),
            "evsl_amd_spmm_index_mode": (I, [I, I, I]),
            "evsl_amd_set_pattern_codes": (I, [I]),
            "evsl_amd_pattern_codes": (I, []),
            "evsl_amd_spmm_streaming": (I, [I, I, I]),
            "kpmdos_batch": (I, [I, C.POINTER(C.POINTER(CsrMat)), dp, I, I, I, I,
                                 C.POINTER(C.c_ulonglong), I, dp, dp]),
            "evsl_amd_server_create": (VP, [C.c_size_t, I]),
            "evsl_amd_server_destroy": (V, [VP]),
            "evsl_amd_server_add_matrix": (I, [VP, C.POINTER(CsrMat), D, D]),
            "evsl_amd_server_submit_slice": (I, [VP, I, D, D, I, D, pp, dp]),
            "evsl_amd_server_submit_dos": (I, [VP, I, D, D, I, I, I, I, C.c_ulonglong]),
            "evsl_amd_server_step": (I, [VP]),
            "evsl_amd_server_status": (I, [VP, I, dp, dp, dp]),
            "evsl_amd_server_slice_result": (I, [VP, I, ip, ddp, ddp, ddp]),
            "evsl_amd_server_dos_result": (I, [VP, I, dp, dp]),
            "evsl_amd_server_stats": (I, [VP, ip, ip, C.POINTER(C.c_size_t),
                                          C.POINTER(C.c_size_t)]),
            "evsl_amd_cheb_av_cols": (I, [I, pp, dp, dp, I]),
            "evsl_amd_cheb_av_cols_noalloc": (I, [I, pp, VP, VP, I]),
            "evsl_amd_set_grouped_filter": (I, [I]),
            "evsl_amd_set_mixed_capped_fp64": (I, [I]),
            "evsl_amd_set_fp32_basis": (I, [I]),
            "ChebLanNr_batch": (I, [I, dp, I, D, dp, pp, ip, ddp, ddp, ddp, VP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        assert L.evsl_amd_stats_size() == C.sizeof(AmdStats), "evslAmdStats ABI mismatch"
        self._mat = None
        self._mv = None

    # ---- lifecycle / operator ----------------------------------------------
    def start(self):
        return self.lib.EVSLStart()

    def finish(self):
        return self.lib.EVSLFinish()

    def set_matrix(self, A):
        """A: evsl_amd.types.HostCsr."""
        self._mat = A
        return self.lib.SetAMatrix(C.byref(A.csr))

    def set_matvec(self, n, func, data=None):
        self._mv = MVFUNC(func) if not isinstance(func, MVFUNC) else func
        return self.lib.SetAMatvec(n, self._mv, data)

    def last_error(self):
        return self.lib.evsl_amd_last_error().decode()

    # ---- device memory -------------------------------------------------------
    def to_device(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        p = self.lib._evsl_Malloc_device(x.nbytes)
        if not p:
            raise MemoryError(self.last_error())
        self.lib.evsl_memcpy_host_to_device(p, x.ctypes.data, x.nbytes)
        return p

    def from_device(self, p, count):
        out = np.empty(count)
        if count:
            self.lib.evsl_memcpy_device_to_host(out.ctypes.data, p, out.nbytes)
        return out

    def free_device(self, p):
        self.lib._evsl_Free_device(p)

    # ---- filters -----------------------------------------------------------
    def default_pol(self):
        pol = PolParams()
        self.lib.set_pol_def(C.byref(pol))
        return pol

    def find_pol(self, intv, pol):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        return self.lib.find_pol(dptr(intv), C.byref(pol))

    def free_pol(self, pol):
        self.lib.free_pol(C.byref(pol))

    # ---- DOS / slicing ---------------------------------------------------------
    def kpmdos(self, Mdeg, damping, nvec, intv, vec_kind=None, seed=0):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        mu = np.empty(Mdeg + 1)
        ecnt = C.c_double()
        if vec_kind is None:
            err = self.lib.kpmdos(Mdeg, damping, nvec, dptr(intv), dptr(mu), C.byref(ecnt))
        else:
            err = self.lib.kpmdos_ex(Mdeg, damping, nvec, dptr(intv), vec_kind, seed,
                                     dptr(mu), C.byref(ecnt))
        return err, mu, ecnt.value

    def spslicer(self, mu, intv, n_int, npts):
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        sli = np.empty(n_int + 1)
        err = self.lib.spslicer(dptr(sli), dptr(mu), len(mu) - 1, dptr(intv), n_int, npts)
        return err, sli

    def spslicer2(self, xi, yi, n_int):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        yi = np.ascontiguousarray(yi, dtype=np.float64).copy()
        sli = np.empty(n_int + 1)
        self.lib.spslicer2(dptr(xi), dptr(yi), n_int, len(xi), dptr(sli))
        return sli

    def LanDos(self, nvec, msteps, npts, intv):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        x, y, neig = np.empty(npts), np.empty(npts), C.c_double()
        err = self.lib.LanDos(nvec, msteps, npts, dptr(x), dptr(y), C.byref(neig), dptr(intv))
        return err, x, y, neig.value

    # ---- solvers -----------------------------------------------------------------
    def _collect(self, n, nev, lam, W, res):
        if nev > 0:
            lam_a = np.ctypeslib.as_array(lam, shape=(nev,)).copy()
            res_a = np.ctypeslib.as_array(res, shape=(nev,)).copy()
            Y = self.from_device(C.cast(W, C.c_void_p).value, n * nev).reshape(nev, n).T
        else:
            lam_a, res_a, Y = np.empty(0), np.empty(0), np.empty((n, 0))
        if lam:
            self.lib._evsl_Free(C.cast(lam, C.c_void_p))
        if res:
            self.lib._evsl_Free(C.cast(res, C.c_void_p))
        if W:
            self.lib._evsl_Free_device(C.cast(W, C.c_void_p))
        return lam_a, Y, res_a

    def ChebLanNr(self, intv, maxit, tol, vinit, pol, n=None):
        """vinit: numpy array (host) or an int device pointer (then give n)."""
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        if isinstance(vinit, int):
            vp = C.cast(vinit, c_double_p)
            assert n is not None
        else:
            vinit = np.ascontiguousarray(vinit, dtype=np.float64)
            vp = dptr(vinit)
            n = len(vinit)
        nev = C.c_int()
        lam, W, res = c_double_p(), c_double_p(), c_double_p()
        err = self.lib.ChebLanNr(dptr(intv), maxit, tol, vp, C.byref(pol), C.byref(nev),
                                 C.byref(lam), C.byref(W), C.byref(res), None)
        if err:
            return err, None, None, None
        return (0,) + self._collect(n, nev.value, lam, W, res)

    def ChebLanTr(self, lanm, nev, intv, maxit, tol, vinit, pol):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        vinit = np.ascontiguousarray(vinit, dtype=np.float64)
        nev2 = C.c_int()
        lam, W, res = c_double_p(), c_double_p(), c_double_p()
        err = self.lib.ChebLanTr(lanm, nev, dptr(intv), maxit, tol, dptr(vinit), C.byref(pol),
                                 C.byref(nev2), C.byref(lam), C.byref(W), C.byref(res), None)
        if err:
            return err, None, None, None
        return (0,) + self._collect(len(vinit), nev2.value, lam, W, res)

    def ChebSI(self, nev, intv, maxit, tol, V0, pol, fstats=None):
        """V0: n x nev start block (host array, column-major = V0.T row-major)."""
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        V0 = np.asfortranarray(V0, dtype=np.float64)
        n = V0.shape[0]
        nevo = C.c_int()
        lam, W, res = c_double_p(), c_double_p(), c_double_p()
        err = self.lib.ChebSI(nev, dptr(intv), maxit, tol,
                              V0.ctypes.data_as(c_double_p), C.byref(pol), C.byref(nevo),
                              C.byref(lam), C.byref(W), C.byref(res), fstats)
        if err:
            return err, None, None, None
        return (0,) + self._collect(n, nevo.value, lam, W, res)

    def cheb_av_cols(self, pols, X, prec=PREC_FP64):
        """Y[:, j] = p_j(A) X[:, j] with the blocked filter (X: n x ncols host
        array; pols: one PolParams per column, same cc and dd).  Returns
        (err, Y)."""
        X = np.asfortranarray(X, dtype=np.float64)
        ncols = X.shape[1]
        assert len(pols) == ncols
        P = (PolParams * ncols)(*pols)
        Y = np.zeros_like(X, order="F")
        err = self.lib.evsl_amd_cheb_av_cols(ncols, P, X.ctypes.data_as(c_double_p),
                                             Y.ctypes.data_as(c_double_p), prec)
        return err, Y

    def cheb_av_cols_noalloc(self, pols, Xd, Yd, prec=PREC_FP64):
        """The allocation-free form of cheb_av_cols: Xd, Yd are device
        pointers to n x len(pols) column-major blocks.  Returns the error code."""
        P = (PolParams * len(pols))(*pols)
        return self.lib.evsl_amd_cheb_av_cols_noalloc(len(pols), P, C.c_void_p(Xd),
                                                      C.c_void_p(Yd), prec)

    def kpmdos_batch(self, mats, intvs, Mdeg, damping, nvec, vec_kind=VEC_RADEMACHER,
                     seeds=None, prec=PREC_FP64):
        """KPM densities of several matrices in one batched run.  mats: list
        of HostCsr (repeats allowed), intvs: list of 4-vectors.  Returns
        (err, mu [nreq x (Mdeg+1)], ecnt [nreq])."""
        nreq = len(mats)
        P = (C.POINTER(CsrMat) * nreq)(*[C.pointer(A.csr) for A in mats])
        iv = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64)
                                                  for x in intvs]))
        sd = None
        if seeds is not None:
            sd = (C.c_ulonglong * nreq)(*[int(x) for x in seeds])
        mu = np.zeros((nreq, Mdeg + 1))
        ecnt = np.zeros(nreq)
        err = self.lib.kpmdos_batch(nreq, P, dptr(iv), Mdeg, damping, nvec, vec_kind, sd, prec,
                                    mu.ctypes.data_as(c_double_p),
                                    ecnt.ctypes.data_as(c_double_p))
        return err, mu, ecnt

    def ChebLanNr_batch(self, intvs, maxit, tol, vinit, pols, n=None, fstats=None):
        """All slices of one matrix in one batched run.  intvs: list of
        4-vectors, pols: list of PolParams (one per slice).  Returns
        (err, [(lam, Y, res) per slice])."""
        ns = len(pols)
        iv = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in intvs]))
        if isinstance(vinit, int):
            vp = C.cast(vinit, c_double_p)
            assert n is not None
        else:
            vinit = np.ascontiguousarray(vinit, dtype=np.float64)
            vp = dptr(vinit)
            n = len(vinit)
        P = (PolParams * ns)(*pols)
        nev = (C.c_int * ns)()
        lam, W, res = (c_double_p * ns)(), (c_double_p * ns)(), (c_double_p * ns)()
        err = self.lib.ChebLanNr_batch(ns, dptr(iv), maxit, tol, vp, P, nev, lam, W, res, fstats)
        out = [self._collect(n, nev[s], lam[s], W[s], res[s]) for s in range(ns)]
        return err, out

    def LanBounds(self, msteps, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lmin, lmax = C.c_double(), C.c_double()
        err = self.lib.LanBounds(msteps, dptr(v), C.byref(lmin), C.byref(lmax))
        return err, lmin.value, lmax.value

    def LanTrbounds(self, lanm, maxit, tol, v, bndtype, fstats=None):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lmin, lmax = C.c_double(), C.c_double()
        err = self.lib.LanTrbounds(lanm, maxit, tol, dptr(v), bndtype, C.byref(lmin),
                                   C.byref(lmax), fstats)
        return err, lmin.value, lmax.value

    def read_mm(self, path, idxin=1):
        """Matrix Market file -> (err, HostCsr) through the library's reader
        and cooMat_to_csrMat, the ingest path of TESTS/PLanN/MMPLanN.c:102-112."""
        coo = CooMat()
        err = self.lib.evsl_amd_read_coo_mm(str(path).encode(), idxin, 0, C.byref(coo))
        if err:
            return err, None
        csr = CsrMat()
        err = self.lib.cooMat_to_csrMat(0, C.byref(coo), C.byref(csr))
        self.lib.free_coo(C.byref(coo))
        if err:
            return err, None
        n, nnz = csr.nrows, csr.nnz
        ia = np.ctypeslib.as_array(csr.ia, shape=(n + 1,)).copy()
        ja = np.ctypeslib.as_array(csr.ja, shape=(nnz,)).copy()
        a = np.ctypeslib.as_array(csr.a, shape=(nnz,)).copy()
        self.lib.free_csr(C.byref(csr))
        return 0, HostCsr(n, ia, ja, a)

    # ---- extensions -----------------------------------------------------------------
    def set_precision(self, prec):
        return self.lib.evsl_amd_set_filter_precision(prec)

    def set_refine_passes(self, k):
        return self.lib.evsl_amd_set_refine_passes(k)

    def set_refine_guard(self, g):
        return self.lib.evsl_amd_set_refine_guard(g)

    def set_refine_widen(self, f):
        return self.lib.evsl_amd_set_refine_widen(f)

    def set_fp32_basis(self, on):
        return self.lib.evsl_amd_set_fp32_basis(int(on))

    def stats(self):
        s = AmdStats()
        self.lib.evsl_amd_get_stats(C.byref(s))
        return s.as_dict()

    def reset_stats(self):
        self.lib.evsl_amd_reset_stats()

    def bench_filter(self, B, prec, mode, nsteps, cc=6.0, dd=6.0):
        ms, by = C.c_double(), C.c_double()
        err = self.lib.evsl_amd_bench_filter(B, prec, mode, nsteps, cc, dd, C.byref(ms),
                                             C.byref(by))
        return err, ms.value, by.value

    def bench_dense(self, kind, n, k, m, reps=10):
        """(err, ms per repetition, bytes, flops) of a dense kernel (kind 0 a
        Gram-Schmidt pass, 1 the Ritz GEMM, 2 the Gram product)."""
        ms, by, fl = C.c_double(), C.c_double(), C.c_double()
        err = self.lib.evsl_amd_bench_dense(kind, n, k, m, reps, C.byref(ms), C.byref(by),
                                            C.byref(fl))
        return err, ms.value, by.value, fl.value

    def bench_copy(self, nbytes, reps=20):
        """(err, ms per copy) of a device copy of nbytes; 2*nbytes move."""
        ms = C.c_double()
        err = self.lib.evsl_amd_bench_copy(nbytes, reps, C.byref(ms))
        return err, ms.value

    def spmm_vec(self, B, prec, mode):
        return self.lib.evsl_amd_spmm_vec(B, prec, mode)

    def spmm_layout(self, B, prec, mode):
        """(layout 10*V + R the key launches now on the current matrix,
        decided)."""
        d = C.c_int()
        lay = self.lib.evsl_amd_spmm_layout(B, prec, mode, C.byref(d))
        return lay, bool(d.value)

    def set_spmm_layout(self, B, prec, mode, layout):
        """layout > 0 forces a candidate, 0 the fixed policy, -1 tunes."""
        return self.lib.evsl_amd_set_spmm_layout(B, prec, mode, layout)

    def set_spmm_tuning(self, on):
        return self.lib.evsl_amd_set_spmm_tuning(1 if on else 0)

    def spmm_kernel_name(self, B, prec, mode):
        """Template name of the fused filter kernel for (B, prec, mode) on
        the current matrix, as rocprofv3 reports it."""
        vt = "float" if prec == PREC_MIXED and mode != 2 else "double"
        ix = self.lib.evsl_amd_spmm_index_mode(B, prec, mode)
        nt = "true" if self.lib.evsl_amd_spmm_streaming(B, prec, mode) else "false"
        return (f"k_cheb_spmm<{vt}, {B}, {self.spmm_vec(B, prec, mode)}, {mode}, true, {ix}, "
                f"{nt}>")

    def set_pattern_codes(self, on):
        return self.lib.evsl_amd_set_pattern_codes(1 if on else 0)

    def pattern_codes(self):
        return self.lib.evsl_amd_pattern_codes()


class Server:
    """The request loop of include/evsl_amd.h (evsl_amd_server_*), over an
    EVSL instance whose context is started."""

    def __init__(self, ev, mem_budget, prec=PREC_FP64):
        self.ev = ev
        self.lib = ev.lib
        self.h = self.lib.evsl_amd_server_create(int(mem_budget), prec)
        if not self.h:
            raise RuntimeError(ev.last_error())
        self._keep = []   # host CSR arrays of registered matrices
        self.n = {}
        self.mdeg = {}

    def close(self):
        if self.h:
            self.lib.evsl_amd_server_destroy(self.h)
            self.h = None

    def add_matrix(self, A, lmin, lmax):
        self._keep.append(A)
        mid = self.lib.evsl_amd_server_add_matrix(self.h, C.byref(A.csr), lmin, lmax)
        if mid >= 0:
            self.n[mid] = A.n
        return mid

    def submit_slice(self, mat, a, b, maxit, tol, vinit, pol=None):
        v = np.ascontiguousarray(vinit, dtype=np.float64)
        return self.lib.evsl_amd_server_submit_slice(self.h, mat, a, b, maxit, tol,
                                                     C.byref(pol) if pol is not None else None,
                                                     dptr(v))

    def submit_dos(self, mat, a, b, Mdeg, damping, nvec, vec_kind=VEC_RADEMACHER, seed=0):
        rid = self.lib.evsl_amd_server_submit_dos(self.h, mat, a, b, Mdeg, damping, nvec,
                                                  vec_kind, seed)
        if rid >= 0:
            self.mdeg[rid] = Mdeg
        return rid

    def step(self):
        return self.lib.evsl_amd_server_step(self.h)

    def status(self, rid):
        ts, t0, td = C.c_double(), C.c_double(), C.c_double()
        st = self.lib.evsl_amd_server_status(self.h, rid, C.byref(ts), C.byref(t0), C.byref(td))
        return st, ts.value, t0.value, td.value

    def stats(self):
        q, r = C.c_int(), C.c_int()
        m, pk = C.c_size_t(), C.c_size_t()
        self.lib.evsl_amd_server_stats(self.h, C.byref(q), C.byref(r), C.byref(m), C.byref(pk))
        return {"queued": q.value, "running": r.value, "mem_in_use": m.value, "mem_peak": pk.value}

    def slice_result(self, rid, n):
        nev = C.c_int()
        lam, Y, res = c_double_p(), c_double_p(), c_double_p()
        err = self.lib.evsl_amd_server_slice_result(self.h, rid, C.byref(nev), C.byref(lam),
                                                    C.byref(Y), C.byref(res))
        k = nev.value
        lam_h = np.array([lam[i] for i in range(k)]) if k else np.zeros(0)
        res_h = np.array([res[i] for i in range(k)]) if k else np.zeros(0)
        Yh = np.zeros((n, 0))
        if k:
            Yh = self.ev.from_device(C.cast(Y, C.c_void_p).value, n * k).reshape(k, n).T
            self.lib._evsl_Free_device(C.cast(Y, C.c_void_p))
        if lam:
            self.lib._evsl_Free(C.cast(lam, C.c_void_p))
        if res:
            self.lib._evsl_Free(C.cast(res, C.c_void_p))
        return err, lam_h, Yh, res_h

    def dos_result(self, rid):
        mu = np.zeros(self.mdeg[rid] + 1)
        ec = C.c_double()
        err = self.lib.evsl_amd_server_dos_result(self.h, rid, dptr(mu), C.byref(ec))
        return err, mu, ec.value
