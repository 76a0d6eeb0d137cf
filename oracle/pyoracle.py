"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check the library's results (and, in the
cpu_baseline leg, to time the CPU path).  The product package evsl_amd never
imports this module.
"""
import ctypes as C
import os

import numpy as np

from evsl_amd.types import MVFUNC, PolParams, c_double_p, c_int_p, dptr, iptr

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Oracle:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = L = C.CDLL(path)
        D, I, V = C.c_double, C.c_int, None
        dp, ip = c_double_p, c_int_p
        pp = C.POINTER(PolParams)
        ddp = C.POINTER(c_double_p)
        sig = {
            "ora_set_csr": (V, [I, ip, ip, dp]),
            "ora_set_matvec": (V, [I, MVFUNC, C.c_void_p]),
            "ora_matvec": (V, [dp, dp]),
            "ora_rand_double": (V, [I, dp]),
            "ora_randn_double": (V, [I, dp]),
            "ora_linspace": (V, [D, D, I, dp]),
            "ora_dampcf": (I, [I, I, dp]),
            "ora_chebxPltd": (I, [I, dp, I, dp, dp]),
            "ora_set_pol_def": (V, [pp]),
            "ora_find_pol": (I, [dp, pp]),
            "ora_free_pol": (V, [pp]),
            "ora_ChebAv": (I, [pp, dp, dp, dp]),
            "ora_kpmdos": (I, [I, I, I, dp, dp, dp]),
            "ora_intChx": (V, [I, dp, I, dp, dp]),
            "ora_spslicer": (I, [dp, dp, I, dp, I, I]),
            "ora_CGS_DGKS": (V, [I, I, I, dp, dp, dp, dp]),
            "ora_tridiag_eig": (I, [I, dp, dp, dp, dp]),
            "ora_sym_eig": (I, [I, dp, I, dp, I, dp]),
            "ora_ChebLanNr": (I, [dp, I, D, dp, pp, ip, ddp, ddp, ddp]),
            "ora_ChebSI": (I, [I, dp, I, D, dp, pp, ip, ddp, ddp, ddp]),
            "ora_ChebLanTr": (I, [I, I, dp, I, D, dp, pp, ip, ddp, ddp, ddp]),
            "ora_LanBounds": (I, [I, dp, dp, dp]),
            "ora_LanTrbounds": (I, [I, I, D, dp, I, dp, dp]),
            "ora_LanDos": (I, [I, I, I, dp, dp, dp, dp]),
            "ora_simpson": (V, [dp, dp, I]),
            "ora_spslicer2": (V, [dp, dp, I, I, dp]),
            "ora_free": (V, [C.c_void_p]),
            "ora_timer": (D, []),
            "ora_last_lanczos_steps": (I, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self._keep = None

    # ---- operator -----------------------------------------------------
    def set_csr(self, A):
        """A: evsl_amd.types.HostCsr (kept alive by the caller)."""
        self._keep = A
        self.lib.ora_set_csr(A.n, iptr(A.ia), iptr(A.ja), dptr(A.a))
        self.n = A.n

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.lib.ora_matvec(dptr(x), dptr(y))
        return y

    # ---- vectors --------------------------------------------------------
    def rand_double(self, n):
        v = np.empty(n)
        self.lib.ora_rand_double(n, dptr(v))
        return v

    def randn_double(self, n):
        v = np.empty(n)
        self.lib.ora_randn_double(n, dptr(v))
        return v

    # ---- filters --------------------------------------------------------
    def dampcf(self, m, damping):
        jac = np.empty(m + 1)
        self.lib.ora_dampcf(m, damping, dptr(jac))
        return jac

    def chebxPltd(self, mu, x):
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
        y = np.empty_like(x)
        self.lib.ora_chebxPltd(len(mu) - 1, dptr(mu), len(x), dptr(x), dptr(y))
        return y

    def default_pol(self):
        pol = PolParams()
        self.lib.ora_set_pol_def(C.byref(pol))
        return pol

    def find_pol(self, intv, pol):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        return self.lib.ora_find_pol(dptr(intv), C.byref(pol))

    def free_pol(self, pol):
        self.lib.ora_free_pol(C.byref(pol))

    def ChebAv(self, pol, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        y = np.empty_like(v)
        w = np.empty(3 * len(v))
        self.lib.ora_ChebAv(C.byref(pol), dptr(v), dptr(y), dptr(w))
        return y

    # ---- DOS / slicing ----------------------------------------------------
    def kpmdos(self, Mdeg, damping, nvec, intv):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        mu = np.empty(Mdeg + 1)
        ecnt = C.c_double()
        err = self.lib.ora_kpmdos(Mdeg, damping, nvec, dptr(intv), dptr(mu), C.byref(ecnt))
        return err, mu, ecnt.value

    def intChx(self, mu, xi):
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        yi = np.empty_like(xi)
        self.lib.ora_intChx(len(mu) - 1, dptr(mu), len(xi), dptr(xi), dptr(yi))
        return yi

    def spslicer(self, mu, intv, n_int, npts):
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        sli = np.empty(n_int + 1)
        err = self.lib.ora_spslicer(dptr(sli), dptr(mu), len(mu) - 1, dptr(intv), n_int, npts)
        return err, sli

    def LanDos(self, nvec, msteps, npts, intv):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        x, y, neig = np.empty(npts), np.empty(npts), C.c_double()
        err = self.lib.ora_LanDos(nvec, msteps, npts, dptr(x), dptr(y), C.byref(neig), dptr(intv))
        return err, x, y, neig.value

    def simpson(self, xi, yi):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        yi = np.ascontiguousarray(yi, dtype=np.float64).copy()
        self.lib.ora_simpson(dptr(xi), dptr(yi), len(xi))
        return yi

    def spslicer2(self, xi, yi, n_int):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        yi = np.ascontiguousarray(yi, dtype=np.float64).copy()
        sli = np.empty(n_int + 1)
        self.lib.ora_spslicer2(dptr(xi), dptr(yi), n_int, len(xi), dptr(sli))
        return sli

    # ---- dense ------------------------------------------------------------
    def CGS_DGKS(self, Q, v, i_max=2):
        """Q: (n, k) array; returns (v_out, beta)."""
        Qf = np.asfortranarray(Q, dtype=np.float64)
        n, k = Qf.shape
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        w = np.empty(max(k, 1))
        nrm = C.c_double()
        self.lib.ora_CGS_DGKS(n, k, i_max, Qf.ctypes.data_as(c_double_p), dptr(v),
                              C.byref(nrm), dptr(w))
        return v, nrm.value

    def tridiag_eig(self, d, e, vectors=False):
        d = np.ascontiguousarray(d, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        n = len(d)
        ev = np.empty(n)
        Z = np.empty((n, n), order="F") if vectors else None
        info = self.lib.ora_tridiag_eig(
            n, dptr(d), dptr(e), dptr(ev),
            Z.ctypes.data_as(c_double_p) if vectors else None)
        return (info, ev, Z) if vectors else (info, ev)

    def sym_eig(self, A):
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        Q = np.empty((n, n), order="F")
        lam = np.empty(n)
        info = self.lib.ora_sym_eig(n, A.ctypes.data_as(c_double_p), n,
                                    Q.ctypes.data_as(c_double_p), n, dptr(lam))
        return info, lam, Q

    # ---- solvers ----------------------------------------------------------
    def _take(self, n, nev, lam, W, res):
        if nev > 0:
            out = (np.ctypeslib.as_array(lam, shape=(nev,)).copy(),
                   np.ctypeslib.as_array(W, shape=(nev, n)).copy().T,
                   np.ctypeslib.as_array(res, shape=(nev,)).copy())
        else:
            out = (np.empty(0), np.empty((n, 0)), np.empty(0))
        for p in (lam, W, res):
            if p:
                self.lib.ora_free(C.cast(p, C.c_void_p))
        return out

    def ChebLanNr(self, intv, maxit, tol, vinit, pol):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        vinit = np.ascontiguousarray(vinit, dtype=np.float64)
        nev = C.c_int()
        lam, W, res = c_double_p(), c_double_p(), c_double_p()
        err = self.lib.ora_ChebLanNr(dptr(intv), maxit, tol, dptr(vinit), C.byref(pol),
                                     C.byref(nev), C.byref(lam), C.byref(W), C.byref(res))
        if err:
            return err, None, None, None
        return (0,) + self._take(len(vinit), nev.value, lam, W, res)

    def ChebSI(self, nev, intv, maxit, tol, V0, pol):
        """V0: n x nev start block."""
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        V0 = np.asfortranarray(V0, dtype=np.float64)
        nevo = C.c_int()
        lam, W, res = c_double_p(), c_double_p(), c_double_p()
        err = self.lib.ora_ChebSI(nev, dptr(intv), maxit, tol, V0.ctypes.data_as(c_double_p),
                                  C.byref(pol), C.byref(nevo), C.byref(lam), C.byref(W),
                                  C.byref(res))
        if err:
            return err, None, None, None
        return (0,) + self._take(V0.shape[0], nevo.value, lam, W, res)

    def ChebLanTr(self, lanm, nev, intv, maxit, tol, vinit, pol):
        intv = np.ascontiguousarray(intv, dtype=np.float64)
        vinit = np.ascontiguousarray(vinit, dtype=np.float64)
        nev2 = C.c_int()
        lam, W, res = c_double_p(), c_double_p(), c_double_p()
        err = self.lib.ora_ChebLanTr(lanm, nev, dptr(intv), maxit, tol, dptr(vinit),
                                     C.byref(pol), C.byref(nev2), C.byref(lam),
                                     C.byref(W), C.byref(res))
        if err:
            return err, None, None, None
        return (0,) + self._take(len(vinit), nev2.value, lam, W, res)

    def LanBounds(self, msteps, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lmin, lmax = C.c_double(), C.c_double()
        err = self.lib.ora_LanBounds(msteps, dptr(v), C.byref(lmin), C.byref(lmax))
        return err, lmin.value, lmax.value

    def LanTrbounds(self, lanm, maxit, tol, v, bndtype):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lmin, lmax = C.c_double(), C.c_double()
        err = self.lib.ora_LanTrbounds(lanm, maxit, tol, dptr(v), bndtype, C.byref(lmin),
                                       C.byref(lmax))
        return err, lmin.value, lmax.value

    def last_lanczos_steps(self):
        return self.lib.ora_last_lanczos_steps()

    def timer(self):
        return self.lib.ora_timer()
