"""ctypes mirrors of the C interface's data types (include/evsl.h).

The layouts follow EVSL 1.1.1's INC/struct.h:14-98 (CPU layout), so the same
struct can be handed to libevsl_amd.so and to any other library built
against that header.
"""
import ctypes as C

import numpy as np

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)


class CsrMat(C.Structure):
    """csrMat (INC/struct.h:29-42, CPU layout)."""
    _fields_ = [("owndata", C.c_int), ("nrows", C.c_int), ("ncols", C.c_int),
                ("nnz", C.c_int), ("ia", c_int_p), ("ja", c_int_p),
                ("a", c_double_p)]


class CooMat(C.Structure):
    """cooMat (INC/struct.h:14-21)."""
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("nnz", C.c_int),
                ("ir", c_int_p), ("jc", c_int_p), ("vv", c_double_p)]


class PolParams(C.Structure):
    """polparams (INC/struct.h:68-98)."""
    _fields_ = [("max_deg", C.c_int), ("min_deg", C.c_int), ("damping", C.c_int),
                ("thresh_ext", C.c_double), ("thresh_int", C.c_double),
                ("intvtol", C.c_double), ("type", C.c_int), ("mu", c_double_p),
                ("cc", C.c_double), ("dd", C.c_double), ("gam", C.c_double),
                ("bar", C.c_double), ("deg", C.c_int)]

    def coefficients(self):
        """mu[0..deg] as a numpy copy."""
        return np.ctypeslib.as_array(self.mu, shape=(self.deg + 1,)).copy()


MVFUNC = C.CFUNCTYPE(None, c_double_p, c_double_p, C.c_void_p)


def dptr(a):
    """double* view of a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    """int* view of a C-contiguous int32 numpy array."""
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int_p)


class HostCsr:
    """A CSR matrix held in numpy arrays together with its csrMat view.

    The arrays stay referenced by this object for as long as the csrMat is
    in use (SetAMatrix stores the pointer, as the reference does,
    SRC/evsl.c:95-101)."""

    def __init__(self, n, ia, ja, a):
        self.n = int(n)
        self.ia = np.ascontiguousarray(ia, dtype=np.int32)
        self.ja = np.ascontiguousarray(ja, dtype=np.int32)
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.csr = CsrMat(0, self.n, self.n, int(self.ia[-1]), iptr(self.ia),
                          iptr(self.ja), dptr(self.a))

    @property
    def nnz(self):
        return int(self.ia[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.a, self.ja, self.ia), shape=(self.n, self.n))

    @classmethod
    def from_scipy(cls, m):
        m = m.tocsr()
        m.sort_indices()
        return cls(m.shape[0], m.indptr, m.indices, m.data)
