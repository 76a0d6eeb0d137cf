import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def libc():
    return ctypes.CDLL(None)


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import LIB_PATH, Oracle, build
    if not os.path.exists(LIB_PATH):
        build()
    return Oracle()


def libc_rand_double(libc, n):
    """U[-1,1] from libc rand(), the mapping of SRC/vect.c:14-20."""
    rmax = 2147483647.0  # glibc RAND_MAX
    t = rmax / 2.0
    r = np.fromiter((libc.rand() for _ in range(n)), dtype=np.float64, count=n)
    return (r - t) / t


@pytest.fixture(scope="session")
def ev():
    """The product library with a started context (GPU tests only)."""
    from evsl_amd.evsl import EVSL
    lib = EVSL()
    assert lib.start() == 0, lib.last_error()
    yield lib
    lib.finish()
