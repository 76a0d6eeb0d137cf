"""The drop-in claim from C: drivers with the reference's own call sequences
(TESTS/PLanN/LapPLanN.c, TESTS/Landos/LanDos.c), written against
include/evsl.h, compile with -Wall -Wextra -Werror and link against
libevsl_amd.so; on the GPU they run and check closed-form answers.  exDOS
(SRC/exDOS.c:17-69) is host code and is checked here against the smoothed
DOS in closed form."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

LIBDIR = os.path.join(ROOT, "evsl_amd")
DRIVERS = os.path.join(ROOT, "tests", "drivers")


def compile_driver(name, outdir):
    exe = os.path.join(str(outdir), name)
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), "-I", DRIVERS,
           os.path.join(DRIVERS, name + ".c"), "-o", exe,
           "-L", LIBDIR, "-levsl_amd", "-Wl,-rpath," + LIBDIR, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    out = tmp_path_factory.mktemp("drivers")
    return {n: compile_driver(n, out) for n in ("lap_plann", "landos_drv")}


@pytest.fixture(scope="module")
def lib():
    path = os.path.join(LIBDIR, "libevsl_amd.so")
    if not os.path.exists(path):
        import __graft_entry__
        __graft_entry__.build()
    return ctypes.CDLL(path)


def test_drivers_compile_and_link(drivers):
    for exe in drivers.values():
        assert os.access(exe, os.X_OK)


def smoothed_dos(vals, npts, intv):
    """exDOS's curve from its definition: Gaussians exp(-(x - t)^2 / (2 s^2)),
    s = H / sqrt(8 ln 1.25), H = (lmax - lmin) / (min(n, 30) - 1), cut at
    s sqrt(-2 ln 1e-8), for the eigenvalues t in [a, b], scaled by
    1 / (n sqrt(2 pi s^2)).  Each Gaussian is placed on the points of the
    reference's walk: from grid index int((t - a - width) / h) (at least 0),
    x advanced by repeated addition of h while x <= min(t + width, b)."""
    a, b, lm, lM = intv
    n = len(vals)
    s = (lM - lm) / (min(n, 30) - 1) / np.sqrt(8 * np.log(1.25))
    s2 = 2 * s * s
    width = s * np.sqrt(-2 * np.log(1e-8))
    x = np.linspace(a, b, npts)
    x[0], x[-1] = a, b
    h = x[1] - x[0]
    y = np.zeros(npts)
    for t in vals:
        if not a <= t <= b:
            continue
        j = max(int((t - a - width) / h), 0)
        xi = a + j * h
        while xi <= min(t + width, b) and j < npts:
            y[j] += np.exp(-(xi - t) * (xi - t) / s2)
            j += 1
            xi += h
    return x, y / (n * np.sqrt(s2 * np.pi))


def test_exdos_closed_form(lib):
    L = lib
    dp = ctypes.POINTER(ctypes.c_double)
    L.exDOS.restype = ctypes.c_int
    L.exDOS.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, dp, dp]
    rng = np.random.default_rng(4)
    for n, npts, intv in ((500, 200, [0.0, 12.0, 0.0, 12.0]),
                          (37, 311, [1.0, 3.5, -0.5, 4.0]),
                          (12, 64, [-0.2, 0.9, -1.0, 1.0])):
        vals = np.sort(rng.uniform(intv[2], intv[3], n))
        x, y = np.empty(npts), np.empty(npts)
        iv = np.array(intv, dtype=np.float64)
        ret = L.exDOS(vals.ctypes.data_as(dp), n, npts, x.ctypes.data_as(dp),
                      y.ctypes.data_as(dp), iv.ctypes.data_as(dp))
        assert ret == 0
        xe, ye = smoothed_dos(vals, npts, intv)
        np.testing.assert_allclose(x, xe, rtol=0, atol=1e-14 * max(1, abs(intv[1])))
        np.testing.assert_allclose(y, ye, rtol=0, atol=1e-14 * ye.max())
        # against the plain grid curve (every point within the cut-off), the
        # walk differs only by the Gaussian tails it starts early (< 1e-8)
        # and, at x = b, where the accumulated x passes b
        xg = np.linspace(intv[0], intv[1], npts)
        s = (intv[3] - intv[2]) / (min(n, 30) - 1) / np.sqrt(8 * np.log(1.25))
        yg = sum(np.exp(-(xg - t) ** 2 / (2 * s * s)) * (np.abs(xg - t) <= s * np.sqrt(-2 * np.log(1e-8)))
                 for t in vals if intv[0] <= t <= intv[1]) / (n * np.sqrt(2 * s * s * np.pi))
        assert np.max(np.abs(y - yg)[:-1]) < 1e-8 * ye.max()
    # no eigenvalue in [a, b]: return 1
    vals = np.array([5.0, 6.0])
    iv = np.array([0.0, 1.0, 0.0, 8.0])
    x, y = np.empty(50), np.empty(50)
    assert L.exDOS(vals.ctypes.data_as(dp), 2, 50, x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                   iv.ctypes.data_as(dp)) == 1


def run(exe, *args):
    print("driver:", os.path.basename(exe), *args, flush=True)
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:])
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("args", [(30, 30, 1, 0.4, 0.8, 2), (12, 12, 12, 1.0, 1.6, 3)])
def test_lap_plann_driver(drivers, args):
    r = run(drivers["lap_plann"], *args)
    assert r.returncode == 0 and "PASS" in r.stdout


@pytest.mark.gpu
def test_landos_driver(drivers):
    r = run(drivers["landos_drv"], 20)
    assert r.returncode == 0 and "PASS" in r.stdout
