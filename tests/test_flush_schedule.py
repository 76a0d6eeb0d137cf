"""The y-flush schedule of the fused Chebyshev step (evsl_dev::y_flush in
evsl_amd/csrc/dev.h), compiled from the header itself: for every degree m
the launches k = 1..m add each term T_1..T_m of the series exactly once, in
order, never look further back than the two previous degrees, and the first
launch sets y.  cheb_vec_bytes counts, per row and column, xk (gathered
once), xkm1 and xkp1 of the vector size plus the FP64 y in and out on a
flush (the first launch: xk, xkp1 and y out)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include <cstdio>
#include "dev.h"
int main() {
  for (int m = 1; m <= 40; ++m)
    for (int k = 1; k <= m; ++k)
      std::printf("%d %d %d %g %g\n", m, k, evsl_dev::y_flush(k, m),
                  evsl_dev::cheb_vec_bytes(evsl_dev::y_flush(k, m), k == 1, 4),
                  evsl_dev::cheb_vec_bytes(evsl_dev::y_flush(k, m), k == 1, 8));
  return 0;
}
'''


@pytest.fixture(scope="module")
def schedule(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("flush")
    src, exe = d / "t.cpp", d / "t"
    src.write_text(SRC)
    subprocess.run([gxx, "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "evsl_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = [tuple(int(t) for t in line.split()[:3]) + tuple(float(t) for t in line.split()[3:])
            for line in out.splitlines()]
    return rows


def test_each_term_added_once_in_order(schedule):
    by_m = {}
    for m, k, yf, b4, b8 in schedule:
        by_m.setdefault(m, []).append((k, yf, b4, b8))
    for m, launches in by_m.items():
        added = []
        for k, yf, b4, b8 in launches:
            assert 0 <= yf <= 3
            if k == 1:
                # y = mu0 x + mu1 T_1: reads xk, writes xkp1 and the FP64 y
                assert yf == 1 and b4 == 2 * 4 + 8 and b8 == 3 * 8
            assert yf <= k  # look-back stays within the series
            added.extend(range(k - yf + 1, k + 1))
            if k > 1:
                assert b4 == 3 * 4 + (16 if yf else 0)
                assert b8 == 3 * 8 + (16 if yf else 0)
        assert added == list(range(1, m + 1)), (m, added)
        # every launch but the first and the last flushes at most every third degree
        flushes = [k for k, yf, _, _ in launches if yf]
        assert flushes[-1] == m
