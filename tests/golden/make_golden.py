"""Writes the golden fixtures of the test suite (run from the repository root:
python tests/golden/make_golden.py).

A.mtx is the reference's own test matrix (MATRICES/A.mtx, read by
TESTS/PLanN/MMPLanN.c with matfile line "A.mtx stiff1 MM1 0.1 0.3 5"), kept
as input data.  Its expected outputs come from LAPACK through numpy:
A_eigs.npy holds every eigenvalue (ascending) of the symmetric matrix the
file defines, so interval counts and extreme eigenvalues are fixed
independently of any Lanczos code."""
import os

import numpy as np
import scipy.io

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    A = scipy.io.mmread(os.path.join(HERE, "A.mtx")).toarray()
    assert np.array_equal(A, A.T)
    w = np.linalg.eigvalsh(A)
    np.save(os.path.join(HERE, "A_eigs.npy"), w)
    print(f"A.mtx: n {A.shape[0]}, eigenvalues [{w[0]:.15e}, {w[-1]:.15e}], "
          f"{np.count_nonzero((w >= 0.1) & (w <= 0.3))} in [0.1, 0.3]")


if __name__ == "__main__":
    main()
