"""Writes the golden vectors under tests/golden/ from the reference itself.

The reference routines are the flang build of the reference's own sources
(oracle/_ref/libreflapack.so, made by `make -C oracle ref`); the inputs come
from the library's seeded generator (lapack_amd_dlarnv), so every file here
can be regenerated bit for bit:

  dgetrf_n1024_u.npz   n = 1024, A uniform(-1,1) seed 1988, b seed 1989:
                       reference DGETRF IPIV and INFO, DGETRS('N') x
  dgetrf_n16384_normal_ipiv.npy
                       n = 16384, A N(0,1) seed 16384 (the benchmark's matrix):
                       reference DGETRF IPIV (about 12 minutes on one core)

Usage: python tests/golden/make_golden.py [--large]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import lapack_amd as la  # noqa: E402  (generator only: no GPU needed)
import oracle  # noqa: E402


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref is not built (needs the reference sources): make -C oracle ref")
    n = 1024
    a = la.matrix(2, 1988, n, n)
    b = la.matrix(2, 1989, n, 1)
    lu, ipiv, info = oracle.ref_dgetrf(a)[:3]
    x, _ = oracle.ref_dgetrs(lu, ipiv, b)
    np.savez(os.path.join(HERE, "dgetrf_n1024_u.npz"), ipiv=np.asarray(ipiv, np.int32), info=np.int32(info),
             x=np.asarray(x[:, 0], np.float64))
    print("wrote dgetrf_n1024_u.npz")
    if "--large" in sys.argv:
        n = 16384
        a = la.matrix(3, 16384, n, n)
        _, ipiv, info = oracle.ref_dgetrf(a)[:3]
        assert info == 0
        np.save(os.path.join(HERE, "dgetrf_n16384_normal_ipiv.npy"), np.asarray(ipiv, np.int32))
        print("wrote dgetrf_n16384_normal_ipiv.npy")


if __name__ == "__main__":
    main()
