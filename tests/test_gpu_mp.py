"""GPU parity of DSGESV (FP32 LU + FP64 refinement, FP64 fallback).

The reference's own driver test for it (TESTING/LIN/ddrvab.f:313-345) runs
DSGESV and accepts the solution when DGET08's ratio
||B - A X||_inf / (||A||_inf ||X||_inf eps) stays below the threshold 30.
DSGESV's stopping test (SRC/dsgesv.f:337-342) is that same ratio per column
bounded by SQRT(N); the bar here is MAX(30, 2*SQRT(N)), doubled because the
checker recomputes the residual with its own rounding (an error of the
same order as the residual being measured).
On top of that the GPU result is held to the oracle's:
  * ITER and INFO: the same branch of SRC/dsgesv.f is taken (overflow -2,
    FP32-singular -3, no convergence -31, else the iteration count).  The
    FP32 factors agree with the reference to rounding, not bitwise, so the
    iteration at which the stopping test first holds may move by one;
  * IPIV of the FP32 factorization (SGETRF) identical to the oracle's, up to
    a documented near-tie; IPIV of DGETRF on the fallback branches;
  * A unchanged when refinement succeeded (ITER >= 0).
"""
import numpy as np
import pytest

import lapack_amd as la
import oracle

from gpu_util import from_dev, ivec, to_dev
from test_gpu_lu import assert_ipiv

pytestmark = pytest.mark.gpu

THRESH = 30.0


def dget08_bar(n, it):
    """DDRVAB's bar (TESTING/LIN/ddrvab.f:362-365): DGET08 < sqrt(N) when the
    refinement claimed success (ITER >= 0), < THRESH after the FP64 fallback."""
    return np.sqrt(n) if it >= 0 else THRESH


def gpu_dsgesv(a, b):
    n, nrhs = b.shape
    d_a = to_dev(a)
    d_b = to_dev(b)
    d_x = to_dev(np.zeros((n, nrhs)))
    d_p = ivec(n)
    info, it = la.dsgesv(n, nrhs, d_a, n, d_p, d_b, n, d_x, n)
    return from_dev(d_x, (n, nrhs)), d_p.cpu().numpy()[:n], it, info, from_dev(d_a, (n, n))


@pytest.mark.parametrize("n,nrhs,seed", [(1, 1, 1), (17, 2, 2), (256, 1, 3), (600, 5, 4), (2048, 16, 5)])
def test_dsgesv_refines_like_reference(dev, n, nrhs, seed):
    a = la.matrix(3, 4000 + seed, n, n)
    b = la.matrix(3, 5000 + seed, n, nrhs)
    x, p, it, info, a_after = gpu_dsgesv(a, b)
    xr, pr, itr, infor, _ = oracle.dsgesv(a, b)
    assert info == infor == 0
    assert itr >= 0 and it >= 0 and abs(it - itr) <= 1, (it, itr)
    lu_s, _, _ = oracle.sgetrf(a)
    assert_ipiv(p, pr, lu_s.astype(np.float64), unit=2.0 ** -24)
    assert np.array_equal(a_after, a), "A must be unchanged when refinement succeeds"
    assert oracle.dget08(a, x, b) < dget08_bar(n, it)
    # forward agreement with the oracle's refined solution
    assert np.max(np.abs(x - xr)) <= 1e3 * np.finfo(np.float64).eps * np.linalg.cond(a, np.inf) * np.max(np.abs(xr))


def test_dsgesv_overflow_falls_back(dev):
    n = 300
    a = la.matrix(2, 3, n, n)
    a[3, 4] = 1e39
    b = la.matrix(2, 1, n, 2)
    x, p, it, info, lu = gpu_dsgesv(a, b)
    _, pr, itr, infor, lur = oracle.dsgesv(a, b)
    assert (it, info) == (itr, infor) == (-2, 0)
    assert_ipiv(p, pr, lur)
    assert oracle.dget01(a, lu, p) < THRESH and oracle.dget02(a, x, b) < THRESH


def test_dsgesv_rhs_overflow_falls_back(dev):
    n = 64
    a = la.matrix(2, 13, n, n)
    b = la.matrix(2, 14, n, 1)
    b[5, 0] = -4e38
    x, p, it, info, lu = gpu_dsgesv(a, b)
    assert (it, info) == (-2, 0)
    assert oracle.dget02(a, x, b) < THRESH


def test_dsgesv_fp32_singular(dev):
    n = 200
    a = la.matrix(2, 4, n, n)
    a[:, 7] = 0
    b = la.matrix(2, 1, n, 1)
    _, _, it, info, _ = gpu_dsgesv(a, b)
    _, _, itr, infor, _ = oracle.dsgesv(a, b)
    assert (it, info) == (itr, infor) == (-3, 8)


def test_dsgesv_no_convergence(dev):
    n = 120
    q1, _ = np.linalg.qr(la.matrix(3, 5, n, n))
    q2, _ = np.linalg.qr(la.matrix(3, 6, n, n))
    a = np.asfortranarray(q1 @ np.diag(np.logspace(0, 14, n)) @ q2.T)
    b = la.matrix(2, 1, n, 1)
    x, p, it, info, lu = gpu_dsgesv(a, b)
    _, pr, itr, infor, lur = oracle.dsgesv(a, b)
    assert (it, info) == (itr, infor) == (-31, 0)
    assert_ipiv(p, pr, lur)
    assert oracle.dget02(a, x, b) < THRESH


def test_dsgesv_host_pointers_and_zero_rhs(dev):
    n = 96
    a = la.matrix(2, 40, n, n)
    b = la.matrix(2, 41, n, 3)
    x = np.zeros((n, 3), order="F")
    p = np.zeros(n, dtype=np.int32)
    a_in = a.copy(order="F")
    info, it = la.dsgesv(n, 3, a_in, n, p, b, n, x, n)
    assert info == 0 and it >= 0
    assert np.array_equal(a_in, a)
    assert oracle.dget08(a, x, b) < THRESH
    # NRHS = 0 still factors in FP32 and reports ITER = 0 (SRC/dsgesv.f:337-348)
    p0 = np.zeros(n, dtype=np.int32)
    info, it = la.dsgesv(n, 0, a.copy(order="F"), n, p0, b, n, x, n)
    assert (info, it) == (0, 0)
    _, ps, _ = oracle.sgetrf(a)
    assert np.array_equal(p0, ps)


def test_dsgesv_argument_errors(dev):
    a = np.zeros((4, 4), order="F")
    b = np.zeros((4, 1), order="F")
    p = np.zeros(4, dtype=np.int32)
    assert la.dsgesv(-1, 1, a, 4, p, b, 4, b, 4)[0] == -1
    assert la.dsgesv(4, -1, a, 4, p, b, 4, b, 4)[0] == -2
    assert la.dsgesv(4, 1, a, 3, p, b, 4, b, 4)[0] == -4
    assert la.dsgesv(4, 1, a, 4, p, b, 3, b, 4)[0] == -7
    assert la.dsgesv(4, 1, a, 4, p, b, 4, b, 3)[0] == -9
