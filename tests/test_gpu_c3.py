"""The mixed-precision workload at its specified sizes, on the GPU:

  * A = Q1 diag(s) Q2^T at n = 8192, Q1 and Q2 from Householder QR of seeded
    Gaussian matrices, s geometric from 1 to 1e6 (kappa = 1e6), factored in
    FP32 ('S') and with FP16 / BF16 trailing updates after DGEEQUB ('H', 'B'),
    refined in FP64 with the GMRES-IR fallback;
  * A, b i.i.d. N(0,1) at n = 32768 through DSGESV (SRC/dsgesv.f) and the
    FP32 refinement of lapack_amd_dgesv_ir.

Bars, for every variant that ends with ITER >= 0:
  * DSGESV's stopping test per column, max|r| <= max|x| ||A||_inf eps sqrt(n)
    (SRC/dsgesv.f:337-342), on the backward error the solver returns;
  * DDRVAB's bar: DGET08 ||b - A x||_inf / (||A||_inf ||x||_inf eps) < sqrt(n)
    (TESTING/LIN/ddrvab.f:362-365), recomputed here in FP64 independently of
    the solver (numpy at n = 8192, torch FP64 products on the device at
    n = 32768);
  * the workload's target: normwise backward error <= n * eps64;
  * forward error against the FP64 solution within kappa * n * eps.  The FP64
    solution is the library's plain FP64 DGESV (whose IPIV the tests pin to
    the reference DGETRF at n = 1024 and 16384), cross-checked at n = 8192
    against numpy's LAPACK solve (an independent FP64 solver).
The iteration counts (refinement steps, GMRES inner iterations) are printed.
"""
import numpy as np
import pytest

import lapack_amd as la

from gpu_util import from_dev, ivec, to_dev
from test_gpu_ir import conditioned

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


@pytest.fixture(scope="module")
def c3_conditioned():
    n = 8192
    a = conditioned(n, 1e6, 8192, mode=3)
    b = la.matrix(3, 8194, n, 1)
    x_np = np.linalg.solve(a, b)
    return a, b, x_np


def fp64_dgesv(a, b):
    n, nrhs = b.shape
    d_a, d_b = to_dev(a), to_dev(b)
    d_p = ivec(n)
    assert la.dgesv(n, nrhs, d_a, n, d_p, d_b, n) == 0
    return from_dev(d_b, (n, nrhs))


def dget08_np(a, x, b):
    r = b - a @ x
    return np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max() * EPS)


@pytest.mark.parametrize("prec", ["S", "H", "B"])
def test_conditioned_n8192_geometric(dev, c3_conditioned, prec):
    a, b, x_np = c3_conditioned
    n = a.shape[0]
    d_a, d_b = to_dev(a), to_dev(b)
    d_x = to_dev(np.zeros((n, 1)))
    d_p = ivec(n)
    berr = np.zeros(1)
    info, it, git = la.dgesv_ir(n, 1, d_a, n, d_p, d_b, n, d_x, n, prec, berr)
    x = from_dev(d_x, (n, 1))
    assert info == 0
    assert it >= 0, (prec, it, git)
    assert berr[0] <= EPS * np.sqrt(n) * (1 + 1e-12)
    r08 = dget08_np(a, x, b)
    assert r08 < np.sqrt(n), r08
    assert r08 * EPS <= n * EPS
    x64 = fp64_dgesv(a, b)
    kappa = 1e6
    fwd = np.abs(x - x64).max() / np.abs(x64).max()
    fwd_np = np.abs(x - x_np).max() / np.abs(x_np).max()
    assert fwd <= kappa * n * EPS and fwd_np <= kappa * n * EPS, (fwd, fwd_np)
    print(f"n={n} kappa=1e6 geometric prec {prec}: ITER {it}, GMRES iterations {git}, "
          f"backward error {r08 * EPS:.3e} (DGET08 {r08:.2f}), forward error vs FP64 DGESV {fwd:.3e}, "
          f"vs numpy {fwd_np:.3e}")


def test_dsgesv_n32768_normal(dev):
    import torch
    n = 32768
    a = la.matrix(3, 32768, n, n)
    b = la.matrix(3, 32769, n, 1)
    d_a0 = torch.from_numpy(a.T).cuda()   # row-major view of the column-major bytes
    del a
    d_b = torch.from_numpy(b.T.copy()).cuda()
    anorm = d_a0.abs().sum(dim=0).max().item()   # ||A||_inf: column sums of A^T

    def dget08(x):
        r = d_b - x @ d_a0
        return (r.abs().max() / (anorm * x.abs().max() * EPS)).item()

    d_a = torch.empty_like(d_a0)
    p = ivec(n)
    d_a.copy_(d_a0)
    x64 = d_b.clone()
    assert la.dgesv(n, 1, d_a, n, p, x64, n) == 0
    r64 = dget08(x64)
    assert r64 < 30.0
    d_a.copy_(d_a0)
    x = torch.zeros_like(d_b)
    info, it = la.dsgesv(n, 1, d_a, n, p, d_b, n, x, n)
    assert info == 0 and it >= 0, (info, it)
    r08 = dget08(x)
    assert r08 < np.sqrt(n), r08
    fwd = ((x - x64).abs().max() / x64.abs().max()).item()
    assert fwd <= 1e-6, fwd
    d_a.copy_(d_a0)
    x2 = torch.zeros_like(d_b)
    berr = np.zeros(1)
    info, it2, git = la.dgesv_ir(n, 1, d_a, n, p, d_b, n, x2, n, "S", berr)
    assert info == 0 and it2 >= 0
    r08b = dget08(x2)
    assert r08b < np.sqrt(n) and berr[0] <= EPS * np.sqrt(n) * (1 + 1e-12)
    print(f"n=32768 N(0,1): DGESV DGET08 {r64:.2f}; DSGESV ITER {it}, DGET08 {r08:.2f}, forward error {fwd:.3e}; "
          f"FP32 IR ITER {it2}, GMRES {git}, DGET08 {r08b:.2f}")
