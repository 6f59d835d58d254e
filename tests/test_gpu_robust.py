"""GPU checks of the edge cases and failure paths of the LU path, against the
oracle (the reference's loop order, bit-identical to the flang build of the
reference):

  * the IDAMAX rule on NaN columns (BLAS/SRC/idamax.f:101-107: a NaN at the
    head of the column wins, NaNs elsewhere never do);
  * the |pivot| < sfmin branch of DGETRF2 (SRC/dgetrf2.f:204-210: divide
    instead of multiplying by the reciprocal) -- checked bitwise on the
    column it produces;
  * the M/N/NB grid of TESTING/dtest.in:2-9;
  * matrix kinds of DCHKGE's types 1-4 and 8-11 (TESTING/LIN/dlatb4.f:437-477:
    diagonal, triangular, cond 2, cond sqrt(0.1/eps) and 0.1/eps, scaled
    near underflow and overflow), generated here with the same singular
    value spectra (MODE = 3, geometric) and norms;
  * panels taller than the register-resident kernels hold (m > 32768) and the
    rectangular shape with m > 16384 where the multi-workgroup panel runs
    beside the update;
  * the bounded waits: every wait of a factorization or solve is made to give
    up from the host (lapack_amd_test_fail_waits), no kernel hangs; the
    asynchronous call leaves A / B exactly as passed and reports through
    INFO and the device's sticky status; the synchronous call redoes the work
    with the forms that never wait and returns the reference's result.
Bars: IPIV identical (or a documented near-tie), INFO equal, DGET01 / DGET02
below 30 (TESTING/dtest.in:13).
"""
import os

import numpy as np
import pytest

import lapack_amd as la
import oracle

from gpu_util import from_dev, ivec, to_dev
from test_gpu_lu import THRESH, assert_ipiv, gpu_getrf

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def test_nan_columns_follow_idamax(dev):
    """NaNs off the head of a column are never chosen; a NaN that reaches the
    head wins (and then every later column is NaN): same IPIV as the reference."""
    for n in (50, 1500):
        a = la.matrix(2, 404 + n, n, n)
        a[7, 3] = np.nan           # below the head of column 3: ignored by the search
        a[n - 1, 10] = np.nan
        lu_r, p_r, info_r = oracle.dgetrf(a)
        lu, p, info = gpu_getrf(a)
        assert info == info_r
        assert np.array_equal(p, p_r), np.nonzero(p != p_r)[0][:5]
        # where the reference's factors are NaN, so are the GPU's
        assert np.array_equal(np.isnan(lu), np.isnan(lu_r))


def test_nan_at_head_wins(dev):
    n = 40
    a = la.matrix(2, 405, n, n)
    a[0, 0] = np.nan
    lu_r, p_r, info_r = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert p_r[0] == 1 and np.array_equal(p, p_r)
    assert info == info_r


@pytest.mark.parametrize("n", [40, 1500])
def test_subnormal_pivot_divides(dev, n):
    """Column 0 scaled to subnormal values: its pivot is below sfmin = 2^-1022,
    so the multipliers are x / pivot (SRC/dgetrf2.f:207-210), bitwise the
    reference's; the rest of the factorization keeps its IPIV and ratios."""
    a = la.matrix(2, 406 + n, n, n)
    a[:, 0] *= 2.0 ** -1060
    assert 0 < np.abs(a[:, 0]).max() < 2.0 ** -1022
    lu_r, p_r, info_r = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info == info_r == 0
    assert np.array_equal(p, p_r)
    # the multipliers of the first column: one IEEE division each, as the reference
    assert np.array_equal(lu[1:, 0].view(np.int64), lu_r[1:, 0].view(np.int64))
    # the reciprocal would differ: (1/p)*x != x/p somewhere for these values
    piv = lu_r[0, 0]
    recip = (1.0 / piv) * a[:, 0]
    assert not np.isfinite(1.0 / piv) or not np.array_equal(recip, a[:, 0] / piv)


def test_dtest_in_grid(dev):
    """TESTING/dtest.in:2-9: M, N in {0,1,2,3,5,10,50}, NB in {1,3,20} plus the
    default, uniform(-1,1) entries: IPIV, INFO and DGET01 against the reference."""
    vals = [0, 1, 2, 3, 5, 10, 50]
    try:
        for nb in (1, 3, 20, 0):
            la.set_tuning(nb, 0)
            for m in vals:
                for n in vals:
                    a = la.matrix(2, 1988 + 100 * m + n, max(m, 1), max(n, 1))[:m, :n]
                    a = np.asfortranarray(a)
                    if m == 0 or n == 0:
                        ip = np.full(max(1, min(m, n)), -7, dtype=np.int32)
                        assert la.dgetrf(m, n, np.zeros(1), max(1, m), ip) == 0
                        continue
                    lu_r, p_r, info_r = oracle.dgetrf(a)
                    lu, p, info = gpu_getrf(a)
                    assert info == info_r, (m, n, nb)
                    assert np.array_equal(p, p_r), (m, n, nb)
                    assert oracle.dget01(a, lu, p) < THRESH, (m, n, nb)
    finally:
        la.set_tuning(0, 0)


def dchkge_kind(imat, n, seed):
    """A matrix of DCHKGE type imat (TESTING/LIN/dlatb4.f:437-477): singular
    values geometric from 1 to 1/cond (DLATMS MODE = 3), scaled to ANORM;
    type 1 diagonal, 2 upper, 3 lower triangular, others dense."""
    eps = 2.0 ** -53
    badc2 = 0.1 / eps
    badc1 = np.sqrt(badc2)
    small = 0.25 * (2.0 ** -1022 / eps)
    large = 1.0 / small
    cond = badc1 if imat == 8 else badc2 if imat == 9 else 2.0
    anorm = small if imat == 10 else large if imat == 11 else 1.0
    rng = np.random.default_rng(seed)
    d = cond ** (-np.arange(n) / max(n - 1, 1))
    d *= rng.choice([-1.0, 1.0], n)
    if imat == 1:
        a = np.diag(d)
    elif imat in (2, 3):
        t = np.triu(rng.uniform(-1, 1, (n, n)), 1) / n
        a = t + np.diag(d)
        if imat == 3:
            a = a.T
    else:
        q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
        q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        a = (q1 * d) @ q2.T
    a = a * (anorm / np.abs(a).max())
    return np.asfortranarray(a)


@pytest.mark.parametrize("n", [300, 1500])
@pytest.mark.parametrize("imat", [1, 2, 3, 4, 8, 9, 10, 11])
def test_dchkge_matrix_kinds(dev, imat, n):
    a = dchkge_kind(imat, n, 1988 + imat + n)
    b = np.asfortranarray(a @ la.matrix(2, 77 + imat, n, 2))   # B = A * XACT, as DLARHS
    lu_r, p_r, info_r = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info == info_r
    assert_ipiv(p, p_r, lu_r, lu)
    assert oracle.dget01(a, lu, p) < THRESH
    d_lu, d_p, d_b = to_dev(lu), to_dev(p.reshape(-1, 1).astype(np.int32)), to_dev(b)
    assert la.dgetrs("N", n, 2, d_lu, n, d_p, d_b, n) == 0
    x = from_dev(d_b, (n, 2))
    assert oracle.dget02(a, x, b) < THRESH


@pytest.mark.parametrize("m,n", [(20000, 512), (40000, 192)])
def test_tall_rectangular(dev, m, n):
    """m > 16384: the multi-workgroup panel beside the update without the XCD
    split; m > 32768: leaves taller than 64 register-resident workgroups take
    the column-at-a-time form."""
    a = la.matrix(3, 7000 + m + n, m, n)
    lu_r, p_r, info_r = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info == info_r == 0
    assert np.array_equal(p, p_r)
    assert oracle.dget01(a, lu, p) < THRESH


def test_panel_without_xcd_split(dev, monkeypatch):
    """LAPACK_AMD_PANEL_XCD=0: the multi-workgroup panel shares the device
    with the update instead of owning XCD 0; same IPIV as the reference."""
    monkeypatch.setenv("LAPACK_AMD_PANEL_XCD", "0")
    n = 3000
    a = la.matrix(3, 3001, n, n)
    lu_r, p_r, _ = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info == 0
    assert np.array_equal(p, p_r)
    assert oracle.dget01(a, lu, p) < THRESH


def test_quick_return_writes_info(dev):
    """M = 0 or N = 0: INFO = 0 on the device (SRC/dgetrf.f:161), whatever it held."""
    d_info = ivec(1)
    d_info.fill_(77)
    d_a = ivec(4).double()
    assert la.dgetrf_async(0, 5, d_a, 1, ivec(1), d_info) == 0
    assert int(d_info.cpu()[0]) == 0
    d_info.fill_(77)
    assert la.sgetrf_async(5, 0, d_a.float(), 5, ivec(1), d_info) == 0
    assert int(d_info.cpu()[0]) == 0


def test_getrf_wait_failure_async_leaves_a(dev):
    """Every bounded wait of one factorization gives up (forced from the
    host): INFO = LA_ERR_SYNC, A exactly as passed, sticky status set; the next
    call factors normally."""
    import torch
    n = 2048
    a = la.matrix(3, 2049, n, n)
    d_a = to_dev(a)
    d_p = ivec(n)
    d_info = ivec(1)
    la.device_status(clear=True)
    la.test_fail_waits(1)
    assert la.dgetrf_async(n, n, d_a, n, d_p, d_info) == 0
    torch.cuda.synchronize()
    assert int(d_info.cpu()[0]) == la.LA_ERR_SYNC
    assert np.array_equal(from_dev(d_a, (n, n)).view(np.int64), a.view(np.int64))
    assert la.device_status(clear=True) & 1
    assert la.device_status(clear=True) == 0
    assert la.dgetrf_async(n, n, d_a, n, d_p, d_info) == 0
    torch.cuda.synchronize()
    assert int(d_info.cpu()[0]) == 0
    _, p_r, _ = oracle.dgetrf(a)
    assert np.array_equal(d_p.cpu().numpy(), p_r)


@pytest.mark.parametrize("n", [2048, 5000])
def test_getrf_wait_failure_sync_recovers(dev, n):
    """The synchronous entry point redoes the call with the panel form that
    never waits: the reference's IPIV and DGET01 < 30."""
    a = la.matrix(3, 5000 + n, n, n)
    aa = np.array(a, order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    la.device_status(clear=True)
    la.test_fail_waits(1)
    info = la.dgetrf(n, n, aa, n, ipiv)
    assert info == 0
    assert la.device_status(clear=True) & 1   # the first attempt gave up
    lu_r, p_r, _ = oracle.dgetrf(a)
    assert np.array_equal(ipiv, p_r)
    assert oracle.dget01(a, aa, ipiv) < THRESH


@pytest.mark.parametrize("trans", ["N", "T"])
def test_getrs_wait_failure(dev, trans):
    """Every wait of a triangular sweep gives up: the asynchronous solve leaves
    B as passed and sets the sticky status; the synchronous solve redoes it
    without waits and meets DGET02."""
    import torch
    n, nrhs = 1500, 3
    a = la.matrix(2, 1501, n, n)
    b = la.matrix(2, 1502, n, nrhs)
    lu, p, info = gpu_getrf(a)
    assert info == 0
    d_lu, d_p = to_dev(lu), to_dev(p.reshape(-1, 1).astype(np.int32))
    d_b = to_dev(b)
    la.device_status(clear=True)
    la.test_fail_waits(1)
    assert la.dgetrs_async(trans, n, nrhs, d_lu, n, d_p, d_b, n) == 0
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_b, (n, nrhs)).view(np.int64), b.view(np.int64))
    assert la.device_status(clear=True) & 1
    la.test_fail_waits(1)
    bb = np.array(b, order="F")
    assert la.dgetrs(trans, n, nrhs, d_lu, n, d_p, bb, n) == 0
    assert oracle.dget02(a, bb, b, trans) < THRESH
    la.device_status(clear=True)


def test_dsgesv_reports_wait_failure(dev):
    """A mixed-precision driver whose inner solve gave up returns LA_ERR_SYNC."""
    n = 1200
    a = la.matrix(2, 1201, n, n)
    b = la.matrix(2, 1202, n, 1)
    aa = np.array(a, order="F")
    x = np.zeros((n, 1), order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    la.device_status(clear=True)
    la.test_fail_waits(1)
    info, _ = la.dsgesv(n, 1, aa, n, ipiv, b, n, x, n)
    assert info == la.LA_ERR_SYNC
    la.test_fail_waits(0)
    assert la.device_status(clear=True) == 0
    aa = np.array(a, order="F")
    info, it = la.dsgesv(n, 1, aa, n, ipiv, b, n, x, n)
    assert info == 0 and it > 0
