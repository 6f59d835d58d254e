"""GPU checks of equilibration (DGEEQU / DGEEQUB / DLAQGE) and of the
mixed-precision solver with refinement and GMRES-IR.

Equilibration is pinned bitwise to the oracle (itself pinned bitwise to the
reference build and to DCHKEQ's closed form): R, C, ROWCND, COLCND, AMAX,
INFO and the scaled matrix.

The mixed-precision solver has no reference routine beyond DSGESV, whose
contract it keeps: per-column stopping test max|r| <= max|x| ||A||_inf eps
sqrt(n) (SRC/dsgesv.f:337-342), ITER codes, FP64 fallback.  Bars: the
returned backward error meets that test; DGET08 (TESTING/LIN/dget08.f)
recomputed on the CPU stays below max(30, 2 sqrt(n)); the forward error
against the FP64 oracle's solution stays within cond(A) * n * eps.
"""
import numpy as np
import pytest

import lapack_amd as la
import oracle

from gpu_util import from_dev, ivec, to_dev

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.uint64)


@pytest.mark.parametrize("radix", [False, True])
@pytest.mark.parametrize("shape", [(30, 25), (257, 130), (1000, 1000)])
def test_geequ_bitwise(dev, radix, shape):
    m, n = shape
    a = la.matrix(3, 21 + m, m, n) * np.logspace(-30, 30, m)[:, None] * np.logspace(5, -5, n)[None, :]
    a = np.asfortranarray(a)
    r = np.zeros(m)
    c = np.zeros(n)
    info, rowcnd, colcnd, amax = la.dgeequ(m, n, a, m, r, c, radix)
    r1, c1, rc1, cc1, am1, info1 = oracle.dgeequ(a, radix)
    assert info == info1 == 0
    assert np.array_equal(bits(r), bits(r1)) and np.array_equal(bits(c), bits(c1))
    assert (rowcnd, colcnd, amax) == (rc1, cc1, am1)
    if radix:
        assert np.all(np.frexp(r)[0] == 0.5) and np.all(np.frexp(c)[0] == 0.5)
    # DLAQGE on the device copy, same decision and bytes as the oracle
    d_a = to_dev(a)
    equed = la.dlaqge(m, n, d_a, m, to_dev(r), to_dev(c), rowcnd, colcnd, amax)
    x1, e1 = oracle.dlaqge(a, r1, c1, rc1, cc1, am1)
    assert equed == e1
    assert np.array_equal(bits(from_dev(d_a, (m, n))), bits(x1))


def test_geequ_zero_row_and_column(dev):
    a = la.matrix(2, 5, 40, 30)
    a[7, :] = 0.0
    r, c = np.zeros(40), np.zeros(30)
    info = la.dgeequ(40, 30, a, 40, r, c)[0]
    assert info == 8 == oracle.dgeequ(a)[5]
    a = la.matrix(2, 6, 40, 30)
    a[:, 11] = 0.0
    info = la.dgeequ(40, 30, a, 40, r, c, True)[0]
    assert info == 40 + 12 == oracle.dgeequ(a, True)[5]


def conditioned(n, cond, seed, mode=3):
    """A = Q1 diag(s) Q2^T, Q from Householder QR of seeded Gaussians; the
    singular values follow DLATMS's modes (TESTING/MATGEN/dlatms.f:158-161):
    3 geometric 1 .. 1/cond, 4 arithmetic 1 .. 1/cond."""
    q1, _ = np.linalg.qr(la.matrix(3, seed, n, n))
    q2, _ = np.linalg.qr(la.matrix(3, seed + 1, n, n))
    if mode == 3:
        s = np.logspace(0, -np.log10(cond), n)
    else:
        s = 1.0 - np.arange(n) / (n - 1.0) * (1.0 - 1.0 / cond)
    return np.asfortranarray((q1 * s) @ q2.T)


def solve_ir(a, b, prec):
    n, nrhs = b.shape
    d_a = to_dev(a)
    d_b = to_dev(b)
    d_x = to_dev(np.zeros((n, nrhs)))
    d_p = ivec(n)
    berr = np.zeros(nrhs)
    info, it, git = la.dgesv_ir(n, nrhs, d_a, n, d_p, d_b, n, d_x, n, prec, berr)
    return from_dev(d_x, (n, nrhs)), it, git, info, berr


def check_solution(a, b, x, berr, it):
    n = a.shape[0]
    assert it >= 0, it
    assert np.all(berr <= EPS * np.sqrt(n) * (1 + 1e-12)), berr
    assert oracle.dget08(a, x, b) < np.sqrt(n)  # DDRVAB's bar for ITER >= 0 (ddrvab.f:362-365)
    lu, p, _ = oracle.dgetrf(a)
    xr, _ = oracle.dgetrs(lu, p, b)
    kappa = np.linalg.cond(a, np.inf)
    fwd = np.max(np.abs(x - xr)) / np.max(np.abs(xr))
    assert fwd <= max(1e-14, kappa * n * EPS), (fwd, kappa)
    return fwd


@pytest.mark.parametrize("prec", ["S", "H", "B"])
def test_ir_well_conditioned(dev, prec):
    n, nrhs = 1024, 3
    a = la.matrix(3, 700, n, n)
    b = la.matrix(3, 701, n, nrhs)
    x, it, git, info, berr = solve_ir(a, b, prec)
    assert info == 0
    fwd = check_solution(a, b, x, berr, it)
    print(f"prec {prec}: iter {it}, gmres {git}, berr {berr.max():.2e}, fwd {fwd:.2e}")


@pytest.mark.parametrize("env", [{"LAPACK_AMD_UPD_BIG_MIN": "1"},
                                 {"LAPACK_AMD_UPD_BIG_MIN": "1", "LAPACK_AMD_UPD_DMA": "0"}],
                         ids=["direct_lds", "staged"])
@pytest.mark.parametrize("prec", ["S", "H", "B"])
def test_ir_big_update_kernels(dev, monkeypatch, env, prec):
    """Every trailing update of the FP32-stored factorization that fits on the 256 x 128 kernels (tile
    threshold 1): the direct global-to-LDS kernels (FP32, and FP16 / BF16 operands), or with
    LAPACK_AMD_UPD_DMA=0 the register-staged ones; edge tiles in both dimensions at n = 1000."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    n, nrhs = 1000, 2
    a = la.matrix(3, 720, n, n)
    b = la.matrix(3, 721, n, nrhs)
    x, it, git, info, berr = solve_ir(a, b, prec)
    assert info == 0
    check_solution(a, b, x, berr, it)


@pytest.mark.parametrize("prec", ["S", "H", "B"])
def test_ir_conditioned_1e6(dev, prec):
    """The conditioned variant of the mixed-precision workload (kappa = 1e6) at
    n = 1024, arithmetic spectrum: FP16 and BF16 factors cannot refine
    classically (kappa * u > 1), GMRES-IR must carry them."""
    n = 1024
    a = conditioned(n, 1e6, 710, mode=4)
    b = la.matrix(3, 712, n, 1)
    x, it, git, info, berr = solve_ir(a, b, prec)
    assert info == 0
    fwd = check_solution(a, b, x, berr, it)
    if prec in ("H", "B"):
        assert git > 0
    print(f"prec {prec}: iter {it}, gmres {git}, berr {berr.max():.2e}, fwd {fwd:.2e}")


@pytest.mark.parametrize("prec", ["S", "H", "B"])
def test_ir_conditioned_geometric(dev, prec):
    """kappa = 1e6 with a geometric spectrum (the workload's conditioned
    variant, here at n = 1024): about half the singular values lie below the
    FP16/BF16 unit roundoff, so classic refinement stalls and GMRES-IR must
    converge -- ITER >= 0, no FP64 fallback -- with DSGESV's stopping test,
    DDRVAB's bar and the forward error against the oracle."""
    n = 1024
    a = conditioned(n, 1e6, 710, mode=3)
    b = la.matrix(3, 712, n, 1)
    x, it, git, info, berr = solve_ir(a, b, prec)
    assert info == 0
    assert it >= 0, (prec, it, git)
    fwd = check_solution(a, b, x, berr, it)
    if prec in ("H", "B"):
        assert git > 0
    print(f"prec {prec}: iter {it}, gmres {git}, berr {berr.max():.2e}, fwd {fwd:.2e}")


def test_ir_fallbacks_and_errors(dev):
    n = 200
    b = la.matrix(2, 1, n, 1)
    a = la.matrix(2, 3, n, n)
    a[3, 4] = 1e39                     # beyond FP32: ITER -2, FP64 fallback
    x, it, git, info, berr = solve_ir(a, b, "S")
    assert (it, info) == (-2, 0) and oracle.dget02(a, x, b) < 30
    a = la.matrix(2, 4, n, n)
    a[:, 7] = 0.0                      # singular: ITER -3, FP64 INFO
    for prec in "SHB":
        _, it, _, info, _ = solve_ir(a, b, prec)
        assert (it, info) == (-3, 8), prec
    z = np.zeros((4, 4), order="F")
    zb = np.zeros((4, 1), order="F")
    p = np.zeros(4, dtype=np.int32)
    assert la.dgesv_ir(4, 1, z, 4, p, zb, 4, zb, 4, "Q")[0] == -10
    assert la.dgesv_ir(-1, 1, z, 4, p, zb, 4, zb, 4, "S")[0] == -1
    assert la.dgesv_ir(4, 1, z, 4, p, zb, 4, zb, 3, "S")[0] == -9


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (130, 3), (700, 2)])
def test_dgerfs_against_oracle(dev, trans, n, nrhs):
    """DGERFS on the GPU (residuals, |op(A)||x| and the solves there, the
    reference's decisions and DLACN2 on the host) against the oracle's
    DGERFS from the same starting point.  The GPU sums in another order, so
    BERR (a ratio of rounding-level residuals) agrees in magnitude, not bits:
    both stay below (n+1) eps (DGET07's scale); FERR must bound the true
    error of a known solution and agree with the oracle's within 10x."""
    a = la.matrix(2, 900 + n, n, n)
    xact = la.matrix(2, 901 + n, n, nrhs)
    b = np.asfortranarray(a.T @ xact if trans == "T" else a @ xact)
    lu, p, _ = oracle.dgetrf(a)
    x0, _ = oracle.dgetrs(lu, p, b, trans)
    x0 = np.asfortranarray(x0.astype(np.float32).astype(np.float64))
    xr, ferr_r, berr_r, info_r = oracle.dgerfs(a, lu, p, b, x0, trans)
    d_x = to_dev(x0)
    ferr = np.zeros(nrhs)
    berr = np.zeros(nrhs)
    info = la.dgerfs(trans, n, nrhs, to_dev(a), n, to_dev(lu), n, to_dev(p.astype(np.int32)).view(-1), to_dev(b), n,
                     d_x, n, ferr, berr)
    assert info == info_r == 0
    x = from_dev(d_x, (n, nrhs))
    eps = 2.0 ** -53
    kappa = np.linalg.cond(a, np.inf)
    assert np.max(np.abs(x - xr)) <= 10 * kappa * eps * np.max(np.abs(xr))
    assert np.all(berr <= (n + 1) * eps) and np.all(berr_r <= (n + 1) * eps)
    err = np.max(np.abs(x - xact), axis=0) / np.max(np.abs(x), axis=0)
    assert np.all(err <= ferr + 30 * kappa * eps), (err, ferr)
    assert np.all(ferr <= 10 * ferr_r) and np.all(ferr_r <= 10 * ferr), (ferr, ferr_r)
