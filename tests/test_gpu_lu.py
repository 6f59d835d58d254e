"""GPU parity of the plain FP64/FP32 LU path against the oracle.

Bars (the reference's own test ratios):
  * IPIV identical to the reference DGETRF (integer output).  A divergence is
    accepted only at a near-tie, traced both ways at the first differing step
    k: in the reference's factors the row the GPU chose has a multiplier
    |L| >= 1 - tol_k, and (when the GPU's factors are given) in the GPU's
    factors the row the reference chose has |L| >= 1 - tol_k, where tol_k is
    the rounding two LU implementations may each leave in a candidate of the
    Schur complement at step k, relative to the pivot:
        tol_k = 10 * gamma_k * max_{i>=k} (|L||U|)_{ik} / |U_kk|,
        gamma_k = (k+1) u / (1 - (k+1) u)
    (the standard componentwise bound |dA^(k)| <= gamma_k |L||U|), with the
    unit roundoff u of the factoring precision: 2^-53 for FP64 factors,
    2^-24 for FP32 -- one bar for both.
  * DGET01 ||P*L*U - A||_1/(n ||A||_1 eps) < 30 (TESTING/LIN/dget01.f, threshold
    TESTING/dtest.in:13), DGET02 ||B - A X||_1/(||A||_1 ||X||_1 eps) < 30.
  * GEMM: DBLAT3's componentwise ratio |ct - cc|/(eps*g) < 16
    (BLAS/TESTING/dblat3.f:2940-2960).
  * INFO equal (exact zero pivots).
"""
import numpy as np
import torch
import pytest

import lapack_amd as la
import oracle

from gpu_util import from_dev, ivec, to_dev

pytestmark = pytest.mark.gpu

THRESH = 30.0


def first_mismatch(p, q):
    d = np.nonzero(np.asarray(p) != np.asarray(q))[0]
    return int(d[0]) if len(d) else -1


def near_tie(lu_ref, k):
    """1 - second largest |L(i,k)|, i > k, of the reference factors: the gap
    between the chosen pivot and the runner-up at step k."""
    col = np.abs(lu_ref[k + 1:, k])
    if len(col) == 0:
        return 1.0
    return 1.0 - col.max()


def final_row(ipiv, k, r):
    """Row of the final factors holding the content that sat in row r when
    step k began (0-based), after step k's and every later interchange."""
    p = int(ipiv[k]) - 1
    pos = p if r == k else (k if r == p else r)
    for j in range(k + 1, len(ipiv)):
        q = int(ipiv[j]) - 1
        if pos == j:
            pos = q
        elif pos == q:
            pos = j
    return pos


def tie_tolerance(lu, k, unit):
    """10 * gamma_k * max_{i>=k} (|L||U|)_{ik} / |U_kk| from factors lu."""
    lu = np.asarray(lu, dtype=np.float64)
    ukk = abs(float(lu[k, k]))
    if ukk == 0.0:
        return np.inf
    lcol = np.concatenate(([1.0], np.abs(lu[k + 1:, k])))  # column k of L, rows k.. (unit diagonal)
    g = np.abs(lu[k:, :k]) @ np.abs(lu[:k, k]) + lcol * ukk
    gam = (k + 1) * unit / (1.0 - (k + 1) * unit)
    return 10.0 * gam * float(g.max()) / ukk


def assert_ipiv(ipiv_gpu, ipiv_ref, lu_ref, lu_gpu=None, unit=2.0 ** -53):
    """IPIV equal, or a near-tie at the first differing step traced in the
    reference's factors and, with lu_gpu, in the GPU's (see the module
    docstring).  unit: unit roundoff of the factoring precision."""
    k = first_mismatch(ipiv_gpu, ipiv_ref)
    if k < 0:
        return
    tol = tie_tolerance(np.asarray(lu_ref, dtype=np.float64), k, unit)
    r_gpu, r_ref = int(ipiv_gpu[k]) - 1, int(ipiv_ref[k]) - 1
    gap_ref = 1.0 - abs(float(lu_ref[final_row(ipiv_ref, k, r_gpu), k]))
    msg = f"IPIV differs at step {k}: rows {r_gpu + 1} (GPU) / {r_ref + 1} (reference)"
    assert gap_ref <= tol, f"{msg}; in the reference's factors the GPU's row has gap {gap_ref:.3e} > {tol:.3e}"
    if lu_gpu is not None:
        gap_gpu = 1.0 - abs(float(lu_gpu[final_row(ipiv_gpu, k, r_ref), k]))
        assert gap_gpu <= tol, f"{msg}; in the GPU's factors the reference's row has gap {gap_gpu:.3e} > {tol:.3e}"
    print(f"near-tie at step {k}: gaps {gap_ref:.3e} (reference factors)"
          + (f", {gap_gpu:.3e} (GPU factors)" if lu_gpu is not None else "") + f" <= {tol:.3e}")


def gpu_getrf(a, dtype=np.float64):
    m, n = a.shape
    d_a = to_dev(np.asarray(a, dtype=dtype))
    d_p = ivec(min(m, n))
    fn = la.dgetrf if dtype == np.float64 else la.sgetrf
    info = fn(m, n, d_a, max(1, m), d_p)
    return from_dev(d_a, (m, n)), d_p.cpu().numpy()[:min(m, n)], info


def dmmch_ratio(ta, tb, alpha, a, b, beta, c, cc, eps):
    opa = a.T if ta == "T" else a
    opb = b.T if tb == "T" else b
    ct = oracle.dgemm(ta, tb, alpha, a, b, beta, c)
    g = abs(alpha) * (np.abs(opa) @ np.abs(opb)) + abs(beta) * np.abs(c)
    err = np.abs(ct - cc) / eps
    nz = g != 0
    err[nz] /= g[nz]
    return err.max() if err.size else 0.0


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (128, 128, 16), (129, 130, 131), (300, 70, 257)])
def test_dgemm_dblat3_ratio(dev, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k) if ta == "N" else (k, m)))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n) if tb == "N" else (n, k)))
    c = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    for alpha, beta in [(1.0, 0.0), (-1.0, 1.0), (0.7, -1.3)]:
        d_a, d_b, d_c = to_dev(a), to_dev(b), to_dev(c)
        rc = la.dgemm(ta, tb, m, n, k, alpha, d_a, a.shape[0], d_b, b.shape[0], beta, d_c, m)
        assert rc == 0
        cc = from_dev(d_c, (m, n))
        assert dmmch_ratio(ta, tb, alpha, a, b, beta, c, cc, 2.0 ** -52) < 16.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["d", "s"])
@pytest.mark.parametrize("m,n,k,pad,big_min", [
    (1000, 1300, 64, 4, "1"),   # direct-to-LDS kernel: padded leading dimensions, edge tiles in M and N
    (1001, 1300, 64, 0, "1"),   # M not a multiple of the 16-byte vector: the 256 x 128 register-staged kernel
    (1000, 1300, 72, 0, "1"),   # K not a multiple of 16: the same
    (4100, 4200, 64, 0, None),  # the default tile threshold, direct-to-LDS kernel
])
def test_gemm_update_form(dev, monkeypatch, dtype, m, n, k, pad, big_min):
    """C - A*B (TRANSA = TRANSB = 'N', ALPHA = -1, BETA = 1: the trailing update of SRC/dgetrf.f:212-216)
    on the update kernels, in the DBLAT3 ratio (BLAS/TESTING/dblat3.f:2940-2960) against the oracle's
    DGEMM of the same (rounded) inputs, eps = 2^-52 or 2^-23."""
    if big_min is not None:
        monkeypatch.setenv("LAPACK_AMD_UPD_BIG_MIN", big_min)
    rng = np.random.default_rng(m + 7 * n + 31 * k + pad)
    a = rng.uniform(-1, 1, (m, k)).astype(dtype)
    b = rng.uniform(-1, 1, (k, n)).astype(dtype)
    c = rng.uniform(-1, 1, (m, n)).astype(dtype)
    lda, ldb, ldc = m + pad, k + pad, m + pad
    ap = np.zeros((lda, k), dtype, order="F"); ap[:m] = a
    bp = np.zeros((ldb, n), dtype, order="F"); bp[:k] = b
    cp = np.zeros((ldc, n), dtype, order="F"); cp[:m] = c
    d_a, d_b, d_c = to_dev(ap), to_dev(bp), to_dev(cp)
    if dtype == np.float64:
        assert la.dgemm("N", "N", m, n, k, -1.0, d_a, lda, d_b, ldb, 1.0, d_c, ldc) == 0
    else:
        assert la.sgemm_async("N", "N", m, n, k, -1.0, d_a, lda, d_b, ldb, 1.0, d_c, ldc) == 0
        torch.cuda.synchronize()
    cc = from_dev(d_c, (ldc, n))
    assert np.array_equal(cc[m:], cp[m:]), "rows past M were written"
    eps = 2.0 ** -52 if dtype == np.float64 else 2.0 ** -23
    r = dmmch_ratio("N", "N", -1.0, a.astype(np.float64), b.astype(np.float64), 1.0, c.astype(np.float64),
                    cc[:m].astype(np.float64), eps)
    assert r < 16.0, r


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["d", "s"])
@pytest.mark.parametrize("which", ["a", "b", "c"])
def test_gemm_update_form_unaligned(dev, monkeypatch, dtype, which):
    """C - A*B with one operand starting one element past a 16-byte boundary: the register-staged kernel
    takes it (the direct-to-LDS kernel needs 16-byte aligned operands); DBLAT3 ratio as above."""
    monkeypatch.setenv("LAPACK_AMD_UPD_BIG_MIN", "1")
    m, n, k = 1000, 1300, 64
    rng = np.random.default_rng(77 + ord(which))
    a = rng.uniform(-1, 1, (m, k)).astype(dtype)
    b = rng.uniform(-1, 1, (k, n)).astype(dtype)
    c = rng.uniform(-1, 1, (m, n)).astype(dtype)

    def dev_off(x, off):
        t = torch.zeros(x.size + off, dtype=torch.float64 if dtype == np.float64 else torch.float32, device="cuda")
        t[off:] = torch.from_numpy(np.asfortranarray(x).ravel(order="F")).cuda()
        return t, t[off:]

    offs = {"a": 0, "b": 0, "c": 0}
    offs[which] = 1
    ta, d_a = dev_off(a, offs["a"])
    tb, d_b = dev_off(b, offs["b"])
    tc, d_c = dev_off(c, offs["c"])
    if dtype == np.float64:
        assert la.dgemm("N", "N", m, n, k, -1.0, d_a, m, d_b, k, 1.0, d_c, m) == 0
    else:
        assert la.sgemm_async("N", "N", m, n, k, -1.0, d_a, m, d_b, k, 1.0, d_c, m) == 0
        torch.cuda.synchronize()
    cc = d_c.cpu().numpy().reshape(n, m).T
    eps = 2.0 ** -52 if dtype == np.float64 else 2.0 ** -23
    r = dmmch_ratio("N", "N", -1.0, a.astype(np.float64), b.astype(np.float64), 1.0, c.astype(np.float64),
                    cc.astype(np.float64), eps)
    assert r < 16.0, r
    assert float(tc[0]) == 0.0 or offs["c"] == 0  # the element before C untouched


@pytest.mark.parametrize("occ", ["1", "2"])
@pytest.mark.parametrize("n", [1024, 2048])
def test_sgetrf_parity_big_update_kernels(dev, monkeypatch, n, occ):
    """SGETRF with every trailing update that fits on the direct-to-LDS kernel (tile threshold 1), one
    or two workgroups per CU; IPIV against the reference SGETRF as test_sgetrf_parity."""
    monkeypatch.setenv("LAPACK_AMD_UPD_BIG_MIN", "1")
    monkeypatch.setenv("LAPACK_AMD_UPD_DMA_OCC", occ)
    test_sgetrf_parity(dev, n)


def test_dgemm_beta_zero_does_not_read_c(dev):
    m, n, k = 70, 40, 30
    rng = np.random.default_rng(1)
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    c = np.full((m, n), np.nan, order="F")
    d_c = to_dev(c)
    assert la.dgemm("N", "N", m, n, k, 1.0, to_dev(a), m, to_dev(b), k, 0.0, d_c, m) == 0
    cc = from_dev(d_c, (m, n))
    assert np.isfinite(cc).all()
    np.testing.assert_allclose(cc, a @ b, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 10, 50, 63, 64, 65, 127, 128, 129, 255, 256, 257, 300, 513, 1024])
def test_dgetrf_square_parity(dev, n):
    a = la.matrix(2, 1988 + n, n, n)
    lu_r, p_r, info_r = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info == info_r == 0
    assert_ipiv(p, p_r, lu_r, lu)
    assert oracle.dget01(a, lu, p) < THRESH


@pytest.mark.parametrize("m,n", [(1, 5), (5, 1), (70, 30), (30, 70), (300, 257), (257, 300), (1000, 64)])
def test_dgetrf_rect_parity(dev, m, n):
    a = la.matrix(3, 1991 + m + n, m, n)
    lu_r, p_r, info_r = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info == info_r
    assert_ipiv(p, p_r, lu_r, lu)
    assert oracle.dget01(a, lu, p) < THRESH


@pytest.mark.parametrize("nb,leaf", [(1, 1), (3, 3), (20, 8), (64, 64), (128, 32), (256, 64)])
def test_dgetrf_block_size_sweep(dev, nb, leaf):
    """NB/leaf sweep through the XLAENV-style hook (TESTING/dtest.in:8-9 sweeps NB)."""
    n = 200
    a = la.matrix(2, 4242, n, n)
    lu_r, p_r, _ = oracle.dgetrf(a)
    la.set_tuning(nb, leaf)
    try:
        lu, p, info = gpu_getrf(a)
    finally:
        la.set_tuning(0, 0)
    assert info == 0
    assert_ipiv(p, p_r, lu_r, lu)
    assert oracle.dget01(a, lu, p) < THRESH


@pytest.mark.parametrize("kind", ["first", "last", "half"])
def test_dgetrf_zero_columns_info(dev, kind):
    """DCHKGE types 5-7 (TESTING/LIN/dchkge.f:325-347): INFO is the first zero column."""
    n = 90
    a = la.matrix(2, 77, n, n)
    if kind == "first":
        a[:, 0] = 0.0
        izero = 1
    elif kind == "last":
        a[:, n - 1] = 0.0
        izero = n
    else:
        a[:, n // 2:] = 0.0
        izero = n // 2 + 1
    lu_r, p_r, info_r = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info_r == izero
    assert info == izero
    assert_ipiv(p, p_r, lu_r, lu)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (255, 1), (257, 15), (1024, 16)])
def test_dgetrs_parity(dev, trans, n, nrhs):
    a = la.matrix(2, 1000 + n, n, n)
    b = la.matrix(2, 2000 + n, n, nrhs)
    lu_r, p_r, _ = oracle.dgetrf(a)
    lu, p, info = gpu_getrf(a)
    assert info == 0
    d_lu, d_p, d_b = to_dev(lu), to_dev(p.reshape(-1, 1).astype(np.int32)), to_dev(b)
    assert la.dgetrs(trans, n, nrhs, d_lu, n, d_p, d_b, n) == 0
    x = from_dev(d_b, (n, nrhs))
    assert oracle.dget02(a, x, b, trans) < THRESH
    x_r, _ = oracle.dgetrs(lu_r, p_r, b, trans)
    # same LU up to rounding -> same solution up to cond(A)*eps
    np.testing.assert_allclose(x, x_r, rtol=1e-9 * n, atol=1e-9 * n * np.abs(x_r).max())


def test_dgesv_host_pointers(dev):
    n, nrhs = 300, 4
    a = la.matrix(2, 5, n, n)
    b = la.matrix(2, 6, n, nrhs)
    aa = np.array(a, order="F")
    bb = np.array(b, order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    info = la.dgesv(n, nrhs, aa, n, ipiv, bb, n)
    assert info == 0
    lu_r, p_r, _ = oracle.dgetrf(a)
    assert_ipiv(ipiv, p_r, lu_r)
    assert oracle.dget01(a, aa, ipiv) < THRESH
    assert oracle.dget02(a, bb, b) < THRESH


def test_dgetrf_host_with_lda(dev):
    m, n, lda = 100, 80, 123
    a = la.matrix(2, 8, m, n)
    buf = np.full((lda, n), 7.0, order="F")
    buf[:m, :] = a
    ipiv = np.zeros(min(m, n), dtype=np.int32)
    info = la.dgetrf(m, n, buf, lda, ipiv)
    assert info == 0
    assert np.all(buf[m:, :] == 7.0)  # rows beyond M untouched
    lu_r, p_r, _ = oracle.dgetrf(a)
    assert_ipiv(ipiv, p_r, lu_r)
    assert oracle.dget01(a, np.asfortranarray(buf[:m, :]), ipiv) < THRESH


@pytest.mark.parametrize("incx", [1, -1, 2, -2])
def test_dlaswp_bitwise(dev, incx):
    m, n = 60, 9
    a = la.matrix(2, 9, m, n)
    k1, k2 = 3, 40
    rng = np.random.default_rng(incx + 10)
    ipiv = np.zeros(k1 + (k2 - k1) * abs(incx) + 1, dtype=np.int32)
    ipiv[:] = rng.integers(1, m + 1, size=len(ipiv))
    ref = oracle.dlaswp(a, k1, k2, ipiv, incx)
    if oracle.ref_available():
        assert np.array_equal(ref, oracle.ref_dlaswp(a, k1, k2, ipiv, incx))
    aa = np.array(a, order="F")
    la.dlaswp(n, aa, m, k1, k2, ipiv, incx)
    assert np.array_equal(aa, ref)


@pytest.mark.parametrize("m,n,form,incx", [(5003, 700, "getrf", 1), (5003, 700, "any", 1), (5003, 700, "any", -1),
                                           (300, 1100, "getrf", -1), (17000, 3, "any", 1)])
def test_dlaswp_large_bitwise(dev, m, n, form, incx):
    """DLASWP of a full sequence of M interchanges: the row count not a multiple
    of 256, more columns than the in-place kernel's workgroups (each handles
    several), targets at or after their step (DGETRF's form) or anywhere, and
    more rows than one column in LDS holds (17000 FP64: gather and copy)."""
    a = la.matrix(2, 77 + m + n, m, n)
    rng = np.random.default_rng(m + n + incx)
    k1, k2 = 1, m
    if form == "getrf":
        ipiv = np.array([rng.integers(k, m + 1) for k in range(1, m + 1)], dtype=np.int32)
    else:
        ipiv = rng.integers(1, m + 1, size=m).astype(np.int32)
    ref = oracle.dlaswp(a, k1, k2, ipiv, incx)
    aa = np.array(a, order="F")
    assert la.dlaswp(n, aa, m, k1, k2, ipiv, incx) == 0
    assert np.array_equal(aa, ref)


@pytest.mark.parametrize("n", [64, 300, 1024])
def test_sgetrf_parity(dev, n):
    a = la.matrix(2, 3000 + n, n, n).astype(np.float32)
    lu_r, p_r, _ = oracle.sgetrf(a)
    lu, p, info = gpu_getrf(a, np.float32)
    assert info == 0
    assert_ipiv(p, p_r, lu_r.astype(np.float64), lu.astype(np.float64), unit=2.0 ** -24)
    # SGET01-style ratio in single precision
    l = np.tril(lu.astype(np.float64), -1) + np.eye(n)
    u = np.triu(lu.astype(np.float64))
    pa = oracle.dlaswp(a.astype(np.float64), 1, n, p, 1)
    r = np.abs(l @ u - pa).sum(0).max() / (n * np.abs(a.astype(np.float64)).sum(0).max() * 2.0 ** -24)
    assert r < THRESH


def test_determinism(dev):
    """Two runs give the same bits (LDS / cross-workgroup races would not)."""
    n = 700
    a = la.matrix(3, 31, n, n)
    lu1, p1, _ = gpu_getrf(a)
    lu2, p2, _ = gpu_getrf(a)
    assert np.array_equal(p1, p2)
    assert np.array_equal(lu1.view(np.int64), lu2.view(np.int64))


def test_golden_ipiv_n1024(dev):
    """GPU DGETRF on the golden input reproduces the reference's IPIV and
    solves to the reference's x within the DGET02 bar."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "dgetrf_n1024_u.npz"))
    n = 1024
    a = la.matrix(2, 1988, n, n)
    b = la.matrix(2, 1989, n, 1)
    lu, ipiv, info = gpu_getrf(a)
    assert info == 0 and np.array_equal(ipiv, g["ipiv"])
    d_a, d_b = to_dev(lu), to_dev(b)
    assert la.dgetrs("N", n, 1, d_a, n, to_dev(ipiv.astype(np.int32)).view(-1), d_b, n) == 0
    x = from_dev(d_b, (n, 1))
    assert oracle.dget02(a, x, b) < THRESH
    assert np.max(np.abs(x[:, 0] - g["x"])) <= 1e-10 * np.max(np.abs(g["x"]))


def test_golden_ipiv_n16384(dev):
    """The benchmark's matrix (n = 16384, N(0,1), seed 16384): GPU IPIV equal to
    the reference DGETRF's (tests/golden/dgetrf_n16384_normal_ipiv.npy)."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "dgetrf_n16384_normal_ipiv.npy")
    if not os.path.exists(path):
        pytest.skip("golden n=16384 IPIV not generated")
    ref = np.load(path)
    n = 16384
    a = la.matrix(3, 16384, n, n)
    d_a = to_dev(a)
    del a
    d_p = ivec(n)
    assert la.dgetrf(n, n, d_a, n, d_p) == 0
    assert np.array_equal(d_p.cpu().numpy(), ref)
