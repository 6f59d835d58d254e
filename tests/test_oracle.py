"""CPU tests pinning the oracle (oracle/lu_oracle.cpp) before it is trusted.

Three kinds of evidence pin the oracle:
  * bitwise equality with the reference Fortran itself, built with flang
    from its own sources into oracle/_ref (DGETRF/DGETRF2 path, DGETRS N/T,
    DLASWP, DGEMM, DTRSM, IDAMAX, DSGESV, DGEEQU/DGEEQUB, DLAQGE);
  * closed forms and known answers the reference tests use (DCHKEQ's
    equilibration matrix, TESTING/LIN/dchkeq.f:107-173; zero-column INFO of
    DCHKGE types 5-7, dchkge.f:325-347; DSGESV ITER codes, dsgesv.f:153-165);
  * independent brute force on tiny inputs (textbook partial pivoting in
    pure Python) and mathematical invariants (P*L*U = A to rounding).
A deliberately planted mistake (swapped operand, dropped term, wrong pivot
rule) fails at least one of them: see test_planted_mistakes_are_caught.
"""
import ctypes

import numpy as np
import pytest

import lapack_amd as la
import oracle

from rng_np import dlarnv as np_dlarnv

ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference sources)")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64 if a.dtype == np.float64 else np.uint32)


# ---------------------------------------------------------------- generator

@pytest.mark.parametrize("idist", [1, 2, 3])
def test_generator_c_matches_numpy_restatement(idist):
    for seed, off, n in [(1988, 0, 5000), (7, 123456789, 777), (2 ** 40 + 3, 2 ** 33, 333)]:
        c = la.dlarnv(idist, seed, n, offset=off)
        p = np_dlarnv(idist, seed, n, offset=off)
        assert np.array_equal(bits(c), bits(p))


def test_generator_moments():
    x = la.dlarnv(3, 11, 400000)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    u = la.dlarnv(2, 12, 400000)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01
    # blocks of a matrix are independent of how they are produced
    m = la.matrix(3, 5, 40, 30)
    assert np.array_equal(bits(np.asfortranarray(m)).ravel(order="F"), bits(la.dlarnv(3, 5, 1200)))


# ------------------------------------------------- bitwise vs reference build

@ref
@pytest.mark.parametrize("m,n", [(1, 1), (1, 7), (7, 1), (5, 5), (63, 63), (64, 64), (65, 65), (130, 70),
                                 (70, 130), (200, 200), (257, 257)])
def test_dgetrf_bitwise_vs_reference(m, n):
    a = la.matrix(2, 1988 + m * 7 + n, m, n)
    lu1, p1, i1 = oracle.dgetrf(a)
    lu2, p2, i2 = oracle.ref_dgetrf(a)
    assert i1 == i2
    assert np.array_equal(p1, p2)
    assert np.array_equal(bits(lu1), bits(lu2))


@ref
def test_dgetrf2_bitwise_and_zero_columns():
    for kind in range(3):
        a = la.matrix(2, 99 + kind, 50, 50)
        if kind == 0:
            a[:, 0] = 0
        elif kind == 1:
            a[:, 49] = 0
        else:
            a[:, 25:] = 0
        lu1, p1, i1 = oracle.dgetrf(a)
        lu2, p2, i2 = oracle.ref_dgetrf(a)
        assert i1 == i2 == [1, 50, 26][kind]
        assert np.array_equal(p1, p2) and np.array_equal(bits(lu1), bits(lu2))


@ref
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgetrs_bitwise_vs_reference(trans):
    n = 150
    a = la.matrix(3, 5, n, n)
    b = la.matrix(3, 6, n, 7)
    lu, p, _ = oracle.dgetrf(a)
    x1, _ = oracle.dgetrs(lu, p, b, trans)
    x2, _ = oracle.ref_dgetrs(lu, p, b, trans)
    assert np.array_equal(bits(x1), bits(x2))


@ref
@pytest.mark.parametrize("incx", [1, -1, 2, -3])
def test_dlaswp_bitwise_vs_reference(incx):
    m, n = 40, 37
    a = la.matrix(2, 3, m, n)
    rng = np.random.default_rng(abs(incx))
    k1, k2 = 2, 30
    ipiv = rng.integers(1, m + 1, size=k1 + (k2 - k1) * abs(incx) + 2).astype(np.int32)
    assert np.array_equal(oracle.dlaswp(a, k1, k2, ipiv, incx), oracle.ref_dlaswp(a, k1, k2, ipiv, incx))


@ref
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_dgemm_bitwise_vs_reference(ta, tb):
    m, n, k = 23, 17, 11
    a = la.matrix(2, 1, m, k) if ta == "N" else la.matrix(2, 1, k, m)
    b = la.matrix(2, 2, k, n) if tb == "N" else la.matrix(2, 2, n, k)
    c = la.matrix(2, 3, m, n)
    for alpha, beta in [(1.0, 0.0), (-1.0, 1.0), (0.5, -2.0), (0.0, 3.0)]:
        assert np.array_equal(bits(oracle.dgemm(ta, tb, alpha, a, b, beta, c)),
                              bits(oracle.ref_dgemm(ta, tb, alpha, a, b, beta, c)))


@ref
@pytest.mark.parametrize("uplo,trans,diag", [(u, t, d) for u in "UL" for t in "NT" for d in "UN"])
def test_dtrsm_left_bitwise_vs_reference(uplo, trans, diag):
    m, n = 19, 6
    a = la.matrix(2, 4, m, m) + 4 * np.eye(m)
    b = la.matrix(2, 5, m, n)
    assert np.array_equal(bits(oracle.dtrsm_left(uplo, trans, diag, 0.75, a, b)),
                          bits(oracle.ref_dtrsm("L", uplo, trans, diag, 0.75, a, b)))


@ref
def test_idamax_rule_vs_reference():
    cases = [
        [1.0, -3.0, 3.0, 2.0],          # tie -> first
        [np.nan, 5.0, 1.0],             # NaN first -> 1
        [1.0, np.nan, 0.5, np.nan],     # NaN elsewhere never wins
        [0.0, -0.0, 0.0],
        [np.inf, -np.inf, 1.0],
        [-2.0],
    ]
    for c in cases:
        x = np.array(c)
        assert oracle.idamax(x) == oracle.ref_idamax(x)
    assert oracle.idamax(np.array([1.0, -3.0, 3.0, 2.0])) == 2
    assert oracle.idamax(np.array([np.nan, 5.0])) == 1
    assert oracle.idamax(np.array([1.0, np.nan, 0.5])) == 1


@ref
@pytest.mark.parametrize("n,nrhs", [(1, 1), (13, 2), (45, 1), (132, 15)])
def test_dsgesv_bitwise_vs_reference(n, nrhs):
    a = la.matrix(2, 2006 + n, n, n)
    b = la.matrix(2, 2007 + n, n, nrhs)
    x1, p1, it1, inf1, a1 = oracle.dsgesv(a, b)
    x2, p2, it2, inf2, a2 = oracle.ref_dsgesv(a, b)
    assert (it1, inf1) == (it2, inf2)
    assert it1 >= 0
    assert np.array_equal(p1, p2) and np.array_equal(bits(x1), bits(x2)) and np.array_equal(bits(a1), bits(a2))


@ref
@pytest.mark.parametrize("radix", [False, True])
def test_dgeequ_dlaqge_bitwise_vs_reference(radix):
    a = la.matrix(3, 21, 30, 25) * np.logspace(-30, 30, 30)[:, None] * np.logspace(5, -5, 25)[None, :]
    r1 = oracle.dgeequ(a, radix)
    r2 = oracle.ref_dgeequ(a, radix)
    for u, v in zip(r1, r2):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    x1, e1 = oracle.dlaqge(a, *r1[:5])
    x2, e2 = oracle.ref_dlaqge(a, *r2[:5])
    assert e1 == e2 and np.array_equal(bits(x1), bits(x2))


@ref
def test_reference_error_exits_record_xerbla():
    """DERRGE's illegal calls (TESTING/LIN/derrge.f:131-184) against the real
    reference with the recording XERBLA -- the contract the library copies."""
    a = np.zeros(4)
    ip = np.zeros(2, dtype=np.int32)
    b = np.zeros(4)
    for args, expect in [((-1, 0, 1), 1), ((0, -1, 1), 2), ((2, 1, 1), 4)]:
        oracle.ref_xerbla_reset()
        info = oracle.ref_dgetrf_raw(args[0], args[1], a.ctypes.data, args[2], ip.ctypes.data)
        assert info == -expect
        assert oracle.ref_xerbla_last()[:2] == ("DGETRF", expect)
    for args, expect in [(("/", 0, 0, 1, 1), 1), (("N", -1, 0, 1, 1), 2), (("N", 0, -1, 1, 1), 3),
                         (("N", 2, 1, 1, 2), 5), (("N", 2, 1, 2, 1), 8)]:
        oracle.ref_xerbla_reset()
        t, n, nrhs, lda, ldb = args
        info = oracle.ref_dgetrs_raw(t, n, nrhs, a.ctypes.data, lda, ip.ctypes.data, b.ctypes.data, ldb)
        assert info == -expect
        assert oracle.ref_xerbla_last()[:2] == ("DGETRS", expect)


# ------------------------------------------------- closed forms, brute force

def brute_lu(a):
    """Textbook LU with partial pivoting (first max |.|), pure Python."""
    a = [list(map(float, row)) for row in np.asarray(a)]
    m, n = len(a), len(a[0])
    piv = []
    for k in range(min(m, n)):
        p = k
        for i in range(k + 1, m):
            if abs(a[i][k]) > abs(a[p][k]):
                p = i
        piv.append(p + 1)
        a[k], a[p] = a[p], a[k]
        if a[k][k] != 0:
            for i in range(k + 1, m):
                a[i][k] /= a[k][k]
        for i in range(k + 1, m):
            for j in range(k + 1, n):
                a[i][j] -= a[i][k] * a[k][j]
    return np.array(a), np.array(piv, dtype=np.int32)


@pytest.mark.parametrize("m,n,seed", [(1, 1, 0), (2, 2, 1), (3, 3, 2), (4, 6, 3), (6, 4, 4), (6, 6, 5)])
def test_dgetrf_vs_brute_force(m, n, seed):
    a = la.matrix(2, seed, m, n)
    lu, p, info = oracle.dgetrf(a)
    blu, bp = brute_lu(a)
    assert info == 0
    assert np.array_equal(p, bp)
    np.testing.assert_allclose(lu, blu, rtol=1e-12, atol=1e-13)


def test_lu_reconstructs_pa():
    n = 120
    a = la.matrix(3, 77, n, n)
    lu, p, _ = oracle.dgetrf(a)
    l = np.tril(lu, -1) + np.eye(n)
    u = np.triu(lu)
    pa = oracle.dlaswp(a, 1, n, p, 1)
    np.testing.assert_allclose(l @ u, pa, atol=1e-12)
    assert oracle.dget01(a, lu, p) < 30


def test_dget_ratios_closed_form():
    # exact factors -> DGET01 = 0; perturb one entry by d -> ||.||_1 / (n ||A||_1 eps)
    a = np.asfortranarray(np.array([[4.0, 2.0], [2.0, 3.0]]))
    lu, p, _ = oracle.dgetrf(a)
    assert oracle.dget01(a, lu, p) == 0.0
    lu2 = lu.copy()
    lu2[1, 1] += 1e-12
    anorm = 5.0 + 1e-16
    expect = (1e-12 / 2) / 6.0 / 2.0 ** -53
    assert abs(oracle.dget01(a, lu2, p) - expect) / expect < 1e-3
    # DGET02 for exact x is 0; for a residual r: max_j ||r_j||_1/(||A||_1 ||x_j||_1 eps)
    x = np.asfortranarray(np.array([[1.0], [2.0]]))
    b = a @ x
    assert oracle.dget02(a, x, b) == 0.0
    b2 = b.copy()
    b2[0, 0] += 3e-13
    d = b2[0, 0] - b[0, 0]  # the representable perturbation (exact)
    assert abs(oracle.dget02(a, x, b2) / (d / (6.0 * 3.0) / 2.0 ** -53) - 1) < 1e-12
    assert abs(oracle.dget08(a, x, b2) / (d / (6.0 * 2.0) / 2.0 ** -53) - 1) < 1e-12
    assert oracle.dlange("1", a) == 6.0 and oracle.dlange("I", a) == 6.0 and oracle.dlange("M", a) == 4.0


def test_dlange_frobenius():
    """DLANGE('F') (SRC/dlange.f:190-200, DLASSQ of SRC/dlassq.f90): closed
    forms -- sqrt(m n) for a matrix of ones, exact for a scaled 3-4-5
    triangle at magnitudes where the plain sum of squares over- or
    underflows (Blue's big and small accumulators) -- and, where the
    reference build exists, bitwise equal to it on random, huge, tiny and
    mixed matrices."""
    for m, n in ((1, 1), (3, 7), (64, 5)):
        assert oracle.dlange("F", np.ones((m, n))) == pytest.approx(np.sqrt(m * n), rel=1e-15)
    for s in (1.0, 1e300, 1e-300, 2.0 ** -1070):
        a = np.asfortranarray([[3.0 * s], [4.0 * s]])
        assert oracle.dlange("F", a) == pytest.approx(5.0 * s, rel=1e-15), s
        assert oracle.dlange("E", a) == oracle.dlange("F", a)
    # the unscaled sum of squares would give 0 and inf at these magnitudes
    assert np.sqrt(np.sum(np.array([3e-300, 4e-300]) ** 2)) == 0.0
    assert np.isinf(np.sqrt(np.sum(np.array([3e300, 4e300]) ** 2)))
    if oracle.ref_available():
        rng = np.random.default_rng(7)
        mats = [rng.standard_normal((37, 20)), rng.standard_normal((300, 3)) * 1e250,
                rng.standard_normal((5, 9)) * 1e-250, np.zeros((4, 4))]
        mixed = rng.standard_normal((30, 30))
        mixed[::3] *= 1e200
        mixed[1::3] *= 1e-200
        mats.append(mixed)
        for a in mats:
            a = np.asfortranarray(a)
            assert oracle.dlange("F", a) == oracle.ref_dlange("F", a)
            for norm in "M1OI":
                assert oracle.dlange(norm, a) == oracle.ref_dlange(norm, a)


def test_dchkeq_closed_form():
    """TESTING/LIN/dchkeq.f:107-173: A(i,j) = 10^(i+j) (-1)^(i+j)."""
    eps = 2.0 ** -52
    for m in range(1, 6):
        for n in range(1, 6):
            a = np.asfortranarray([[10.0 ** (i + j) * (-1) ** (i + j) for j in range(1, n + 1)]
                                   for i in range(1, m + 1)])
            r, c, rowcnd, colcnd, amax, info = oracle.dgeequ(a)
            assert info == 0
            worst = max(abs(rowcnd - 10.0 ** -(m - 1)) / 10.0 ** -(m - 1),
                        abs(colcnd - 10.0 ** -(n - 1)) / 10.0 ** -(n - 1),
                        abs(amax - 10.0 ** (n + m)) / 10.0 ** (n + m),
                        max(abs(r[i - 1] - 10.0 ** -(i + n)) / 10.0 ** -(i + n) for i in range(1, m + 1)),
                        max(abs(c[j - 1] - 10.0 ** (n - j)) / 10.0 ** (n - j) for j in range(1, n + 1)))
            assert worst / eps < 30
    a = np.ones((5, 5), order="F")
    a[3, :] = 0
    assert oracle.dgeequ(a)[5] == 4
    a = np.ones((5, 5), order="F")
    a[:, 3] = 0
    assert oracle.dgeequ(a)[5] == 5 + 4
    # DGEEQUB factors are exact powers of two
    r, c, *_ = oracle.dgeequ(la.matrix(3, 8, 20, 20) * 1e7, radix=True)
    assert np.all(np.frexp(r)[0] == 0.5) and np.all(np.frexp(c)[0] == 0.5)


def test_dsgesv_iter_codes():
    n = 40
    b = la.matrix(2, 1, n, 1)
    # well conditioned -> ITER >= 0
    assert oracle.dsgesv(la.matrix(2, 2, n, n), b)[2] >= 0
    # entry beyond FLT_MAX -> ITER = -2, FP64 fallback solves it
    a = la.matrix(2, 3, n, n)
    a[3, 4] = 1e39
    x, _, it, info, _ = oracle.dsgesv(a, b)
    assert it == -2 and info == 0 and oracle.dget02(a, x, b) < 30
    # singular in FP32 (exact zero column) -> ITER = -3 and FP64 INFO
    a = la.matrix(2, 4, n, n)
    a[:, 7] = 0
    x, _, it, info, _ = oracle.dsgesv(a, b)
    assert it == -3 and info == 8
    # cond ~ 1/eps_single**2: refinement cannot converge -> -ITERMAX-1
    q1, _ = np.linalg.qr(la.matrix(3, 5, n, n))
    q2, _ = np.linalg.qr(la.matrix(3, 6, n, n))
    a = np.asfortranarray(q1 @ np.diag(np.logspace(0, 14, n)) @ q2.T)
    x, _, it, info, _ = oracle.dsgesv(a, b)
    assert it == -31 and info == 0


def test_planted_mistakes_are_caught():
    """A checker that accepts wrong factors pins nothing: plant typical
    mistakes in a copy of correct factors and watch the checks fail."""
    n = 60
    a = la.matrix(2, 9, n, n)
    lu, p, _ = oracle.dgetrf(a)
    assert oracle.dget01(a, lu, p) < 30
    bad = lu.copy()
    bad[10, 3] = -bad[10, 3]                       # sign error in L
    assert oracle.dget01(a, bad, p) > 30
    bad = lu.copy()
    bad[5, 40] *= 1 + 1e-9                          # perturbed U entry
    assert oracle.dget01(a, bad, p) > 30
    p2 = p.copy()
    p2[[4, 5]] = p2[[5, 4]]                         # transposed interchanges
    assert oracle.dget01(a, lu, p2) > 30
    assert oracle.dget01(a, np.asfortranarray(lu.T), p) > 30   # transposed operand
    b = la.matrix(2, 10, n, 2)
    x, _ = oracle.dgetrs(lu, p, b)
    assert oracle.dget02(a, x, b) < 30
    assert oracle.dget02(a, x * (1 + 1e-10), b) > 30
    assert oracle.dget02(np.asfortranarray(a.T), x, b) > 30


# ---------------------------------------------------------------- golden vectors

def test_oracle_reproduces_reference_golden_n1024():
    """tests/golden/dgetrf_n1024_u.npz holds reference DGETRF IPIV and DGETRS x
    (flang build of the reference sources, tests/golden/make_golden.py) for
    A uniform(-1,1) seed 1988, n = 1024, b seed 1989."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "dgetrf_n1024_u.npz"))
    n = 1024
    a = la.matrix(2, 1988, n, n)
    b = la.matrix(2, 1989, n, 1)
    lu, ipiv, info = oracle.dgetrf(a)
    x, _ = oracle.dgetrs(lu, ipiv, b)
    assert info == int(g["info"]) == 0
    assert np.array_equal(ipiv, g["ipiv"])
    assert np.array_equal(x[:, 0].view(np.uint64), g["x"].view(np.uint64))


@ref
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (17, 2), (120, 3)])
def test_dgerfs_bitwise_vs_reference(trans, n, nrhs):
    """DGERFS (refined X, FERR, BERR) of the restatement equals the reference
    build bit for bit, starting from an FP32-accurate solution so that the
    refinement loop and DLACN2's iterations all run."""
    a = la.matrix(2, 3000 + n, n, n)
    b = la.matrix(2, 3001 + n, n, nrhs)
    lu, p, info = oracle.dgetrf(a)
    assert info == 0
    x0, _ = oracle.dgetrs(lu, p, b, trans)
    x0 = x0.astype(np.float32).astype(np.float64)
    r1 = oracle.dgerfs(a, lu, p, b, x0, trans)
    r2 = oracle.ref_dgerfs(a, lu, p, b, x0, trans)
    assert r1[3] == r2[3] == 0
    for u, v in zip(r1[:3], r2[:3]):
        assert np.array_equal(bits(u), bits(v))


def test_dgerfs_bounds():
    """DGET07's checks (TESTING/LIN/dget07.f): BERR below (n+1) eps-scale and
    FERR bounding the true forward error of a known solution."""
    n, nrhs = 80, 2
    a = la.matrix(2, 41, n, n)
    xact = la.matrix(2, 42, n, nrhs)
    b = np.asfortranarray(a @ xact)
    lu, p, _ = oracle.dgetrf(a)
    x0, _ = oracle.dgetrs(lu, p, b)
    x, ferr, berr, info = oracle.dgerfs(a, lu, p, b, x0)
    assert info == 0
    eps = 2.0 ** -53
    assert np.all(berr <= (n + 1) * eps * 30)
    err = np.max(np.abs(x - xact), axis=0) / np.max(np.abs(x), axis=0)
    # the bound of the computed problem holds up to the O(eps) perturbation b carries
    assert np.all(err <= ferr + 30 * eps * np.linalg.cond(a, np.inf))
    assert np.all(ferr < 1e-10)


@pytest.mark.parametrize("trans", ["N", "T"])
def test_dget04_dget07_vs_reference(trans):
    """DGET04 / DGET07 restated (TESTING/LIN/dget04.f, dget07.f) equal the
    reference's own routines built into oracle/_ref, bit for bit."""
    if not oracle.ref_available():
        pytest.skip("reference build absent")
    rng = np.random.default_rng(7)
    for n, nrhs in [(1, 1), (9, 2), (60, 3)]:
        a = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
        xact = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)))
        x = np.asfortranarray(xact + rng.uniform(-1, 1, (n, nrhs)) * 1e-11)
        b = np.asfortranarray((a.T if trans == "T" else a) @ xact)
        ferr = rng.uniform(1e-12, 1e-9, nrhs)
        berr = rng.uniform(0, 1e-15, nrhs)
        for chk in (True, False):
            assert oracle.dget07(a, b, x, xact, ferr, berr, trans, chk) == \
                oracle.ref_dget07(a, b, x, xact, ferr, berr, trans, chk)
        for rcond in (0.25, 1e-7, -1.0):
            assert oracle.dget04(x, xact, rcond) == oracle.ref_dget04(x, xact, rcond)


def test_dget04_dget07_closed_forms():
    eps = 2.0 ** -53
    n = 8
    xact = np.ones((n, 1), order="F") * 2.0
    x = xact.copy(order="F")
    x[3, 0] += 2.0 ** -20
    # DGET04: (max|x - xact| / max|xact|) * RCOND / eps
    assert oracle.dget04(x, xact, 0.5) == (2.0 ** -20 / 2.0) * 0.5 / eps
    assert oracle.dget04(x, xact, -1.0) == 1.0 / eps
    assert oracle.dget04(x, np.zeros((n, 1), order="F"), 0.5) == 1.0 / eps
    # a planted mistake -- the error taken relative to x instead of xact -- changes it
    assert oracle.dget04(x, xact, 0.5) != (2.0 ** -20 / np.abs(x).max()) * 0.5 / eps
    # DGET07 on A = I: |b| + |A||x| = 4 everywhere, so BERR = (n+1) eps gives ratio 1
    a = np.eye(n, order="F")
    b = xact.copy(order="F")
    r0, r1 = oracle.dget07(a, b, x, xact, np.array([1e-3]), np.array([(n + 1) * eps]))
    assert r1 == pytest.approx(1.0, rel=1e-15)
    assert r0 == pytest.approx((2.0 ** -20 / (2.0 + 2.0 ** -20)) / 1e-3, rel=1e-15)
    # FERR smaller than the true error: 1/eps
    r0, _ = oracle.dget07(a, b, x, xact, np.array([1e-9]), np.array([0.0]))
    assert r0 == 1.0 / eps
    # without the FERR check the first ratio is 0
    assert oracle.dget07(a, b, x, xact, np.array([1e-9]), np.array([0.0]), chkferr=False)[0] == 0.0
