"""GPU checks of the routines around the solve: DLANGE, DGECON, DGETRI and the
expert driver DGESVX, against the reference's own routines built from its
sources (oracle/_ref: SRC/dlange.f, dgecon.f, dgetri.f, dgesvx.f with their
callees) on the same inputs.

Bars:
  * DLANGE: 'M', '1', 'I' bitwise the reference's (same summation order);
    'F' within 1e-14 relative (the reference's DLASSQ scales differently);
    a NaN anywhere gives NaN, as SRC/dlange.f's DISNAN tests do;
  * DGECON: the same DLACN2 iteration on factors equal to the reference's
    (the oracle's DGETRF), so RCOND agrees to rounding (1e-10 relative);
    singular factors give RCOND = 0, and the error exits follow
    SRC/dgecon.f:157-185;
  * DGETRI: INFO the first zero of U's diagonal; the inverse's DGET03 ratio
    ||I - A inv(A)||_1 / (n ||A||_1 ||inv(A)||_1 eps) below 30
    (TESTING/LIN/dget03.f) and agreement with the reference's inverse to
    cond(A) n eps;
  * DGESVX: EQUED, the equilibrated A and the scaled B bitwise; IPIV equal;
    RCOND and RPVGRW to rounding; X within DGET02 < 30 and the reference's
    X to cond(A) n eps; FERR / BERR through DGET07 (TESTING/LIN/dget07.f)
    below 30; INFO = N+1 for a matrix with RCOND < eps.
"""
import numpy as np
import pytest

import lapack_amd as la
import oracle

from gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def need_ref():
    if not oracle.ref_available():
        pytest.skip("reference build absent")


@pytest.mark.parametrize("shape", [(1, 1), (37, 20), (300, 300), (1000, 257)])
def test_dlange_norms(dev, shape):
    m, n = shape
    a = la.matrix(3, 50 + m + n, m, n) * 3.0
    d_a = to_dev(a)
    for norm in "M1OI":
        v = la.dlange(norm, m, n, d_a, m)
        assert v == oracle.dlange(norm, a), norm
        assert la.dlange(norm, m, n, a, m) == v  # host pointer
    f = la.dlange("F", m, n, d_a, m)
    assert f == pytest.approx(oracle.dlange("F", a), rel=1e-14)
    a[m // 2, n // 2] = np.nan
    d_a = to_dev(a)
    for norm in "M1IF":
        assert np.isnan(la.dlange(norm, m, n, d_a, m)), norm


@pytest.mark.parametrize("norm", ["1", "I"])
@pytest.mark.parametrize("n", [1, 50, 300, 1500])
def test_dgecon_vs_reference(dev, norm, n):
    need_ref()
    a = la.matrix(2, 60 + n, n, n)
    lu, p, _ = oracle.dgetrf(a)
    anorm = oracle.dlange(norm, a)
    rc_ref, info_ref = oracle.ref_dgecon(lu, anorm, norm)
    info, rc = la.dgecon(norm, n, to_dev(lu), n, anorm)
    assert info == info_ref == 0
    assert rc == pytest.approx(rc_ref, rel=1e-10)


def test_dgecon_singular_and_errors(dev):
    need_ref()
    n = 40
    a = la.matrix(2, 61, n, n)
    a[:, 5] = 0.0
    lu, p, info = oracle.dgetrf(a)
    assert info == 6
    anorm = oracle.dlange("1", a)
    rc_ref, _ = oracle.ref_dgecon(lu, anorm, "1")
    info, rc = la.dgecon("1", n, to_dev(lu), n, anorm)
    assert rc == rc_ref == 0.0
    assert la.dgecon("X", n, lu, n, 1.0)[0] == -1
    assert la.dgecon("1", -1, lu, n, 1.0)[0] == -2
    assert la.dgecon("1", n, lu, n - 1, 1.0)[0] == -4
    assert la.dgecon("1", n, lu, n, -1.0)[0] == -5
    assert la.dgecon("1", n, lu, n, 0.0) == (0, 0.0)
    assert la.dgecon("1", 0, lu, 1, 1.0) == (0, 1.0)


@pytest.mark.parametrize("n", [1, 7, 64, 300, 1100])
def test_dgetri_vs_reference(dev, n):
    need_ref()
    a = la.matrix(2, 70 + n, n, n)
    lu, p, _ = oracle.dgetrf(a)
    inv_ref, info_ref = oracle.ref_dgetri(lu, p)
    d_lu = to_dev(lu)
    assert la.dgetri(n, d_lu, n, to_dev(p.reshape(-1, 1).astype(np.int32))) == info_ref == 0
    inv = from_dev(d_lu, (n, n))
    an = np.abs(a).sum(0).max()
    ain = np.abs(inv).sum(0).max()
    dget03 = np.abs(np.eye(n) - a @ inv).sum(0).max() / (n * an * ain * EPS)
    assert dget03 < 30.0, dget03
    kappa = an * ain
    assert np.abs(inv - inv_ref).max() <= kappa * n * EPS * np.abs(inv_ref).max()


def test_dgetri_singular(dev):
    n = 30
    a = la.matrix(2, 71, n, n)
    a[:, 9] = 0.0
    lu, p, info = oracle.dgetrf(a)
    d_lu = to_dev(lu)
    assert la.dgetri(n, d_lu, n, to_dev(p.reshape(-1, 1).astype(np.int32))) == info == 10
    assert np.array_equal(from_dev(d_lu, (n, n)), lu)  # A untouched (SRC/dgetri.f:150-155)


def _gesvx(fact, trans, a, b, af=None, ipiv=None, equed="N", rr=None, cc=None, device=True):
    n = a.shape[0]
    nrhs = b.shape[1]
    aa = np.array(a, order="F")
    bb = np.array(b, order="F")
    ff = np.zeros((n, n), order="F") if af is None else np.array(af, order="F")
    pp = np.zeros(n, dtype=np.int32) if ipiv is None else np.array(ipiv, dtype=np.int32)
    r = np.zeros(n) if rr is None else np.array(rr)
    c = np.zeros(n) if cc is None else np.array(cc)
    x = np.zeros((n, nrhs), order="F")
    ferr = np.zeros(nrhs)
    berr = np.zeros(nrhs)
    if device:
        d = {k: to_dev(v) for k, v in dict(a=aa, b=bb, f=ff, x=x).items()}
        d_p = to_dev(pp.reshape(-1, 1))
        info, eq, rcond, rpvgrw = la.dgesvx(fact, trans, n, nrhs, d["a"], n, d["f"], n, d_p, equed, r, c, d["b"], n,
                                            d["x"], n, ferr, berr)
        aa, bb = from_dev(d["a"], (n, n)), from_dev(d["b"], (n, nrhs))
        ff, x = from_dev(d["f"], (n, n)), from_dev(d["x"], (n, nrhs))
        pp = d_p.cpu().numpy().reshape(-1)
    else:
        info, eq, rcond, rpvgrw = la.dgesvx(fact, trans, n, nrhs, aa, n, ff, n, pp, equed, r, c, bb, n, x, n, ferr,
                                            berr)
    return dict(info=info, equed=eq, rcond=rcond, rpvgrw=rpvgrw, a=aa, b=bb, af=ff, ipiv=pp, x=x, r=r, c=c,
                ferr=ferr, berr=berr)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("fact", ["N", "E", "F"])
@pytest.mark.parametrize("n", [40, 600])
def test_dgesvx_vs_reference(dev, fact, trans, n):
    need_ref()
    # badly scaled rows and columns, so FACT = 'E' equilibrates ('B')
    a = la.matrix(2, 80 + n, n, n) * np.logspace(-6, 6, n)[:, None] * np.logspace(3, -3, n)[None, :]
    a = np.asfortranarray(a)
    xact = la.matrix(2, 81 + n, n, 3)
    b = np.asfortranarray((a.T if trans == "T" else a) @ xact)
    kw = {}
    if fact == "F":
        # factors of the equilibrated matrix, as a previous FACT = 'E' call leaves them
        pre = oracle.ref_dgesvx("E", trans, a, b)
        kw = dict(af=pre["af"], ipiv=pre["ipiv"], equed=pre["equed"], rr=pre["r"], cc=pre["c"])
        a_in = pre["a"]
    else:
        a_in = a
    ref = oracle.ref_dgesvx(fact, trans, a_in, b, **kw)
    got = _gesvx(fact, trans, a_in, b, **kw, device=(n > 100))
    assert got["info"] == ref["info"]
    assert got["equed"] == ref["equed"]
    assert np.array_equal(got["a"], ref["a"]) and np.array_equal(got["b"], ref["b"])
    if fact != "F":
        assert np.array_equal(got["r"], ref["r"]) and np.array_equal(got["c"], ref["c"])
    assert np.array_equal(got["ipiv"], ref["ipiv"])
    assert got["rcond"] == pytest.approx(ref["rcond"], rel=1e-8)
    assert got["rpvgrw"] == pytest.approx(ref["rpvgrw"], rel=1e-12)
    # the solution of the original system
    x, xr = got["x"], ref["x"]
    assert oracle.dget02(a, x, b, trans) < 30.0
    # DDRVGE's forward-error test (TESTING/LIN/ddrvge.f:598-612): DGET04
    # against XACT with the reciprocal condition number of the matrix as
    # passed -- the non-equilibrated one once EQUED != 'N' -- in the 1-norm
    # for 'N' and the infinity norm for 'T' (ddrvge.f:415-443)
    ord_ = 1 if trans == "N" else np.inf
    rcondc = 1.0 / (np.linalg.norm(a, ord_) * np.linalg.norm(np.linalg.inv(a), ord_))
    r_got, r_ref = oracle.dget04(x, xact, rcondc), oracle.dget04(xr, xact, rcondc)
    assert r_ref < 30.0 and r_got < 30.0, (r_got, r_ref)
    # the solved system is the equilibrated one: (R A C) y = R b, x = C y for
    # 'N'; (R A C)^T y = C b, x = R y for 'T' (SRC/dgesvx.f:537-560).  y agrees
    # with the reference's to the equilibrated matrix's condition
    scale = None
    if trans == "N" and got["equed"] in ("C", "B"):
        scale = got["c"]
    if trans == "T" and got["equed"] in ("R", "B"):
        scale = got["r"]
    if scale is not None:
        y, yr = x / scale[:, None], xr / scale[:, None]
        assert np.abs(y - yr).max() <= 30.0 * n * EPS / got["rcond"] * np.abs(yr).max()
    r0, r1 = oracle.dget07(a, b, x, xact, got["ferr"], got["berr"], trans)
    assert r0 < 30.0 and r1 < 30.0, (r0, r1)


def test_dgesvx_ill_conditioned_info(dev):
    need_ref()
    n = 60
    rng = np.random.default_rng(9)
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = np.asfortranarray((q1 * np.logspace(0, -17, n)) @ q2.T)
    b = np.asfortranarray(rng.standard_normal((n, 1)))
    ref = oracle.ref_dgesvx("N", "N", a, b)
    got = _gesvx("N", "N", a, b)
    assert ref["info"] == got["info"] == n + 1


def test_dgesvx_singular_and_errors(dev):
    need_ref()
    n = 30
    a = la.matrix(2, 91, n, n)
    a[:, 4] = 0.0
    b = la.matrix(2, 92, n, 1)
    ref = oracle.ref_dgesvx("N", "N", a, b)
    got = _gesvx("N", "N", a, b)
    assert got["info"] == ref["info"] == 5
    assert got["rcond"] == 0.0
    assert got["rpvgrw"] == pytest.approx(ref["rpvgrw"], rel=1e-12)
    z = np.zeros((4, 4), order="F")
    p = np.zeros(4, dtype=np.int32)
    v = np.zeros(4)
    args = lambda **k: dict(dict(fact="N", trans="N", n=4, nrhs=1, lda=4, ldaf=4, ldb=4, ldx=4), **k)
    for kw, code in [(dict(fact="X"), -1), (dict(trans="X"), -2), (dict(n=-1), -3), (dict(nrhs=-1), -4),
                     (dict(lda=3), -6), (dict(ldaf=3), -8), (dict(ldb=3), -14), (dict(ldx=3), -16)]:
        k = args(**kw)
        info = la.dgesvx(k["fact"], k["trans"], k["n"], k["nrhs"], z, k["lda"], z, k["ldaf"], p, "N", v, v, z,
                         k["ldb"], z, k["ldx"], v, v)[0]
        assert info == code, (kw, info)
    # FACT = 'F' with EQUED = 'R' and a non-positive scale factor: -11
    assert la.dgesvx("F", "N", 4, 1, z, 4, z, 4, p, "R", v, v, z, 4, z, 4, v, v)[0] == -11
