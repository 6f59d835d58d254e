"""oracle -- the CPU checker for the LU solve path.  TEST INFRASTRUCTURE ONLY.

Two implementations, loaded through ctypes:
  * liboracle.so       -- lu_oracle.cpp, a loop-for-loop C++ restatement of
                          the reference routines (file:line cited there);
  * _ref/libreflapack.so -- the reference Fortran itself compiled with flang
                          from /root/reference (oracle/Makefile), present
                          wherever it was built.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, always as the checker, never as the measured or shipped path.

All arrays are numpy, Fortran (column-major) order; IPIV int32 1-based.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libreflapack.so")
EPS = 2.0 ** -53

_o = None
_r = None


def _ld(path):
    return ctypes.CDLL(path)


def o():
    """The C++ restatement."""
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        L = _ld(ORACLE_SO)
        vp, i, d, f, ch = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_char
        sig = {
            "oracle_idamax": (i, [i, vp, i]),
            "oracle_isamax": (i, [i, vp, i]),
            "oracle_dscal": (None, [i, d, vp, i]),
            "oracle_dlaswp": (None, [i, vp, i, i, i, vp, i]),
            "oracle_dtrsm_left": (None, [ch, ch, ch, i, i, d, vp, i, vp, i]),
            "oracle_dgemm": (None, [ch, ch, i, i, i, d, vp, i, vp, i, d, vp, i]),
            "oracle_dgetrf2": (i, [i, i, vp, i, vp]),
            "oracle_dgetrf": (i, [i, i, vp, i, vp, i]),
            "oracle_dgetrs": (i, [ch, i, i, vp, i, vp, vp, i]),
            "oracle_dlange": (d, [ch, i, i, vp, i]),
            "oracle_slaswp": (None, [i, vp, i, i, i, vp, i]),
            "oracle_sgemm": (None, [ch, ch, i, i, i, f, vp, i, vp, i, f, vp, i]),
            "oracle_strsm_left": (None, [ch, ch, ch, i, i, f, vp, i, vp, i]),
            "oracle_sgetrf": (i, [i, i, vp, i, vp, i]),
            "oracle_sgetrs": (i, [ch, i, i, vp, i, vp, vp, i]),
            "oracle_dlag2s": (i, [i, i, vp, i, vp, i]),
            "oracle_slag2d": (None, [i, i, vp, i, vp, i]),
            "oracle_dsgesv": (i, [i, i, vp, i, vp, vp, i, vp, i, vp, vp, vp, i]),
            "oracle_dgeequ": (i, [i, i, vp, i, vp, vp, vp, vp, vp, i]),
            "oracle_dlaqge": (ch, [i, i, vp, i, vp, vp, d, d, d]),
            "oracle_dget01": (d, [i, i, vp, i, vp, i, vp]),
            "oracle_dget02": (d, [ch, i, i, i, vp, i, vp, i, vp, i]),
            "oracle_dget08": (d, [ch, i, i, i, vp, i, vp, i, vp, i]),
            "oracle_dgerfs": (i, [ch, i, i, vp, i, vp, i, vp, vp, i, vp, i, vp, vp]),
            "oracle_dget04": (d, [i, i, vp, i, vp, i, d]),
            "oracle_dget07": (None, [ch, i, i, vp, i, vp, i, vp, i, vp, i, vp, i, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _o = L
    return _o


def ref_available():
    return os.path.exists(REF_SO)


def r():
    """The flang-built reference (Fortran symbols, arguments by reference)."""
    global _r
    if _r is None:
        if not ref_available():
            raise ImportError(f"{REF_SO} missing (built only where /root/reference exists)")
        _r = _ld(REF_SO)
    return _r


def _p(a):
    return a.ctypes.data if a is not None else None


def _chk(a, dtype):
    assert a.dtype == dtype, a.dtype
    assert a.flags.f_contiguous or a.ndim == 1
    return a


# ---- restatement wrappers (copying; inputs untouched) -----------------------

def dgetrf(a, nb=64):
    """(LU, IPIV, INFO) of the reference DGETRF with block size nb (ILAENV: 64)."""
    lu = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = lu.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    info = o().oracle_dgetrf(m, n, _p(lu), max(1, m), _p(ipiv), nb)
    return lu, ipiv[:min(m, n)], info


def sgetrf(a, nb=64):
    lu = np.array(a, dtype=np.float32, order="F", copy=True)
    m, n = lu.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    info = o().oracle_sgetrf(m, n, _p(lu), max(1, m), _p(ipiv), nb)
    return lu, ipiv[:min(m, n)], info


def dgetrf2(a):
    lu = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = lu.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    info = o().oracle_dgetrf2(m, n, _p(lu), max(1, m), _p(ipiv))
    return lu, ipiv[:min(m, n)], info


def dgetrs(lu, ipiv, b, trans="N"):
    x = np.array(b, dtype=np.float64, order="F", copy=True)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n = lu.shape[0]
    lu = np.asfortranarray(lu, dtype=np.float64)
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    info = o().oracle_dgetrs(trans.encode(), n, x.shape[1], _p(lu), max(1, n), _p(ip), _p(x), max(1, n))
    return x, info


def sgetrs(lu, ipiv, b, trans="N"):
    x = np.array(b, dtype=np.float32, order="F", copy=True)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n = lu.shape[0]
    lu = np.asfortranarray(lu, dtype=np.float32)
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    info = o().oracle_sgetrs(trans.encode(), n, x.shape[1], _p(lu), max(1, n), _p(ip), _p(x), max(1, n))
    return x, info


def dlaswp(a, k1, k2, ipiv, incx=1):
    x = np.array(a, dtype=np.float64, order="F", copy=True)
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    o().oracle_dlaswp(x.shape[1], _p(x), max(1, x.shape[0]), k1, k2, _p(ip), incx)
    return x


def dgemm(ta, tb, alpha, a, b, beta, c):
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    cc = np.array(c, dtype=np.float64, order="F", copy=True)
    m, n = cc.shape
    k = a.shape[1] if ta.upper() == "N" else a.shape[0]
    o().oracle_dgemm(ta.encode(), tb.encode(), m, n, k, alpha, _p(a), max(1, a.shape[0]), _p(b),
                     max(1, b.shape[0]), beta, _p(cc), max(1, m))
    return cc


def dtrsm_left(uplo, trans, diag, alpha, a, b):
    a = np.asfortranarray(a, dtype=np.float64)
    x = np.array(b, dtype=np.float64, order="F", copy=True)
    m, n = x.shape
    o().oracle_dtrsm_left(uplo.encode(), trans.encode(), diag.encode(), m, n, alpha, _p(a), max(1, a.shape[0]),
                          _p(x), max(1, m))
    return x


def idamax(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return o().oracle_idamax(len(x), _p(x), 1)


def dlange(norm, a):
    a = np.asfortranarray(a, dtype=np.float64)
    m, n = a.shape
    return o().oracle_dlange(norm.encode(), m, n, _p(a), max(1, m))


def dsgesv(a, b, nb=64):
    """(X, IPIV, ITER, INFO, A_after) of the reference DSGESV."""
    aa = np.array(a, dtype=np.float64, order="F", copy=True)
    bb = np.asfortranarray(b, dtype=np.float64)
    if bb.ndim == 1:
        bb = bb.reshape(-1, 1, order="F")
    n, nrhs = bb.shape
    x = np.zeros((n, nrhs), dtype=np.float64, order="F")
    ipiv = np.zeros(max(1, n), dtype=np.int32)
    work = np.zeros(max(1, n * nrhs), dtype=np.float64)
    swork = np.zeros(max(1, n * (n + nrhs)), dtype=np.float32)
    it = ctypes.c_int(0)
    info = o().oracle_dsgesv(n, nrhs, _p(aa), max(1, n), _p(ipiv), _p(bb), max(1, n), _p(x), max(1, n),
                             _p(work), _p(swork), ctypes.addressof(it), nb)
    return x, ipiv[:n], it.value, info, aa


def dgeequ(a, radix=False):
    a = np.asfortranarray(a, dtype=np.float64)
    m, n = a.shape
    rr = np.zeros(max(1, m))
    cc = np.zeros(max(1, n))
    rowcnd, colcnd, amax = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    info = o().oracle_dgeequ(m, n, _p(a), max(1, m), _p(rr), _p(cc), ctypes.addressof(rowcnd),
                             ctypes.addressof(colcnd), ctypes.addressof(amax), 1 if radix else 0)
    return rr[:m], cc[:n], rowcnd.value, colcnd.value, amax.value, info


def dlaqge(a, rr, cc, rowcnd, colcnd, amax):
    x = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = x.shape
    eq = o().oracle_dlaqge(m, n, _p(x), max(1, m), _p(np.ascontiguousarray(rr)), _p(np.ascontiguousarray(cc)),
                           rowcnd, colcnd, amax)
    return x, eq.decode()


def dget01(a, lu, ipiv):
    """||L*U - P*A||_1 / (n ||A||_1 eps) (TESTING/LIN/dget01.f)."""
    a = np.asfortranarray(a, dtype=np.float64)
    lu = np.asfortranarray(lu, dtype=np.float64)
    m, n = a.shape
    return o().oracle_dget01(m, n, _p(a), max(1, m), _p(lu), max(1, m), _p(np.ascontiguousarray(ipiv, np.int32)))


def dget02(a, x, b, trans="N"):
    """max_j ||b - op(A) x||_1 / (||A|| ||x||_1 eps) (TESTING/LIN/dget02.f)."""
    a = np.asfortranarray(a, dtype=np.float64)
    x = np.asfortranarray(x, dtype=np.float64).reshape(a.shape[1], -1, order="F")
    b = np.asfortranarray(b, dtype=np.float64).reshape(a.shape[0], -1, order="F")
    m, n = a.shape
    return o().oracle_dget02(trans.encode(), m, n, x.shape[1], _p(a), max(1, m), _p(x), max(1, x.shape[0]),
                             _p(b), max(1, b.shape[0]))


def dget08(a, x, b, trans="N"):
    a = np.asfortranarray(a, dtype=np.float64)
    x = np.asfortranarray(x, dtype=np.float64).reshape(a.shape[1], -1, order="F")
    b = np.asfortranarray(b, dtype=np.float64).reshape(a.shape[0], -1, order="F")
    m, n = a.shape
    return o().oracle_dget08(trans.encode(), m, n, x.shape[1], _p(a), max(1, m), _p(x), max(1, x.shape[0]),
                             _p(b), max(1, b.shape[0]))


def dget04(x, xact, rcond):
    """DGET04 (TESTING/LIN/dget04.f): forward error times RCOND over eps."""
    x = np.asfortranarray(x, dtype=np.float64).reshape(len(x), -1, order="F")
    xact = np.asfortranarray(xact, dtype=np.float64).reshape(len(xact), -1, order="F")
    n, nrhs = x.shape
    return o().oracle_dget04(n, nrhs, _p(x), max(1, n), _p(xact), max(1, n), float(rcond))


def dget07(a, b, x, xact, ferr, berr, trans="N", chkferr=True):
    """DGET07 (TESTING/LIN/dget07.f): (FERR ratio, BERR ratio)."""
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    b = np.asfortranarray(b, dtype=np.float64).reshape(n, -1, order="F")
    x = np.asfortranarray(x, dtype=np.float64).reshape(n, -1, order="F")
    xact = np.asfortranarray(xact, dtype=np.float64).reshape(n, -1, order="F")
    ferr = np.ascontiguousarray(ferr, dtype=np.float64)
    berr = np.ascontiguousarray(berr, dtype=np.float64)
    out = np.zeros(2)
    o().oracle_dget07(trans.encode(), n, x.shape[1], _p(a), max(1, n), _p(b), max(1, n), _p(x), max(1, n), _p(xact),
                      max(1, n), _p(ferr), 1 if chkferr else 0, _p(berr), _p(out))
    return float(out[0]), float(out[1])


def dgerfs(a, lu, ipiv, b, x, trans="N"):
    """(X refined, FERR, BERR, INFO) of the reference DGERFS on factors lu/ipiv."""
    a = np.asfortranarray(a, dtype=np.float64)
    lu = np.asfortranarray(lu, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64).reshape(a.shape[0], -1, order="F")
    xx = np.array(x, dtype=np.float64, order="F", copy=True).reshape(a.shape[0], -1, order="F")
    n, nrhs = b.shape
    ferr = np.zeros(max(1, nrhs))
    berr = np.zeros(max(1, nrhs))
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    info = o().oracle_dgerfs(trans.encode(), n, nrhs, _p(a), max(1, n), _p(lu), max(1, n), _p(ip), _p(b),
                             max(1, n), _p(xx), max(1, n), _p(ferr), _p(berr))
    return xx, ferr[:nrhs], berr[:nrhs], info


# ---- the flang-built reference (Fortran ABI) ---------------------------------

def _ri(v):
    return ctypes.byref(ctypes.c_int(v))


def ref_dgetrf(a):
    lu = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = lu.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    info = ctypes.c_int(0)
    r().dgetrf_(_ri(m), _ri(n), ctypes.c_void_p(_p(lu)), _ri(max(1, m)), ctypes.c_void_p(_p(ipiv)),
                ctypes.byref(info))
    return lu, ipiv[:min(m, n)], info.value


def ref_dget04(x, xact, rcond):
    x = np.asfortranarray(x, dtype=np.float64).reshape(len(x), -1, order="F")
    xact = np.asfortranarray(xact, dtype=np.float64).reshape(len(xact), -1, order="F")
    n, nrhs = x.shape
    resid = ctypes.c_double(0)
    r().dget04_(_ri(n), _ri(nrhs), ctypes.c_void_p(_p(x)), _ri(max(1, n)), ctypes.c_void_p(_p(xact)),
                _ri(max(1, n)), ctypes.byref(ctypes.c_double(rcond)), ctypes.byref(resid))
    return resid.value


def ref_dget07(a, b, x, xact, ferr, berr, trans="N", chkferr=True):
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    b = np.asfortranarray(b, dtype=np.float64).reshape(n, -1, order="F")
    x = np.asfortranarray(x, dtype=np.float64).reshape(n, -1, order="F")
    xact = np.asfortranarray(xact, dtype=np.float64).reshape(n, -1, order="F")
    ferr = np.ascontiguousarray(ferr, dtype=np.float64)
    berr = np.ascontiguousarray(berr, dtype=np.float64)
    out = np.zeros(2)
    chk = ctypes.c_int(1 if chkferr else 0)  # Fortran LOGICAL, 4 bytes: .TRUE. = 1 for flang
    r().dget07_(ctypes.c_char_p(trans.encode()), _ri(n), _ri(x.shape[1]), ctypes.c_void_p(_p(a)), _ri(max(1, n)),
                ctypes.c_void_p(_p(b)), _ri(max(1, n)), ctypes.c_void_p(_p(x)), _ri(max(1, n)),
                ctypes.c_void_p(_p(xact)), _ri(max(1, n)), ctypes.c_void_p(_p(ferr)), ctypes.byref(chk),
                ctypes.c_void_p(_p(berr)), ctypes.c_void_p(_p(out)), ctypes.c_size_t(1))
    return float(out[0]), float(out[1])


def ref_dgecon(lu, anorm, norm="1"):
    """Reference DGECON on the factors lu; returns (RCOND, INFO)."""
    lu = np.asfortranarray(lu, dtype=np.float64)
    n = lu.shape[0]
    work = np.zeros(4 * max(n, 1))
    iwork = np.zeros(max(n, 1), dtype=np.int32)
    rcond = ctypes.c_double(0)
    info = ctypes.c_int(0)
    r().dgecon_(ctypes.c_char_p(norm.encode()), _ri(n), ctypes.c_void_p(_p(lu)), _ri(max(1, n)),
                ctypes.byref(ctypes.c_double(anorm)), ctypes.byref(rcond), ctypes.c_void_p(_p(work)),
                ctypes.c_void_p(_p(iwork)), ctypes.byref(info), ctypes.c_size_t(1))
    return rcond.value, info.value


def ref_dgetri(lu, ipiv):
    """Reference DGETRI; returns (inverse, INFO)."""
    a = np.array(lu, dtype=np.float64, order="F", copy=True)
    n = a.shape[0]
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    lwork = max(1, 64 * n)
    work = np.zeros(lwork)
    info = ctypes.c_int(0)
    r().dgetri_(_ri(n), ctypes.c_void_p(_p(a)), _ri(max(1, n)), ctypes.c_void_p(_p(ip)), ctypes.c_void_p(_p(work)),
                _ri(lwork), ctypes.byref(info))
    return a, info.value


def ref_dgesvx(fact, trans, a, b, af=None, ipiv=None, equed="N", rr=None, cc=None):
    """Reference DGESVX; returns a dict of every output."""
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    n = a.shape[0]
    b = np.array(b, dtype=np.float64, order="F", copy=True).reshape(n, -1, order="F")
    nrhs = b.shape[1]
    af = np.zeros((n, n), order="F") if af is None else np.array(af, dtype=np.float64, order="F", copy=True)
    ip = np.zeros(max(n, 1), dtype=np.int32) if ipiv is None else np.ascontiguousarray(ipiv, dtype=np.int32).copy()
    rr = np.zeros(max(n, 1)) if rr is None else np.array(rr, dtype=np.float64, copy=True)
    cc = np.zeros(max(n, 1)) if cc is None else np.array(cc, dtype=np.float64, copy=True)
    x = np.zeros((n, nrhs), order="F")
    eq = ctypes.create_string_buffer(equed.encode(), 1)
    rcond = ctypes.c_double(0)
    ferr = np.zeros(max(nrhs, 1))
    berr = np.zeros(max(nrhs, 1))
    work = np.zeros(4 * max(n, 1))
    iwork = np.zeros(max(n, 1), dtype=np.int32)
    info = ctypes.c_int(0)
    r().dgesvx_(ctypes.c_char_p(fact.encode()), ctypes.c_char_p(trans.encode()), _ri(n), _ri(nrhs),
                ctypes.c_void_p(_p(a)), _ri(max(1, n)), ctypes.c_void_p(_p(af)), _ri(max(1, n)),
                ctypes.c_void_p(_p(ip)), eq, ctypes.c_void_p(_p(rr)), ctypes.c_void_p(_p(cc)),
                ctypes.c_void_p(_p(b)), _ri(max(1, n)), ctypes.c_void_p(_p(x)), _ri(max(1, n)), ctypes.byref(rcond),
                ctypes.c_void_p(_p(ferr)), ctypes.c_void_p(_p(berr)), ctypes.c_void_p(_p(work)),
                ctypes.c_void_p(_p(iwork)), ctypes.byref(info), ctypes.c_size_t(1), ctypes.c_size_t(1),
                ctypes.c_size_t(1))
    return {"a": a, "af": af, "ipiv": ip[:n], "equed": eq.raw.decode(), "r": rr[:n], "c": cc[:n], "b": b, "x": x,
            "rcond": rcond.value, "ferr": ferr[:nrhs], "berr": berr[:nrhs], "rpvgrw": work[0], "info": info.value}


def ref_sgetrf(a):
    lu = np.array(a, dtype=np.float32, order="F", copy=True)
    m, n = lu.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    info = ctypes.c_int(0)
    r().sgetrf_(_ri(m), _ri(n), ctypes.c_void_p(_p(lu)), _ri(max(1, m)), ctypes.c_void_p(_p(ipiv)),
                ctypes.byref(info))
    return lu, ipiv[:min(m, n)], info.value


def ref_dgetrs(lu, ipiv, b, trans="N"):
    x = np.array(b, dtype=np.float64, order="F", copy=True)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n = lu.shape[0]
    lu = np.asfortranarray(lu, dtype=np.float64)
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    info = ctypes.c_int(0)
    r().dgetrs_(ctypes.c_char_p(trans.encode()), _ri(n), _ri(x.shape[1]), ctypes.c_void_p(_p(lu)), _ri(max(1, n)),
                ctypes.c_void_p(_p(ip)), ctypes.c_void_p(_p(x)), _ri(max(1, n)), ctypes.byref(info),
                ctypes.c_size_t(1))
    return x, info.value


def ref_dsgesv(a, b):
    aa = np.array(a, dtype=np.float64, order="F", copy=True)
    bb = np.asfortranarray(b, dtype=np.float64)
    if bb.ndim == 1:
        bb = bb.reshape(-1, 1, order="F")
    n, nrhs = bb.shape
    x = np.zeros((n, nrhs), dtype=np.float64, order="F")
    ipiv = np.zeros(max(1, n), dtype=np.int32)
    work = np.zeros(max(1, n * nrhs), dtype=np.float64)
    swork = np.zeros(max(1, n * (n + nrhs)), dtype=np.float32)
    it = ctypes.c_int(0)
    info = ctypes.c_int(0)
    r().dsgesv_(_ri(n), _ri(nrhs), ctypes.c_void_p(_p(aa)), _ri(max(1, n)), ctypes.c_void_p(_p(ipiv)),
                ctypes.c_void_p(_p(bb)), _ri(max(1, n)), ctypes.c_void_p(_p(x)), _ri(max(1, n)),
                ctypes.c_void_p(_p(work)), ctypes.c_void_p(_p(swork)), ctypes.byref(it), ctypes.byref(info))
    return x, ipiv[:n], it.value, info.value, aa


def ref_dgerfs(a, lu, ipiv, b, x, trans="N"):
    a = np.asfortranarray(a, dtype=np.float64)
    lu = np.asfortranarray(lu, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64).reshape(a.shape[0], -1, order="F")
    xx = np.array(x, dtype=np.float64, order="F", copy=True).reshape(a.shape[0], -1, order="F")
    n, nrhs = b.shape
    ferr = np.zeros(max(1, nrhs))
    berr = np.zeros(max(1, nrhs))
    work = np.zeros(3 * max(1, n))
    iwork = np.zeros(max(1, n), dtype=np.int32)
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    info = ctypes.c_int(0)
    vp = ctypes.c_void_p
    r().dgerfs_(ctypes.c_char_p(trans.encode()), _ri(n), _ri(nrhs), vp(_p(a)), _ri(max(1, n)), vp(_p(lu)),
                _ri(max(1, n)), vp(_p(ip)), vp(_p(b)), _ri(max(1, n)), vp(_p(xx)), _ri(max(1, n)), vp(_p(ferr)),
                vp(_p(berr)), vp(_p(work)), vp(_p(iwork)), ctypes.byref(info), ctypes.c_size_t(1))
    return xx, ferr[:nrhs], berr[:nrhs], info.value


def ref_dlaswp(a, k1, k2, ipiv, incx=1):
    x = np.array(a, dtype=np.float64, order="F", copy=True)
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    r().dlaswp_(_ri(x.shape[1]), ctypes.c_void_p(_p(x)), _ri(max(1, x.shape[0])), _ri(k1), _ri(k2),
                ctypes.c_void_p(_p(ip)), _ri(incx))
    return x


def ref_dgemm(ta, tb, alpha, a, b, beta, c):
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    cc = np.array(c, dtype=np.float64, order="F", copy=True)
    m, n = cc.shape
    k = a.shape[1] if ta.upper() == "N" else a.shape[0]
    r().dgemm_(ctypes.c_char_p(ta.encode()), ctypes.c_char_p(tb.encode()), _ri(m), _ri(n), _ri(k),
               ctypes.byref(ctypes.c_double(alpha)), ctypes.c_void_p(_p(a)), _ri(max(1, a.shape[0])),
               ctypes.c_void_p(_p(b)), _ri(max(1, b.shape[0])), ctypes.byref(ctypes.c_double(beta)),
               ctypes.c_void_p(_p(cc)), _ri(max(1, m)), ctypes.c_size_t(1), ctypes.c_size_t(1))
    return cc


def ref_dtrsm(side, uplo, trans, diag, alpha, a, b):
    a = np.asfortranarray(a, dtype=np.float64)
    x = np.array(b, dtype=np.float64, order="F", copy=True)
    m, n = x.shape
    r().dtrsm_(ctypes.c_char_p(side.encode()), ctypes.c_char_p(uplo.encode()), ctypes.c_char_p(trans.encode()),
               ctypes.c_char_p(diag.encode()), _ri(m), _ri(n), ctypes.byref(ctypes.c_double(alpha)),
               ctypes.c_void_p(_p(a)), _ri(max(1, a.shape[0])), ctypes.c_void_p(_p(x)), _ri(max(1, m)),
               ctypes.c_size_t(1), ctypes.c_size_t(1), ctypes.c_size_t(1), ctypes.c_size_t(1))
    return x


def ref_idamax(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    r().idamax_.restype = ctypes.c_int
    return r().idamax_(_ri(len(x)), ctypes.c_void_p(_p(x)), _ri(1))


def ref_dlange(norm, a):
    """DLANGE of the reference build (SRC/dlange.f; 'F' through SRC/dlassq.f90)."""
    a = np.asfortranarray(a, dtype=np.float64)
    m, n = a.shape
    work = np.zeros(max(1, m))
    fn = r().dlange_
    fn.restype = ctypes.c_double
    return fn(ctypes.c_char_p(norm.encode()), _ri(m), _ri(n), ctypes.c_void_p(_p(a)), _ri(max(1, m)),
              ctypes.c_void_p(_p(work)), ctypes.c_size_t(1))


def ref_dgeequ(a, radix=False):
    a = np.asfortranarray(a, dtype=np.float64)
    m, n = a.shape
    rr = np.zeros(max(1, m))
    cc = np.zeros(max(1, n))
    rowcnd, colcnd, amax, info = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    fn = r().dgeequb_ if radix else r().dgeequ_
    fn(_ri(m), _ri(n), ctypes.c_void_p(_p(a)), _ri(max(1, m)), ctypes.c_void_p(_p(rr)), ctypes.c_void_p(_p(cc)),
       ctypes.byref(rowcnd), ctypes.byref(colcnd), ctypes.byref(amax), ctypes.byref(info))
    return rr[:m], cc[:n], rowcnd.value, colcnd.value, amax.value, info.value


def ref_dlaqge(a, rr, cc, rowcnd, colcnd, amax):
    x = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = x.shape
    eq = ctypes.create_string_buffer(2)
    r().dlaqge_(_ri(m), _ri(n), ctypes.c_void_p(_p(x)), _ri(max(1, m)),
                ctypes.c_void_p(_p(np.ascontiguousarray(rr))), ctypes.c_void_p(_p(np.ascontiguousarray(cc))),
                ctypes.byref(ctypes.c_double(rowcnd)), ctypes.byref(ctypes.c_double(colcnd)),
                ctypes.byref(ctypes.c_double(amax)), eq, ctypes.c_size_t(1))
    return x, eq.value.decode()


def ref_xerbla_reset():
    r().ref_xerbla_reset()


def ref_xerbla_last():
    buf = ctypes.create_string_buffer(32)
    r().ref_xerbla_last.restype = ctypes.c_int
    info = r().ref_xerbla_last(buf, 32)
    r().ref_xerbla_calls.restype = ctypes.c_int
    return buf.value.decode(), info, r().ref_xerbla_calls()


def ref_dgetrf_raw(m, n, a_ptr, lda, ipiv_ptr):
    """Call the reference DGETRF with raw arguments (error-exit tests)."""
    info = ctypes.c_int(0)
    r().dgetrf_(_ri(m), _ri(n), ctypes.c_void_p(a_ptr), _ri(lda), ctypes.c_void_p(ipiv_ptr), ctypes.byref(info))
    return info.value


def ref_dgetrs_raw(trans, n, nrhs, a_ptr, lda, ipiv_ptr, b_ptr, ldb):
    info = ctypes.c_int(0)
    r().dgetrs_(ctypes.c_char_p(trans.encode()), _ri(n), _ri(nrhs), ctypes.c_void_p(a_ptr), _ri(lda),
                ctypes.c_void_p(ipiv_ptr), ctypes.c_void_p(b_ptr), _ri(ldb), ctypes.byref(info), ctypes.c_size_t(1))
    return info.value
