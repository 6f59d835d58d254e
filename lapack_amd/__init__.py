"""lapack_amd -- LAPACK's dense LU solve path on one MI355X (gfx950).

Python mirror of the reference interface over the C ABI in
include/lapack_amd.h.  The functions take the reference's arguments, in the
reference's order, with the same meaning (SRC/dgetrf.f:105, SRC/dgetrs.f:118,
SRC/dgesv.f:148, SRC/dlaswp.f:112, BLAS/SRC/dgemm.f:213): arrays are
column-major buffers (numpy arrays, or torch tensors / raw integers holding a
device address), IPIV is 1-based, and INFO is returned.

There is no CPU fallback: if liblapack_amd.so is missing or no GPU is present
the calls fail loudly (ImportError at load, or INFO = LA_ERR_NODEVICE).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblapack_amd.so")

LA_ERR_HIP = -1001
LA_ERR_NOMEM = -1002
LA_ERR_SYNC = -1003
LA_ERR_NODEVICE = -1004
LA_ERR_LIMIT = -1005

PHASES = ("transpose", "panel", "laswp", "trsm", "gemm_panel", "gemm", "getrs")

_lib = None

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


def lib():
    """The loaded liblapack_amd.so (built by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the GPU library has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp = ctypes.c_void_p
    i = ctypes.c_int
    d = ctypes.c_double
    f = ctypes.c_float
    ch = ctypes.c_char
    sig = {
        "lapack_amd_dgetrf": (i, [i, i, vp, i, vp]),
        "lapack_amd_sgetrf": (i, [i, i, vp, i, vp]),
        "lapack_amd_dgetrs": (i, [ch, i, i, vp, i, vp, vp, i]),
        "lapack_amd_sgetrs": (i, [ch, i, i, vp, i, vp, vp, i]),
        "lapack_amd_dgesv": (i, [i, i, vp, i, vp, vp, i]),
        "lapack_amd_dlaswp": (i, [i, vp, i, i, i, vp, i]),
        "lapack_amd_dsgesv": (i, [i, i, vp, i, vp, vp, i, vp, i, _c_int_p]),
        "lapack_amd_dgesv_ir": (i, [i, i, vp, i, vp, vp, i, vp, i, ch, _c_int_p, _c_int_p, vp]),
        "lapack_amd_dgerfs": (i, [ch, i, i, vp, i, vp, i, vp, vp, i, vp, i, vp, vp]),
        "lapack_amd_dgeequ": (i, [i, i, vp, i, vp, vp, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
        "lapack_amd_dgeequb": (i, [i, i, vp, i, vp, vp, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
        "lapack_amd_dlaqge": (ch, [i, i, vp, i, vp, vp, d, d, d]),
        "lapack_amd_dgemm": (i, [ch, ch, i, i, i, d, vp, i, vp, i, d, vp, i]),
        "lapack_amd_dgetrf_async": (i, [i, i, vp, i, vp, vp, vp]),
        "lapack_amd_sgetrf_async": (i, [i, i, vp, i, vp, vp, vp]),
        "lapack_amd_dgetrs_async": (i, [ch, i, i, vp, i, vp, vp, i, vp]),
        "lapack_amd_dgetrs_noalloc_async": (i, [ch, i, i, vp, i, vp, vp, i, vp, vp]),
        "lapack_amd_dgetrs_noalloc_bytes": (ctypes.c_size_t, [i, i]),
        "lapack_amd_noalloc_scratch_bytes": (ctypes.c_size_t, []),
        "lapack_amd_sgetrs_async": (i, [ch, i, i, vp, i, vp, vp, i, vp]),
        "lapack_amd_dlaswp_async": (i, [i, vp, i, i, i, vp, i, vp]),
        "lapack_amd_dgemm_async": (i, [ch, ch, i, i, i, d, vp, i, vp, i, d, vp, i, vp]),
        "lapack_amd_sgemm_async": (i, [ch, ch, i, i, i, f, vp, i, vp, i, f, vp, i, vp]),
        "lapack_amd_dgesv_adjoint_async": (i, [i, i, vp, i, vp, vp, i, vp, i, vp, i, vp, i, i, vp]),
        "lapack_amd_batch_plan_create": (i, [ctypes.POINTER(vp), i, vp, ch, i]),
        "lapack_amd_batch_plan_destroy": (None, [vp]),
        "lapack_amd_batch_plan_config": (i, [vp, _c_int_p, _c_int_p, _c_dbl_p]),
        "lapack_amd_batch_plan_set_config": (i, [vp, i, i]),
        "lapack_amd_dgetrf_batched_async": (i, [vp, vp, vp, vp, vp, vp]),
        "lapack_amd_sgetrf_batched_async": (i, [vp, vp, vp, vp, vp, vp]),
        "lapack_amd_dgetrs_batched_async": (i, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "lapack_amd_sgetrs_batched_async": (i, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "lapack_amd_dgesv_adjoint_batched_async": (i, [vp] * 11 + [i, vp]),
        "lapack_amd_set_tuning": (None, [i, i]),
        "lapack_amd_get_tuning": (None, [_c_int_p, _c_int_p]),
        "lapack_amd_xerbla_record": (None, [i]),
        "lapack_amd_xerbla_last": (i, [ctypes.c_char_p, i]),
        "lapack_amd_xerbla_calls": (i, []),
        "lapack_amd_phase_enable": (None, [i]),
        "lapack_amd_phase_reset": (None, []),
        "lapack_amd_phase_read": (None, [_c_dbl_p, ctypes.POINTER(ctypes.c_long), _c_dbl_p]),
        "lapack_amd_dlarnv": (None, [i, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_longlong, vp]),
        "lapack_amd_version": (ctypes.c_char_p, []),
        "lapack_amd_device_status": (i, [i]),
        "lapack_amd_tune_getrf": (i, [i, ctypes.c_char_p, _c_int_p, _c_int_p, _c_dbl_p]),
        "lapack_amd_load_tuning": (i, [ctypes.c_char_p]),
        "lapack_amd_get_tuning_for": (None, [i, _c_int_p, _c_int_p]),
        "lapack_amd_dlange": (d, [ch, i, i, vp, i]),
        "lapack_amd_dgecon": (i, [ch, i, vp, i, d, _c_dbl_p]),
        "lapack_amd_dgetri": (i, [i, vp, i, vp]),
        "lapack_amd_dgesvx": (i, [ch, ch, i, i, vp, i, vp, i, vp, ctypes.c_char_p, vp, vp, vp, i, vp, i, _c_dbl_p,
                                  vp, vp, _c_dbl_p]),
        "lapack_amd_test_fail_waits": (None, [i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _ptr(x):
    """Address of a buffer: numpy array, torch tensor, or an integer address."""
    if x is None:
        return None
    if isinstance(x, (int, np.integer)):
        return int(x)
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError(f"cannot take the address of {type(x)}")


def _stream(s):
    if s is None:
        return None
    if isinstance(s, int):
        return s
    if hasattr(s, "cuda_stream"):
        return s.cuda_stream
    raise TypeError(f"not a stream: {type(s)}")


def _c(ch):
    return ch.encode() if isinstance(ch, str) else ch


# ---- reference-shaped entry points (synchronous, host or device buffers) ----

def dgetrf(m, n, a, lda, ipiv):
    """DGETRF(M, N, A, LDA, IPIV, INFO): returns INFO (SRC/dgetrf.f:105)."""
    return lib().lapack_amd_dgetrf(m, n, _ptr(a), lda, _ptr(ipiv))


def sgetrf(m, n, a, lda, ipiv):
    """SGETRF (SRC/sgetrf.f)."""
    return lib().lapack_amd_sgetrf(m, n, _ptr(a), lda, _ptr(ipiv))


def dgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb):
    """DGETRS(TRANS, N, NRHS, A, LDA, IPIV, B, LDB, INFO) (SRC/dgetrs.f:118)."""
    return lib().lapack_amd_dgetrs(_c(trans), n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb)


def sgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb):
    """SGETRS (SRC/sgetrs.f)."""
    return lib().lapack_amd_sgetrs(_c(trans), n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb)


def dgesv(n, nrhs, a, lda, ipiv, b, ldb):
    """DGESV(N, NRHS, A, LDA, IPIV, B, LDB, INFO) (SRC/dgesv.f:148)."""
    return lib().lapack_amd_dgesv(n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb)


def dsgesv(n, nrhs, a, lda, ipiv, b, ldb, x, ldx):
    """DSGESV (SRC/dsgesv.f:191); returns (INFO, ITER)."""
    it = ctypes.c_int(0)
    info = lib().lapack_amd_dsgesv(n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb, _ptr(x), ldx, ctypes.byref(it))
    return info, it.value


def dgesv_ir(n, nrhs, a, lda, ipiv, b, ldb, x, ldx, prec="S", berr=None):
    """Mixed-precision solve ('S' FP32, 'H' FP16, 'B' BF16 factorization) with FP64
    refinement and GMRES-IR; returns (INFO, ITER, GMRES_ITER)."""
    it, git = ctypes.c_int(0), ctypes.c_int(0)
    info = lib().lapack_amd_dgesv_ir(n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb, _ptr(x), ldx, _c(prec),
                                     ctypes.byref(it), ctypes.byref(git), _ptr(berr))
    return info, it.value, git.value


def dgerfs(trans, n, nrhs, a, lda, af, ldaf, ipiv, b, ldb, x, ldx, ferr, berr):
    """DGERFS (SRC/dgerfs.f:185); X refined in place, FERR/BERR filled; returns INFO."""
    return lib().lapack_amd_dgerfs(_c(trans), n, nrhs, _ptr(a), lda, _ptr(af), ldaf, _ptr(ipiv), _ptr(b), ldb,
                                   _ptr(x), ldx, _ptr(ferr), _ptr(berr))


def dgeequ(m, n, a, lda, r, c, radix=False):
    """DGEEQU (radix=False) / DGEEQUB (radix=True); returns (INFO, ROWCND, COLCND, AMAX)."""
    rc_, cc_, am = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
    fn = lib().lapack_amd_dgeequb if radix else lib().lapack_amd_dgeequ
    info = fn(m, n, _ptr(a), lda, _ptr(r), _ptr(c), ctypes.byref(rc_), ctypes.byref(cc_), ctypes.byref(am))
    return info, rc_.value, cc_.value, am.value


def dlaqge(m, n, a, lda, r, c, rowcnd, colcnd, amax):
    """DLAQGE; returns EQUED as a str."""
    return lib().lapack_amd_dlaqge(m, n, _ptr(a), lda, _ptr(r), _ptr(c), float(rowcnd), float(colcnd),
                                   float(amax)).decode()


def dlange(norm, m, n, a, lda):
    """DLANGE (SRC/dlange.f:114): the norm, computed on the GPU."""
    return lib().lapack_amd_dlange(_c(norm), m, n, _ptr(a), lda)


def dgecon(norm, n, a, lda, anorm):
    """DGECON (SRC/dgecon.f:130) on the factors in a; returns (INFO, RCOND)."""
    rc = ctypes.c_double(0)
    info = lib().lapack_amd_dgecon(_c(norm), n, _ptr(a), lda, float(anorm), ctypes.byref(rc))
    return info, rc.value


def dgetri(n, a, lda, ipiv):
    """DGETRI (SRC/dgetri.f:114): A's factors replaced by inv(A); returns INFO."""
    return lib().lapack_amd_dgetri(n, _ptr(a), lda, _ptr(ipiv))


def dgesvx(fact, trans, n, nrhs, a, lda, af, ldaf, ipiv, equed, r, c, b, ldb, x, ldx, ferr, berr):
    """DGESVX (SRC/dgesvx.f:349); returns (INFO, EQUED, RCOND, RPVGRW)."""
    eq = ctypes.create_string_buffer(_c(equed), 1)
    rcond, rpvgrw = ctypes.c_double(0), ctypes.c_double(0)
    info = lib().lapack_amd_dgesvx(_c(fact), _c(trans), n, nrhs, _ptr(a), lda, _ptr(af), ldaf, _ptr(ipiv), eq,
                                   _ptr(r), _ptr(c), _ptr(b), ldb, _ptr(x), ldx, ctypes.byref(rcond), _ptr(ferr),
                                   _ptr(berr), ctypes.byref(rpvgrw))
    return info, eq.raw.decode(), rcond.value, rpvgrw.value


def dlaswp(n, a, lda, k1, k2, ipiv, incx):
    """DLASWP(N, A, LDA, K1, K2, IPIV, INCX) (SRC/dlaswp.f:112)."""
    return lib().lapack_amd_dlaswp(n, _ptr(a), lda, k1, k2, _ptr(ipiv), incx)


def dgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """DGEMM (BLAS/SRC/dgemm.f:213); returns 0 or -(parameter number)."""
    return lib().lapack_amd_dgemm(_c(transa), _c(transb), m, n, k, float(alpha), _ptr(a), lda, _ptr(b), ldb,
                                  float(beta), _ptr(c), ldc)


# ---- device entry points (asynchronous on a stream) --------------------------

def dgetrf_async(m, n, d_a, lda, d_ipiv, d_info, stream=None):
    return lib().lapack_amd_dgetrf_async(m, n, _ptr(d_a), lda, _ptr(d_ipiv), _ptr(d_info), _stream(stream))


def sgetrf_async(m, n, d_a, lda, d_ipiv, d_info, stream=None):
    return lib().lapack_amd_sgetrf_async(m, n, _ptr(d_a), lda, _ptr(d_ipiv), _ptr(d_info), _stream(stream))


def dgetrs_async(trans, n, nrhs, d_a, lda, d_ipiv, d_b, ldb, stream=None):
    return lib().lapack_amd_dgetrs_async(_c(trans), n, nrhs, _ptr(d_a), lda, _ptr(d_ipiv), _ptr(d_b), ldb,
                                         _stream(stream))


def dgetrs_noalloc_async(trans, n, nrhs, d_a, lda, d_ipiv, d_b, ldb, d_info, stream=None):
    """DGETRS with its workspace in the library's static device array (no
    device allocation); d_info receives 0 or LA_ERR_SYNC."""
    return lib().lapack_amd_dgetrs_noalloc_async(_c(trans), n, nrhs, _ptr(d_a), lda, _ptr(d_ipiv), _ptr(d_b), ldb,
                                                 _ptr(d_info), _stream(stream))


def sgetrs_async(trans, n, nrhs, d_a, lda, d_ipiv, d_b, ldb, stream=None):
    return lib().lapack_amd_sgetrs_async(_c(trans), n, nrhs, _ptr(d_a), lda, _ptr(d_ipiv), _ptr(d_b), ldb,
                                         _stream(stream))


def dlaswp_async(n, d_a, lda, k1, k2, d_ipiv, incx, stream=None):
    return lib().lapack_amd_dlaswp_async(n, _ptr(d_a), lda, k1, k2, _ptr(d_ipiv), incx, _stream(stream))


def dgemm_async(transa, transb, m, n, k, alpha, d_a, lda, d_b, ldb, beta, d_c, ldc, stream=None):
    return lib().lapack_amd_dgemm_async(_c(transa), _c(transb), m, n, k, float(alpha), _ptr(d_a), lda, _ptr(d_b),
                                        ldb, float(beta), _ptr(d_c), ldc, _stream(stream))


def sgemm_async(transa, transb, m, n, k, alpha, d_a, lda, d_b, ldb, beta, d_c, ldc, stream=None):
    return lib().lapack_amd_sgemm_async(_c(transa), _c(transb), m, n, k, float(alpha), _ptr(d_a), lda, _ptr(d_b),
                                        ldb, float(beta), _ptr(d_c), ldc, _stream(stream))


def dgesv_adjoint_async(n, nrhs, d_lu, lda, d_ipiv, d_x, ldx, d_xbar, ldxbar, d_bbar, ldbbar, d_abar, ldabar,
                        accumulate=False, stream=None):
    """Bbar = A^-T Xbar (DGETRS('T') on the saved LU) and Abar = -Bbar X^T."""
    return lib().lapack_amd_dgesv_adjoint_async(n, nrhs, _ptr(d_lu), lda, _ptr(d_ipiv), _ptr(d_x), ldx, _ptr(d_xbar),
                                                ldxbar, _ptr(d_bbar), ldbbar, _ptr(d_abar), ldabar,
                                                1 if accumulate else 0, _stream(stream))


# ---- batched LU ----------------------------------------------------------------

class BatchPlan:
    """A batch of independent systems of uneven order (the library's plan):
    orders n_list, precision 'd' or 's'; autotune times the kernel
    configurations on synthetic matrices of these orders and keeps the
    fastest.  Launches take device arrays: pointers (int64 tensors holding
    device addresses), leading dimensions and right-hand-side counts (int32)."""

    def __init__(self, n_list, precision="d", autotune=True):
        self.precision = precision
        self.batch = len(n_list)
        arr = (ctypes.c_int * max(1, self.batch))(*[int(v) for v in n_list])
        h = ctypes.c_void_p()
        rc = lib().lapack_amd_batch_plan_create(ctypes.byref(h), self.batch, ctypes.cast(arr, ctypes.c_void_p),
                                                 _c(precision), 1 if autotune else 0)
        if rc != 0:
            raise RuntimeError(f"lapack_amd_batch_plan_create returned {rc}")
        self._h = h

    def config(self):
        nb, nt, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        lib().lapack_amd_batch_plan_config(self._h, ctypes.byref(nb), ctypes.byref(nt), ctypes.byref(ms))
        return nb.value, nt.value, ms.value

    def set_config(self, nb, nthreads):
        return lib().lapack_amd_batch_plan_set_config(self._h, nb, nthreads)

    def getrf(self, d_aptr, d_lda, d_ipivptr, d_info, stream=None):
        fn = lib().lapack_amd_dgetrf_batched_async if self.precision == "d" else lib().lapack_amd_sgetrf_batched_async
        return fn(self._h, _ptr(d_aptr), _ptr(d_lda), _ptr(d_ipivptr), _ptr(d_info), _stream(stream))

    def getrs(self, d_nrhs, d_aptr, d_lda, d_ipivptr, d_bptr, d_ldb, stream=None):
        fn = lib().lapack_amd_dgetrs_batched_async if self.precision == "d" else lib().lapack_amd_sgetrs_batched_async
        return fn(self._h, _ptr(d_nrhs), _ptr(d_aptr), _ptr(d_lda), _ptr(d_ipivptr), _ptr(d_bptr), _ptr(d_ldb),
                  _stream(stream))

    def adjoint(self, d_nrhs, d_luptr, d_lda, d_ipivptr, d_xptr, d_ldx, d_xbarptr, d_ldxbar, d_abarptr, d_ldabar,
                accumulate=False, stream=None):
        """Batched adjoint: Xbar_b <- A_b^-T Xbar_b, Abar_b <- -Xbar_b X_b^T (+ Abar_b)."""
        return lib().lapack_amd_dgesv_adjoint_batched_async(
            self._h, _ptr(d_nrhs), _ptr(d_luptr), _ptr(d_lda), _ptr(d_ipivptr), _ptr(d_xptr), _ptr(d_ldx),
            _ptr(d_xbarptr), _ptr(d_ldxbar), _ptr(d_abarptr), _ptr(d_ldabar), 1 if accumulate else 0,
            _stream(stream))

    def close(self):
        if getattr(self, "_h", None):
            lib().lapack_amd_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- tuning, error hook, profiling, generator -------------------------------

def set_tuning(nb=0, leaf=0):
    """Outer block size NB and panel leaf width (0 = default); XLAENV analogue."""
    lib().lapack_amd_set_tuning(int(nb), int(leaf))


def tune_getrf(n, path=None):
    """Autotune NB and the leaf width for order n; returns (nb, leaf, ms)."""
    nb, leaf, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    rc = lib().lapack_amd_tune_getrf(int(n), path.encode() if path else None, ctypes.byref(nb), ctypes.byref(leaf),
                                     ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"lapack_amd_tune_getrf returned {rc}")
    return nb.value, leaf.value, ms.value


def load_tuning(path):
    """Read a tuning table (None empties it); returns the buckets read."""
    return lib().lapack_amd_load_tuning(path.encode() if path else None)


def get_tuning_for(n):
    """(NB, leaf) a factorization of order n would use."""
    nb, leaf = ctypes.c_int(), ctypes.c_int()
    lib().lapack_amd_get_tuning_for(int(n), ctypes.byref(nb), ctypes.byref(leaf))
    return nb.value, leaf.value


def get_tuning():
    nb, leaf = ctypes.c_int(), ctypes.c_int()
    lib().lapack_amd_get_tuning(ctypes.byref(nb), ctypes.byref(leaf))
    return nb.value, leaf.value


def xerbla_record(on=True):
    lib().lapack_amd_xerbla_record(1 if on else 0)


def xerbla_last():
    """(routine name, parameter number, number of calls) of the last XERBLA call."""
    buf = ctypes.create_string_buffer(32)
    info = lib().lapack_amd_xerbla_last(buf, 32)
    return buf.value.decode(), info, lib().lapack_amd_xerbla_calls()


def phase_enable(on=True, level=1):
    """level 1: every phase; level 2: only the trailing update (cheap enough
    to stay on in a timed run)."""
    lib().lapack_amd_phase_enable(int(level) if on else 0)


def phase_reset():
    lib().lapack_amd_phase_reset()


def phase_read():
    """{phase: (ms, calls, flops)} accumulated since the last reset."""
    n = len(PHASES)
    ms = (ctypes.c_double * n)()
    calls = (ctypes.c_long * n)()
    fl = (ctypes.c_double * n)()
    lib().lapack_amd_phase_read(ms, calls, fl)
    return {p: (ms[i], calls[i], fl[i]) for i, p in enumerate(PHASES)}


def dlarnv(idist, seed, n, offset=0, out=None):
    """Counter-based seeded generator: idist 1 U(0,1), 2 U(-1,1), 3 N(0,1)."""
    if out is None:
        out = np.empty(int(n), dtype=np.float64)
    lib().lapack_amd_dlarnv(int(idist), int(seed), int(offset), int(n), out.ctypes.data)
    return out


def matrix(idist, seed, m, n, offset=0):
    """m x n Fortran-ordered matrix from dlarnv, element (i, j) at index offset + i + j*m."""
    return dlarnv(idist, seed, m * n, offset).reshape((n, m)).T


def version():
    return lib().lapack_amd_version().decode()


def device_status(clear=True):
    """Sticky status of the current device: bit 0 = a bounded wait between
    workgroups gave up since the last clear (waits for the device first)."""
    return lib().lapack_amd_device_status(1 if clear else 0)


def test_fail_waits(count=1):
    """Test hook: the next `count` factorizations / solves start with their
    wait-error word set, so every bounded wait in them gives up at once."""
    lib().lapack_amd_test_fail_waits(int(count))
