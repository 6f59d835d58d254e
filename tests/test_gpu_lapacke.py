"""GPU checks of the LAPACKE layer (LAPACKE/src/lapacke_d*.c semantics) over
the library's drop-ins: row-major and column-major, host and device arrays,
against the oracle on the column-major data (IPIV identical, DGET01 / DGET02
below 30, the row-major results the transposes of the column-major ones)."""
import ctypes

import numpy as np
import pytest

import lapack_amd as la
import oracle

from gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu

ROW, COL = 101, 102


def L():
    lib = la.lib()
    vp, i, d, ch = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char
    sig = {
        "LAPACKE_dgetrf": (i, [i, i, i, vp, i, vp]),
        "LAPACKE_dgetrs": (i, [i, ch, i, i, vp, i, vp, vp, i]),
        "LAPACKE_dgesv": (i, [i, i, i, vp, i, vp, vp, i]),
        "LAPACKE_dgetri": (i, [i, i, vp, i, vp]),
        "LAPACKE_dsgesv": (i, [i, i, i, vp, i, vp, vp, i, vp, i, vp]),
        "LAPACKE_dgecon": (i, [i, ch, i, vp, i, d, vp]),
        "LAPACKE_dlange": (d, [i, ch, i, i, vp, i]),
        "LAPACKE_dlaswp": (i, [i, i, vp, i, i, i, vp, i]),
    }
    for k, (r, a) in sig.items():
        getattr(lib, k).restype = r
        getattr(lib, k).argtypes = a
    return lib


def stored(a, layout):
    """A copy of the logical matrix a in the buffer layout LAPACKE is given
    (row- or column-major); never a view of a, which the call overwrites."""
    return np.array(a, order="C" if layout == ROW else "F", copy=True)


def logical(buf, layout):
    return np.asarray(buf)


def ptr(x):
    return x.data_ptr() if hasattr(x, "data_ptr") else x.ctypes.data


def place(buf, device):
    """Host buffer, or a device tensor holding the same bytes."""
    if not device:
        return buf
    import torch
    flat = np.ravel(buf, order="K")
    return torch.from_numpy(flat.copy()).cuda()


def fetch(d, like, device):
    if not device:
        return like
    out = d.cpu().numpy()
    return np.lib.stride_tricks.as_strided(out, shape=like.shape, strides=like.strides).copy()


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("layout", [ROW, COL])
def test_lapacke_getrf_getrs_gesv(dev, layout, device):
    lib = L()
    n, nrhs = 300, 3
    a = la.matrix(2, 501, n, n)
    b = la.matrix(2, 502, n, nrhs)
    lu_r, p_r, _ = oracle.dgetrf(a)
    # DGETRF
    ab = stored(a, layout)
    da = place(ab, device)
    ip = np.zeros(n, dtype=np.int32)
    assert lib.LAPACKE_dgetrf(layout, n, n, ptr(da), n, ip.ctypes.data) == 0
    lu = fetch(da, ab, device)
    assert np.array_equal(ip, p_r)
    assert oracle.dget01(a, np.asfortranarray(lu), ip) < 30
    # DGETRS on those factors
    bb = stored(b, layout)
    db = place(bb, device)
    ld_b = nrhs if layout == ROW else n
    assert lib.LAPACKE_dgetrs(layout, b"N", n, nrhs, ptr(da), n, ip.ctypes.data, ptr(db), ld_b) == 0
    x = fetch(db, bb, device)
    assert oracle.dget02(a, np.asfortranarray(x), b) < 30
    # DGESV in one call
    ab2, bb2 = stored(a, layout), stored(b, layout)
    da2, db2 = place(ab2, device), place(bb2, device)
    ip2 = np.zeros(n, dtype=np.int32)
    assert lib.LAPACKE_dgesv(layout, n, nrhs, ptr(da2), n, ip2.ctypes.data, ptr(db2), ld_b) == 0
    assert np.array_equal(ip2, p_r)
    x2 = fetch(db2, bb2, device)
    assert np.array_equal(x2, x)


@pytest.mark.parametrize("layout", [ROW, COL])
def test_lapacke_getri_gecon_lange_laswp(dev, layout):
    lib = L()
    n = 120
    a = la.matrix(2, 511, n, n)
    lu_r, p_r, _ = oracle.dgetrf(a)
    # DLANGE: row-major swaps the one and infinity norms of the stored transpose
    ab = stored(a, layout)
    for norm in "M1I":
        assert lib.LAPACKE_dlange(layout, norm.encode(), n, n, ab.ctypes.data, n) == oracle.dlange(norm, a)
    # DGECON on the factors
    lub = stored(lu_r, layout)
    rc = ctypes.c_double(0)
    anorm = oracle.dlange("1", a)
    assert lib.LAPACKE_dgecon(layout, b"1", n, lub.ctypes.data, n, anorm, ctypes.byref(rc)) == 0
    if oracle.ref_available():
        assert rc.value == pytest.approx(oracle.ref_dgecon(lu_r, anorm, "1")[0], rel=1e-10)
    # DGETRI
    inv = stored(lu_r, layout)
    assert lib.LAPACKE_dgetri(layout, n, inv.ctypes.data, n, p_r.astype(np.int32).ctypes.data) == 0
    assert np.abs(np.asarray(inv) @ a - np.eye(n)).max() < 1e-10
    # DLASWP of the factorization's interchanges on A reproduces P*A
    pa = stored(a, layout)
    ipv = p_r.astype(np.int32)
    assert lib.LAPACKE_dlaswp(layout, n, pa.ctypes.data, n, 1, n, ipv.ctypes.data, 1) == 0
    assert np.array_equal(np.asarray(pa), oracle.dlaswp(a, 1, n, ipv, 1))


@pytest.mark.parametrize("layout", [ROW, COL])
def test_lapacke_dsgesv(dev, layout):
    lib = L()
    n, nrhs = 200, 2
    a = la.matrix(2, 521, n, n)
    b = la.matrix(2, 522, n, nrhs)
    ab, bb = stored(a, layout), stored(b, layout)
    x = stored(np.zeros((n, nrhs)), layout)
    ip = np.zeros(n, dtype=np.int32)
    it = ctypes.c_int(0)
    ld = nrhs if layout == ROW else n
    assert lib.LAPACKE_dsgesv(layout, n, nrhs, ab.ctypes.data, n, ip.ctypes.data, bb.ctypes.data, ld, x.ctypes.data,
                              ld, ctypes.byref(it)) == 0
    assert it.value >= 0
    assert oracle.dget08(a, np.asfortranarray(x), b) < np.sqrt(n)
