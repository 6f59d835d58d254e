"""CPU-side checks of the C ABI (no GPU needed): the library loads, exports
every entry point include/lapack_amd.h declares, and checks arguments with
the reference's parameter numbers and XERBLA contract (TESTING/LIN/derrge.f
:131-184), before any device work -- so these run without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import lapack_amd as la

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lapack_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip and
                   (n.endswith("_") or n.startswith("lapack_amd_") or n.startswith("LAPACKE_"))})


def test_library_exports_every_declared_symbol():
    lib = la.lib()
    names = declared_functions()
    assert "dgetrf_" in names and "lapack_amd_dgetrf_async" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_kernel_launch_stub_is_defined():
    """Each kernel the host code launches has its launch stub in the library: a kernel template whose host
    pass fails to instantiate leaves an undefined __device_stub__ symbol, which a shared library still links
    with and which fails only when that kernel is first launched."""
    out = os.popen(f"nm -C --undefined-only {la.LIB_PATH}").read()
    assert "__device_stub__" not in out, [l for l in out.splitlines() if "__device_stub__" in l]


def test_xerbla_is_weak_and_overridable():
    out = os.popen(f"nm -D {la.LIB_PATH} | grep ' xerbla_'").read()
    assert " W xerbla_" in out


def _fi(v):
    return ctypes.byref(ctypes.c_int(v))


def test_dgetrf_error_exits_fortran_abi():
    lib = la.lib()
    la.xerbla_record(True)
    a = np.zeros(4)
    ip = np.zeros(2, dtype=np.int32)
    info = ctypes.c_int(0)
    for (m, n, lda), expect in [((-1, 0, 1), 1), ((0, -1, 1), 2), ((2, 1, 1), 4)]:
        la.xerbla_record(True)
        lib.dgetrf_(_fi(m), _fi(n), ctypes.c_void_p(a.ctypes.data), _fi(lda), ctypes.c_void_p(ip.ctypes.data),
                    ctypes.byref(info))
        assert info.value == -expect
        assert la.xerbla_last() == ("DGETRF", expect, 1)
    la.xerbla_record(False)


def test_dgetrs_error_exits_fortran_abi():
    lib = la.lib()
    a = np.zeros(4)
    b = np.zeros(4)
    ip = np.zeros(2, dtype=np.int32)
    info = ctypes.c_int(0)
    for (t, n, nrhs, lda, ldb), expect in [(("/", 0, 0, 1, 1), 1), (("N", -1, 0, 1, 1), 2),
                                           (("N", 0, -1, 1, 1), 3), (("N", 2, 1, 1, 2), 5),
                                           (("N", 2, 1, 2, 1), 8)]:
        la.xerbla_record(True)
        lib.dgetrs_(ctypes.c_char_p(t.encode()), _fi(n), _fi(nrhs), ctypes.c_void_p(a.ctypes.data), _fi(lda),
                    ctypes.c_void_p(ip.ctypes.data), ctypes.c_void_p(b.ctypes.data), _fi(ldb), ctypes.byref(info),
                    ctypes.c_size_t(1))
        assert info.value == -expect
        assert la.xerbla_last()[:2] == ("DGETRS", expect)
    la.xerbla_record(False)


def test_value_api_error_codes_and_quick_returns():
    a = np.zeros(4)
    ip = np.zeros(2, dtype=np.int32)
    assert la.dgetrf(-1, 1, a, 1, ip) == -1
    assert la.dgetrf(1, -1, a, 1, ip) == -2
    assert la.dgetrf(3, 1, a, 2, ip) == -4
    assert la.dgetrf(0, 5, a, 1, ip) == 0       # quick return, no device needed
    assert la.dgetrs("X", 1, 1, a, 1, ip, a, 1) == -1
    assert la.dgetrs("N", 3, 1, a, 2, ip, a, 3) == -5
    assert la.dgetrs("N", 0, 1, a, 1, ip, a, 1) == 0
    assert la.dgesv(2, 1, a, 1, ip, a, 2) == -4
    assert la.dgesv(2, 1, a, 2, ip, a, 1) == -7
    assert la.dgemm("N", "N", 2, 2, 2, 1.0, a, 1, a, 2, 0.0, a, 2) == -8
    assert la.dgemm("Q", "N", 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a, 2) == -1


def test_no_device_is_reported_not_hidden():
    """Without a GPU the library refuses to compute (there is no CPU path)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    a = np.eye(3, order="F")
    ip = np.zeros(3, dtype=np.int32)
    assert la.dgetrf(3, 3, a, 3, ip) == la.LA_ERR_NODEVICE


def test_tuning_hook_roundtrip():
    la.set_tuning(128, 32)
    assert la.get_tuning() == (128, 32)
    la.set_tuning(0, 0)
    assert la.get_tuning() == (256, 64)


def test_lapacke_layout_and_nan_checks_cpu():
    """LAPACKE's checks that come before any device work
    (LAPACKE/src/lapacke_dgetrf.c:40-51, lapacke_dgesv.c, lapacke_dgetrf_work.c):
    a bad layout is -1, a NaN in a host input is minus its position, the
    row-major leading-dimension checks are the _work function's numbers."""
    lib = la.lib()
    vp, i = ctypes.c_void_p, ctypes.c_int
    lib.LAPACKE_dgetrf.restype = i
    lib.LAPACKE_dgetrf.argtypes = [i, i, i, vp, i, vp]
    lib.LAPACKE_dgesv.restype = i
    lib.LAPACKE_dgesv.argtypes = [i, i, i, vp, i, vp, vp, i]
    lib.LAPACKE_dgetrf_work.restype = i
    lib.LAPACKE_dgetrf_work.argtypes = [i, i, i, vp, i, vp]
    lib.LAPACKE_set_nancheck.argtypes = [i]
    a = np.ones((4, 4))
    b = np.ones((4, 1))
    ip = np.zeros(4, dtype=np.int32)
    assert lib.LAPACKE_dgetrf(7, 4, 4, a.ctypes.data, 4, ip.ctypes.data) == -1
    a[2, 1] = np.nan
    assert lib.LAPACKE_dgetrf(102, 4, 4, a.ctypes.data, 4, ip.ctypes.data) == -4
    assert lib.LAPACKE_dgetrf(101, 4, 4, a.ctypes.data, 4, ip.ctypes.data) == -4
    a[2, 1] = 1.0
    b[3, 0] = np.nan
    assert lib.LAPACKE_dgesv(101, 4, 1, a.ctypes.data, 4, ip.ctypes.data, b.ctypes.data, 1) == -7
    # row-major: lda below the row length is -5 of the _work function
    assert lib.LAPACKE_dgetrf_work(101, 4, 4, a.ctypes.data, 3, ip.ctypes.data) == -5
    # a NaN scan switched off reaches the routine's own checks: LDA -4 shifted to -5
    # (the default XERBLA stops the program, as SRC/xerbla.f does: record instead)
    lib.LAPACKE_set_nancheck(0)
    la.xerbla_record(True)
    try:
        assert lib.LAPACKE_dgesv(102, 4, 1, a.ctypes.data, 3, ip.ctypes.data, b.ctypes.data, 4) == -5
        assert la.xerbla_last()[:2] == ("DGESV", 4)
    finally:
        lib.LAPACKE_set_nancheck(1)
        la.xerbla_record(False)


def test_tuning_table_json(tmp_path):
    """A table written in the tuner's format is read back per bucket (the next
    power of two of min(m, n)); explicit tuning overrides it."""
    path = tmp_path / "tune.json"
    path.write_text('{"getrf": {"16384": {"nb": 192, "leaf": 32, "ms": 101.5}, '
                    '"1024": {"nb": 128, "leaf": 64, "ms": 1.2}}}\n')
    assert la.load_tuning(str(path)) == 2
    assert la.get_tuning_for(16384) == (192, 32)
    assert la.get_tuning_for(9000) == (192, 32)
    assert la.get_tuning_for(1000) == (128, 64)
    assert la.get_tuning_for(3000) == (256, 64)   # no entry for 4096: defaults
    la.set_tuning(64, 16)
    try:
        assert la.get_tuning_for(16384) == (64, 16)
    finally:
        la.set_tuning(0, 0)
    assert la.load_tuning(str(tmp_path / "missing.json")) == -1
    assert la.load_tuning(None) == 0
    assert la.get_tuning_for(16384) == (256, 64)
