"""The Fortran caller (lapack_amd/fortran/example_dgesv.f90), built with the
image's flang against liblapack_amd.so: reference-style CALL DGETRF / DGETRS /
DSGESV resolved to the library's drop-in symbols, and the ISO_C_BINDING
module.  It checks its own residual ratios (< 30) and prints OK."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lapack_amd", "fortran",
                   "example_dgesv")


def test_fortran_example_runs(dev):
    if not os.path.exists(EXE):
        pytest.skip("Fortran example not built (needs flang: make -C lapack_amd/fortran)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "OK"
