"""Code-generation guard for the gfx950 kernels (CPU only).

Every .hip source is compiled to device assembly and scanned: the kernels
must write memory only through vector instructions -- no scalar-cache
stores, scalar atomics or scalar-cache write-back/discard -- and must not
spill to scratch in the hot kernels.  Runs where hipcc is present.
"""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lapack_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "--offload-device-only", "-S", "-o", "-"]
FORBIDDEN = re.compile(r"^\s*(s_store_|s_buffer_store|s_scratch_store|s_atomic_|s_buffer_atomic|s_dcache_wb|"
                       r"s_dcache_discard)", re.M)


def _asm(src):
    r = subprocess.run([HIPCC] + FLAGS + [os.path.join(CSRC, src)], capture_output=True, text=True, cwd=CSRC)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_no_scalar_memory_writes():
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        asms = dict(zip(srcs, ex.map(_asm, srcs)))
    bad = {s: FORBIDDEN.findall(a) for s, a in asms.items() if FORBIDDEN.search(a)}
    assert not bad, bad
    # scratch (register spills) would show as a non-zero private segment in the hot kernels
    for s in ("gemm.hip", "panel.hip", "batched.hip", "getrs.hip", "gemm_h.hip"):
        sizes = [int(v) for v in re.findall(r"\.private_segment_fixed_size:\s+(\d+)", asms[s])]
        assert sizes and max(sizes) == 0, (s, sizes)
