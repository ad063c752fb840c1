"""Builds the persistent batched CG kernels to gfx950 assembly and checks that
no loop with a per-lane (exec-mask) exit holds a workgroup barrier.  Such a
loop hung the persistent kernel: thread 0's fetch of the next system was split
into an outer loop while the rest of its wavefront re-read a stale slot."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tnl_amd", "csrc")
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import barrier_loops  # noqa: E402

HIPCC = "/opt/rocm/bin/hipcc"

BAD = """\
_Zbad:
.LBB0_1:                                ; =>This Loop Header: Depth=1
	s_barrier
	ds_read_b32 v2, v6 offset:32
	s_andn2_b64 exec, exec, s[4:5]
	s_cbranch_execnz .LBB0_1
	s_endpgm
"""
GOOD = """\
_Zgood:
.LBB0_1:                                ; =>This Loop Header: Depth=1
	s_barrier
	v_readfirstlane_b32 s0, v2
	s_cmp_lt_i32 s0, s1
	s_cbranch_scc1 .LBB0_1
.LBB0_2:                                ; =>This Inner Loop Header: Depth=1
	global_load_dword v3, v[4:5], off
	s_andn2_b64 exec, exec, s[4:5]
	s_cbranch_execnz .LBB0_2
	s_endpgm
"""


def test_scanner_flags_divergent_barrier_loop():
    assert barrier_loops.scan(BAD) == [("_Zbad", 2, 5)]
    assert barrier_loops.scan(GOOD) == []


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src", ["batched.hip", "blas1.hip", "ir.hip", "fused_cg.hip", "spai.hip", "krylov.hip",
                                 "spmv_window.hip"])
def test_kernels_have_uniform_barrier_loops(tmp_path, src):
    out = tmp_path / (src + ".s")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "--cuda-device-only", "-S",
           "-o", str(out), os.path.join(CSRC, src)]
    subprocess.run(cmd, check=True, capture_output=True)
    assert barrier_loops.scan(out.read_text()) == []
