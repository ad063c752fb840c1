"""Builds every HIP source of the library to gfx950 assembly and checks that no
kernel writes through the scalar data cache: no scalar memory stores or
atomics, and no write-back or discard of that cache (the library writes only
through vector memory instructions).  CPU-only: it reads the compiler's
assembly of this library's own sources."""
import glob
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tnl_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FORBIDDEN = re.compile(r"^\s*(s_store_|s_buffer_store|s_scratch_store|s_atomic_|s_buffer_atomic|s_dcache_wb|"
                       r"s_dcache_discard)", re.M)


def assembly(src, out_dir):
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "--cuda-device-only", "-S",
           "-o", out]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    subprocess.run(cmd + [src], check=True, capture_output=True)
    with open(out) as f:
        return f.read()


def test_scanner_pattern():
    assert FORBIDDEN.search("\ts_store_dword s1, s[2:3], 0x0\n")
    assert FORBIDDEN.search("\ts_dcache_wb\n")
    assert not FORBIDDEN.search("\tglobal_store_dwordx2 v[0:1], v[2:3], off\n\ts_load_dword s0, s[4:5], 0x0\n")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_library_has_no_scalar_cache_writes(tmp_path):
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    assert sources
    with ThreadPoolExecutor(max_workers=8) as pool:
        texts = list(pool.map(lambda s: assembly(s, str(tmp_path)), sources))
    bad = {os.path.basename(s): sorted(set(m.group(1) for m in FORBIDDEN.finditer(t)))
           for s, t in zip(sources, texts) if FORBIDDEN.search(t)}
    assert bad == {}
