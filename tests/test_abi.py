"""CPU checks of the C ABI library: it loads, exports every function that
include/tnlmi.h declares, and its host-side generators are deterministic and
shaped as documented.  No GPU work here."""
import ctypes
import os
import re

import numpy as np

import tnl_amd
from tnl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "tnlmi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tnlmi_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.DECLARED) | {n for n in names if hasattr(L, n)}


def test_version_and_error_string():
    assert "gfx950" in tnl_amd.version()
    assert isinstance(_lib.lib().tnlmi_last_error(), bytes)


def test_status_maps_to_reference_exceptions():
    assert issubclass(tnl_amd.InvalidArgument, ValueError)
    assert issubclass(tnl_amd.OverflowError_, OverflowError)
    assert issubclass(tnl_amd.NotImplementedError_, NotImplementedError)


def test_generators():
    off, col, val = tnl_amd.laplace2d(5, 4)
    assert off[-1] == 5 * 20 - 2 * 5 - 2 * 4 and len(col) == off[-1]
    assert np.all(np.diff(off) >= 3) and np.all(np.diff(off) <= 5)
    for r in range(20):
        c = col[off[r]:off[r + 1]]
        assert np.all(np.diff(c) > 0) and r in c
    off3, col3, val3 = tnl_amd.poisson3d(4, 3, 5)
    assert off3[-1] == 7 * 60 - 2 * (3 * 5 + 4 * 5 + 4 * 3)
    assert np.isclose(val3.sum(), 6 * 60 - (off3[-1] - 60))
    o1, c1, v1 = tnl_amd.zipf_csr(50000, seed=3)
    o2, c2, v2 = tnl_amd.zipf_csr(50000, seed=3)
    assert np.array_equal(o1, o2) and np.array_equal(c1, c2) and np.array_equal(v1, v2)
    assert o1[-1] == 16 * 50000
    sizes = np.diff(o1)
    assert sizes.min() >= 1 and sizes.max() <= 4096
    for r in range(0, 50000, 997):
        c = c1[o1[r]:o1[r + 1]]
        assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < 50000
    u = tnl_amd.uniform(100000, -1, 1, seed=0)
    assert -1 <= u.min() and u.max() < 1 and abs(u.mean()) < 0.01
    g = tnl_amd.normal(100000, seed=0)
    assert abs(g.mean()) < 0.02 and abs(g.std() - 1) < 0.02
