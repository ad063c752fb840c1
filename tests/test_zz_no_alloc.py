"""tnlmi_dot_noalloc (the allocation-free (a, b)) against tnlmi_dot, through
the C ABI: bit-identical results on two calls in a row, the second on a
smaller input, and an empty input.  The entry point keeps its partial sums in
the library's static device array, so it must also work as the first call of
a fresh process."""
import ctypes

import pytest
import torch

from tnl_amd import _lib

pytestmark = pytest.mark.gpu


def _dot(name, a, b):
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(_lib.lib(), name)(a.numel(), ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), stream)
    assert rc == 0, _lib.lib().tnlmi_last_error()
    torch.cuda.synchronize()
    return out.item()


def test_dot_noalloc_matches_dot_on_two_calls_in_a_row():
    g = torch.Generator(device="cpu").manual_seed(5)
    big = 3_000_001
    a = torch.randn(big, generator=g, dtype=torch.float64).cuda()
    b = torch.randn(big, generator=g, dtype=torch.float64).cuda()
    # first call of the process on this device: the static array's address is looked up here
    first = _dot("tnlmi_dot_noalloc", a, b)
    assert first == _dot("tnlmi_dot", a, b)
    exact = float((a.cpu() * b.cpu()).sum())
    assert abs(first - exact) <= 1e-12 * float((a.cpu().abs() * b.cpu().abs()).sum())
    # second call, smaller input: fewer first-pass partials than the call before left in the array
    n2 = 777
    a2, b2 = a[:n2].clone(), b[:n2].clone()
    second = _dot("tnlmi_dot_noalloc", a2, b2)
    assert second == _dot("tnlmi_dot", a2, b2)
    assert abs(second - float((a2.cpu() * b2.cpu()).sum())) <= 1e-13 * float((a2.cpu().abs() * b2.cpu().abs()).sum())
    # an empty input gives 0 (one partial, written by the first pass)
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    assert _dot("tnlmi_dot_noalloc", e, e) == 0.0


def test_dot_noalloc_rejects_bad_arguments():
    a = torch.ones(4, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = _lib.lib().tnlmi_dot_noalloc(-1, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), None)
    assert rc == _lib.INVALID_ARGUMENT
    assert b"dot_noalloc" in _lib.lib().tnlmi_last_error()
