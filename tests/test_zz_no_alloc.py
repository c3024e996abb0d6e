"""lam_dot_noalloc (no device allocation) against lam_dot, through the C ABI.

Two calls in a row, the second on a smaller input, each compared with lam_dot
on the same device buffers: the same kernels and launch shape, so the FP64
results are identical.  Sizes beyond the stated maximum are rejected.
"""
import ctypes

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def test_dot_noalloc_matches_dot_on_two_calls(ctx):
    from lamino_amd import _lib

    lib = ctx.lib
    rng = np.random.default_rng(81)
    for n in (3_000_017, 1_234):
        a = rng.standard_normal(n).astype(np.float32)
        b = rng.standard_normal(n).astype(np.float32)
        ad, bd = ctx.to_device(a), ctx.to_device(b)
        want, got = ctypes.c_double(), ctypes.c_double()
        _lib.check(lib.lam_dot(ctx.h, n, ad.ptr, bd.ptr, ctypes.byref(want)))
        _lib.check(lib.lam_dot_noalloc(ctx.h, n, ad.ptr, bd.ptr, ctypes.byref(got)))
        say(f"n {n}: lam_dot {want.value!r}, lam_dot_noalloc {got.value!r}")
        assert got.value == want.value
        assert abs(got.value - float(np.dot(a.astype(np.float64), b.astype(np.float64)))) <= 1e-9 * float(
            np.sum(np.abs(a.astype(np.float64) * b)))


def test_dot_noalloc_rejects_sizes_beyond_the_maximum(ctx):
    from lamino_amd import _lib

    lib = ctx.lib
    a = ctx.to_device(np.ones(8, np.float32))
    out = ctypes.c_double()
    with pytest.raises(_lib.LaminoError) as e:
        _lib.check(lib.lam_dot_noalloc(ctx.h, (1 << 40) + 1, a.ptr, a.ptr, ctypes.byref(out)))
    assert e.value.code == 1
    with pytest.raises(_lib.LaminoError):
        _lib.check(lib.lam_dot_noalloc(ctx.h, -1, a.ptr, a.ptr, ctypes.byref(out)))
