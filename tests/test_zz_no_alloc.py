"""Allocation-free inference (clr_mlp_infer_static) against clr_mlp_infer,
through the C ABI: two calls in a row, the second smaller (the static
staging array keeps the first call's contents), each bitwise equal to the
allocating entry point (both run the same fused kernel); the size limit is
refused with CLR_E_RANGE."""
import ctypes as C

import numpy as np
import pytest

from tardiis_and_claritas_amd import _lib, abi

pytestmark = pytest.mark.gpu
PF = C.POINTER(C.c_float)


def _f(a):
    return a.ctypes.data_as(PF)


def test_infer_static_matches_infer(gpu_available):
    print('allocation-free inference vs clr_mlp_infer', flush=True)
    ctx = _lib.Context(0)
    lib, h = ctx.lib, ctx.h
    shape = abi.MlpShape(5, 8, 256, 3, 3, 0)
    ctx.check(lib.clr_mlp_init(h, C.byref(shape), 11), 'clr_mlp_init')
    rng = np.random.default_rng(4)
    for n in (60_000, 777):
        x = rng.normal(size=(n, 8)).astype(np.float32)
        m1, s1, m2, s2 = (np.full((n, 3), np.nan, np.float32) for _ in range(4))
        ctx.check(lib.clr_mlp_infer(h, _f(x), n, _f(m1), _f(s1)), 'clr_mlp_infer')
        ctx.check(lib.clr_mlp_infer_static(h, _f(x), n, _f(m2), _f(s2)), 'clr_mlp_infer_static')
        assert np.all(np.isfinite(m1)) and np.all(s1 >= 0)
        np.testing.assert_array_equal(m2, m1)
        np.testing.assert_array_equal(s2, s1)
    # stdev is optional
    m3 = np.zeros((n, 3), np.float32)
    ctx.check(lib.clr_mlp_infer_static(h, _f(x), n, _f(m3), None), 'clr_mlp_infer_static')
    np.testing.assert_array_equal(m3, m1)
    # beyond the 4 MiB staging array: refused, nothing run
    big = 80_000
    xb = np.zeros((big, 8), np.float32)
    mb = np.zeros((big, 3), np.float32)
    assert lib.clr_mlp_infer_static(h, _f(xb), big, _f(mb), None) == abi.E_RANGE
    assert b'queries per call' in lib.clr_last_error(h)
    ctx.close()
