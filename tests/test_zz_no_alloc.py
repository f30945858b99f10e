"""The fixed-point sum that allocates nothing (vlj_fixed_sum_device), through
the C ABI, against vlj_fixed_sum on the same values: exact integers, on two
calls in a row (the second smaller, so stale partial sums of the first
would show), and on an empty input.

Device buffers come from the HIP runtime the library itself is linked to
(libamdhip64), so the pointers are the library's own kind."""
import ctypes as C

import numpy as np
import pytest

from helpers import spread_values

pytestmark = pytest.mark.gpu


def hip():
    h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def device_sum(lj, h, values, scale):
    from virial_particles import _lib

    n = values.shape[0]
    d = C.c_void_p()
    o = C.c_void_p()
    assert h.hipMalloc(C.byref(d), max(n, 1) * 8) == 0
    assert h.hipMalloc(C.byref(o), 8) == 0
    try:
        if n:
            assert h.hipMemcpy(d, values.ctypes.data, n * 8, H2D) == 0
        rc = _lib.lib().vlj_fixed_sum_device(d, n, scale, o, None)
        assert rc == 0, _lib.lib().vlj_last_error(None)
        assert h.hipDeviceSynchronize() == 0
        out = np.zeros(1, dtype=np.int64)
        assert h.hipMemcpy(out.ctypes.data, o, 8, D2H) == 0
        return int(out[0])
    finally:
        h.hipFree(d)
        h.hipFree(o)


def test_no_alloc_sum_equals_the_allocating_one(lj):
    h = hip()
    big = spread_values(1 << 20).astype(np.float64)
    rng = np.random.default_rng(3)
    small = rng.uniform(-1e3, 1e3, 1000)
    for values, scale in ((big, 2.0 ** 32), (small, 2.0 ** 20), (big[: 300_000], 2.0 ** 32)):
        want = lj.fixed_sum(values, scale)
        got = device_sum(lj, h, values, scale)
        assert got == want
    assert device_sum(lj, h, np.zeros(0), 2.0 ** 32) == 0


def test_no_alloc_sum_rejects_oversized_input(lj):
    from virial_particles import _lib

    o = C.c_void_p(16)  # never touched: the size check comes first
    assert _lib.lib().vlj_fixed_sum_device(C.c_void_p(16), (1 << 36) + 1, 1.0, o, None) == _lib.E_INVALID
    assert b"2^36" in _lib.lib().vlj_last_error(None)
