"""tc_check_divergence_noalloc: the no-allocation variant of
tc_check_divergence (its reduction partials in the library's fixed device
scratch, the result written to device memory) returns exactly what
tc_check_divergence returns, on two calls in a row, the second on a smaller
grid, and rejects a grid beyond the scratch."""
import ctypes

import numpy as np
import pytest

from helpers import small_params

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime the library itself is linked against
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return h


def _solver(tc, p, seed):
    s = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, seed, 1e-2)
    s.set("qt", qt); s.set("qr", qr); s.set("qz", qz)
    return s


def test_noalloc_divergence_equals_the_allocating_entry_point(tc):
    hip = _hip()
    L = tc.lib()
    buf = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(buf), 8) == 0
    try:
        for shape, seed in (((96, 64, 49), 3), ((40, 24, 17), 4)):
            p = small_params(tc, *shape)
            s = _solver(tc, p, seed)
            assert hip.hipMemset(buf, 0xFF, 8) == 0  # a NaN pattern that must be overwritten
            assert L.tc_check_divergence_noalloc(s.h, buf) == 0, L.tc_last_error()
            s.synchronize()
            got = ctypes.c_double()
            assert hip.hipMemcpy(ctypes.addressof(got), buf, 8, 2) == 0  # device to host
            want = s.CheckDivergence()
            print(shape, got.value, want, flush=True)
            assert want > 1e-3
            assert got.value == want, (shape, got.value, want)
            s.close()
    finally:
        hip.hipFree(buf)


def test_noalloc_rejects_bad_arguments(tc):
    p = small_params(tc, 40, 24, 17)
    s = _solver(tc, p, 5)
    with pytest.raises(tc.TCError):
        tc._check(tc.lib().tc_check_divergence_noalloc(s.h, None))
    s.close()
