"""Allocation-free inner product (monolis_hip_inner_product_R_noalloc): an
entry point with no counterpart in monolis, for callers that cannot allocate
device memory.  Checked through the C-ABI against monolis_hip_inner_product_R
(the same kernels over the same partial count, so the sums agree bit for bit)
on consecutive calls with a shrinking input."""
import ctypes

import numpy as np
import pytest

from monolis_amd import _lib, api

pytestmark = pytest.mark.gpu


def _noalloc(L, m, dx, dy, out):
    return L.monolis_hip_inner_product_R_noalloc(int(m), ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dy.ptr), out)


def test_noalloc_dot_matches_allocating_entry_point():
    L = _lib.load()
    _lib.check(L.monolis_hip_init(), "monolis_hip_init")
    rng = np.random.default_rng(5)
    sizes = [3_000_001, 777_777, 1]     # consecutive calls, each smaller than the last
    for m in sizes:
        x = rng.uniform(-1, 1, m)
        y = rng.uniform(-1, 1, m)
        dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer.from_numpy(y)
        ref = ctypes.c_double()
        _lib.check(L.monolis_hip_inner_product_R(int(m), ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dy.ptr),
                                                 ctypes.byref(ref)), "inner_product_R")
        got = ctypes.c_double()
        _lib.check(_noalloc(L, m, dx, dy, ctypes.byref(got)), "inner_product_R_noalloc (host sum)")
        assert got.value == ref.value, (m, got.value, ref.value)
        dsum = api.DeviceBuffer(1)
        _lib.check(_noalloc(L, m, dx, dy, ctypes.c_void_p(dsum.ptr)), "inner_product_R_noalloc (device sum)")
        assert dsum.numpy()[0] == ref.value
        exact = float(np.dot(x, y))
        assert abs(got.value - exact) <= 1e-12 * max(1.0, float(np.dot(np.abs(x), np.abs(y))))


def test_noalloc_dot_rejects_host_arrays():
    L = _lib.load()
    _lib.check(L.monolis_hip_init(), "monolis_hip_init")
    x = np.ones(10)
    out = ctypes.c_double()
    rc = L.monolis_hip_inner_product_R_noalloc(10, x.ctypes.data_as(ctypes.c_void_p),
                                               x.ctypes.data_as(ctypes.c_void_p), ctypes.byref(out))
    assert rc != 0
    assert b"device arrays" in L.monolis_hip_last_error()
