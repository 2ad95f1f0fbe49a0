"""The allocation-free entry point pmcd_dot_device_noalloc against
pmcd_dot_device (same kernels, so the same bits), through the C ABI on device
buffers: two calls in a row, the second on a smaller input, then an empty one.
The scratch array keeps the previous call's partials; the second call must not
see them."""
import ctypes

import numpy as np
import pytest

import parmoon2022_amd as pm
from hip_buffers import DeviceArray

pytestmark = pytest.mark.gpu


def _dot(name, x, y, n):
    out = ctypes.c_double(np.nan)
    fn = getattr(pm.lib(), name)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                   ctypes.c_void_p]
    rc = fn(n, x.ptr, y.ptr, ctypes.byref(out), None)
    if rc != 0:
        raise RuntimeError(pm.lib().pmcd_last_error().decode())
    return out.value


def test_dot_noalloc_equals_dot_device():
    rng = np.random.default_rng(5)
    sizes = [3_000_001, 1000, 0]  # 1024 partials, then 4, then 1 (empty)
    xs = [rng.standard_normal(max(n, 1)) for n in sizes]
    ys = [rng.standard_normal(max(n, 1)) for n in sizes]
    dx = [DeviceArray(a) for a in xs]
    dy = [DeviceArray(a) for a in ys]
    for n, x, y, hx, hy in zip(sizes, dx, dy, xs, ys):
        print("n", n, flush=True)
        ref = _dot("pmcd_dot_device", x, y, n)
        got = _dot("pmcd_dot_device_noalloc", x, y, n)
        assert got == ref, (n, got, ref)
        exact = float(np.dot(hx[:n], hy[:n])) if n else 0.0
        bound = 4 * max(n, 1) * np.finfo(float).eps * (float(np.dot(np.abs(hx[:n]), np.abs(hy[:n]))) if n else 0.0)
        assert abs(got - exact) <= bound + 1e-300
    # the large one again after the small ones
    assert _dot("pmcd_dot_device_noalloc", dx[0], dy[0], sizes[0]) == _dot("pmcd_dot_device", dx[0], dy[0],
                                                                          sizes[0])
    for a in dx + dy:
        a.close()
