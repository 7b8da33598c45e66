"""The allocation-free entry point (mkde_density_at_points_noalloc) against
the allocating one (mkde_density_at_points) through the C ABI: identical
densities (the same kernels in the same order), on two calls in a row on one
device, the second with fewer points, for fixed and per-point bandwidths and
both the pilot (self term included) and the leave-one-out form."""
import ctypes

import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory of the caller (the entry point writes into it), from the
    HIP runtime the library already uses."""

    def __init__(self, nbytes):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def fill(self, value, n):
        a = np.full(n, value, np.float32)
        assert self.rt.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0   # host to device

    def read(self, n):
        a = np.empty(n, np.float32)
        assert self.rt.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, a.nbytes, 2) == 0   # device to host
        return a

    def free(self):
        if self.ptr:
            self.rt.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def noalloc(ctx, loo, buf):
    ctx.density_at_points_noalloc(loo, buf.ptr.value)


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("loo", [False, True])
def test_noalloc_equals_allocating_two_calls(adaptive, loo):
    with M.Context(0) as ctx:
        sizes = (120_000, 9_000)   # the second call has fewer points
        buf = DeviceBuffer(4 * sizes[0])
        try:
            for i, n in enumerate(sizes):
                x, y, w = cases.thomas_points(max(1, n // 300), 300, 12.0, 1500.0, seed=20260700 + i)
                x, y, w = x[:n], y[:n], w[:n]
                ctx.points_upload(x, y, w)
                if adaptive:
                    ctx.set_bandwidths(np.random.default_rng(i).uniform(3.0, 20.0, len(x)).astype(np.float32))
                else:
                    ctx.set_bandwidth(9.5)
                ctx.set_ec_weights(np.random.default_rng(i + 7).uniform(1.0, 1.4, len(x)).astype(np.float32))
                ref = ctx.DensityAtPoints(loo=loo)
                buf.fill(-1.0, len(x))
                held = ctx.memory()
                noalloc(ctx, loo, buf)
                assert ctx.memory() == held   # no context buffer was allocated or grown
                got = buf.read(len(x))
                assert np.array_equal(got, ref), (i, n)
        finally:
            buf.free()


def test_noalloc_rejects_too_many_points():
    # more points than the scratch array holds: a clear error, no kernel run
    n = 9_000_000
    r = np.random.default_rng(3)
    with M.Context(0) as ctx:
        ctx.points_upload(r.uniform(0, 1e4, n).astype(np.float32), r.uniform(0, 1e4, n).astype(np.float32))
        ctx.set_bandwidth(5.0)
        buf = DeviceBuffer(4 * 16)
        try:
            with pytest.raises(M.MkdeError) as e:
                noalloc(ctx, False, buf)
            assert e.value.code == M.MKDE_E_RANGE
            assert "scratch" in str(e.value)
        finally:
            buf.free()
