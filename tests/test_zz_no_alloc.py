"""The allocation-free entry point lmw_corrector_forward_noalloc (device arrays in and out, activations in the
library's static per-GPU workspace) against lmw_corrector_forward (host arrays, workspace allocated per size), through
the C ABI: the same kernels on the same inputs give the same bits, on two calls in a row of which the second is
smaller (the workspace keeps the first call's contents), and sizes beyond the workspace are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class DeviceArrays:
    """Device buffers for the test's inputs and outputs, from the HIP runtime the library itself loaded."""

    def __init__(self):
        from lmweather_amd import _lib as L
        L.lib()  # the library first: its HIP runtime is the one found below
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.held.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held = []


def test_noalloc_forward_equals_forward():
    from lmweather_amd import _lib as L
    from lmweather_amd.corrector import Corrector
    dev = DeviceArrays()
    try:
        net = Corrector(42, 64, seed=11)
        rng = np.random.default_rng(4)
        params = (0.2 * rng.standard_normal(net.nparams)).astype(np.float32)
        L.call("lmw_corrector_set_params", net._h, L.fptr(params), params.size)
        for S in (9, 4):  # the second call is smaller: the workspace still holds the first call's activations
            x = rng.standard_normal((S, 3, net.C, 2)).astype(np.float32)
            want, want_res = net.forward(x)
            xd = dev.put(x)
            out = dev.alloc(x.nbytes)
            res = dev.alloc(8 * S)
            L.call("lmw_corrector_forward_noalloc", net._h, xd, S, out, res)
            got = dev.get(out, x.shape, np.float32)
            got_res = dev.get(res, (S,), np.float64)
            print(f"{S} samples: max |correction| {np.max(np.abs(want)):.3e}, residuals {got_res[:3]}")
            assert np.array_equal(got, want)
            assert np.array_equal(got_res, want_res)
            assert np.any(want != 0)
            # without the residual output
            L.call("lmw_corrector_forward_noalloc", net._h, xd, S, out, None)
            assert np.array_equal(dev.get(out, x.shape, np.float32), want)
        # 42 samples of T42 at width 64 need more than the 32 MiB workspace: refused, nothing launched
        big = dev.alloc(42 * 3 * net.C * 2 * 4)
        with pytest.raises(ValueError, match=r"42 samples need 34126336 bytes of workspace, more than the 33554432 bytes it has"):
            L.call("lmw_corrector_forward_noalloc", net._h, big, 42, big, None)
        # host arrays are refused rather than handed to a kernel
        x = np.zeros((1, 3, net.C, 2), dtype=np.float32)
        with pytest.raises(ValueError, match="must be device memory"):
            L.call("lmw_corrector_forward_noalloc", net._h, C.c_void_p(x.ctypes.data), 1, C.c_void_p(x.ctypes.data), None)
    finally:
        dev.free()
