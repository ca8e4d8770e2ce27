"""The warm-start network's loss and gradient without allocation
(upside_mlp_loss_grad_noalloc, intermediates in the library's static device
array) against the allocating entry point (upside_mlp_loss_grad) on the same
parameters and rows: the same kernels, grid and order, so the same bits.  Two
calls in a row, the second with fewer rows, so the array's contents from the
first call cannot leak into the second.  Device buffers are allocated by the
test through the HIP runtime the library uses."""
import ctypes as ct

import numpy as np
import pytest

from conftest import say
from test_warmstart import random_params, synthetic_rows

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class DeviceArrays:
    """hipMalloc'd float32 buffers, freed with the object."""

    def __init__(self):
        self.hip = ct.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.hip.hipFree.argtypes = [ct.c_void_p]
        self.hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = ct.c_void_p()
        assert self.hip.hipMalloc(ct.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def __del__(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


@pytest.fixture(scope="module")
def ue():
    import upside_engine
    lib = upside_engine.load_upside_library()
    lib.upside_mlp_loss_grad_noalloc.restype = ct.c_int
    lib.upside_mlp_loss_grad_noalloc.argtypes = [ct.c_void_p] * 3 + [ct.c_int, ct.c_void_p]
    return upside_engine


def test_noalloc_loss_grad_matches_allocating_entry_point(ue):
    lib = ue.load_upside_library()
    dev = DeviceArrays()
    net = ue.WarmStartNet(seed=2)
    p = random_params(31, scale_out=1.0)
    net.set_params(p)
    dp = dev.put(p)
    for n in (5000, 700):
        say("case: no-allocation loss and gradient, %d rows" % n)
        f, t = synthetic_rows(n, 40 + n)
        net.set_data(f, t)
        loss, g = net.loss_grad()
        df, dt = dev.put(f), dev.put(t)
        out = dev.put(np.full(len(p) + 1, np.nan, np.float32))
        assert lib.upside_mlp_loss_grad_noalloc(dp, df, dt, n, out) == 0
        got = dev.get(out, len(p) + 1)
        say("loss %.7f vs %.7f; gradient max |diff| %.3e" % (got[-1], loss, np.abs(got[:-1] - g).max()))
        assert np.array_equal(got[:-1], g)
        assert got[-1] == np.float32(loss)


def test_noalloc_rejects_sizes_beyond_the_maximum(ue):
    lib = ue.load_upside_library()
    dev = DeviceArrays()
    ptr = dev.put(np.zeros(64, np.float32))
    assert lib.upside_mlp_loss_grad_noalloc(ptr, ptr, ptr, 0, ptr) != 0
    assert lib.upside_mlp_loss_grad_noalloc(ptr, ptr, ptr, (1 << 24) + 1, ptr) != 0
    assert lib.upside_mlp_loss_grad_noalloc(None, ptr, ptr, 16, ptr) != 0
