"""fip_closure_forward_noalloc (the network on caller device buffers, activations in the
library's static device array) against fip_closure_forward: bit for bit, on two calls in a
row, the second with a smaller batch; batches beyond the array's size are rejected."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

MAX_PIXELS = 512 * 512


class DeviceBuffer:
    """A device allocation of the test (through the HIP runtime), with copies to and from
    host arrays."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def test_forward_noalloc_equals_forward():
    from fip_amd.sim import Closure
    rng = np.random.default_rng(41)
    cl = Closure(64, 64, 0, seed=6)
    feat_buf = DeviceBuffer(MAX_PIXELS * 6 * 4)
    out_buf = DeviceBuffer(MAX_PIXELS * 2 * 4)
    try:
        for N, H, W in ((2, 100, 130), (1, 37, 50)):
            print(f"no-alloc forward: N={N} H={H} W={W}", flush=True)
            feat = rng.normal(0, 1, (N, H, W, 6)).astype(np.float32)
            want = cl.forward(feat)
            feat_buf.upload(feat)
            cl.forward_noalloc(feat_buf.ptr.value, N, H, W, out_buf.ptr.value)
            got = out_buf.download((N, H, W, 2))
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (N, H, W)
        # the full size of the array: one 512 x 512 image
        feat = rng.normal(0, 1, (1, 512, 512, 6)).astype(np.float32)
        want = cl.forward(feat)
        feat_buf.upload(feat)
        cl.forward_noalloc(feat_buf.ptr.value, 1, 512, 512, out_buf.ptr.value)
        got = out_buf.download((1, 512, 512, 2))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    finally:
        feat_buf.free()
        out_buf.free()
        cl.close()


def test_forward_noalloc_rejects_batches_beyond_the_scratch():
    from fip_amd.sim import Closure, FipError
    cl = Closure(64, 64, 0, seed=6)
    feat_buf = DeviceBuffer(2 * 512 * 512 * 6 * 4)
    out_buf = DeviceBuffer(2 * 512 * 512 * 2 * 4)
    try:
        with pytest.raises(FipError, match="exceeds the scratch"):
            cl.forward_noalloc(feat_buf.ptr.value, 2, 512, 512, out_buf.ptr.value)
        with pytest.raises(FipError, match="empty"):
            cl.forward_noalloc(feat_buf.ptr.value, 0, 8, 8, out_buf.ptr.value)
    finally:
        feat_buf.free()
        out_buf.free()
        cl.close()
