"""fx_bloom_device_noalloc (static scratch, no allocation) against fx_bloom_device through the C ABI.

Two calls in a row, the second on a smaller canvas, each bit-identical to the allocating entry
point on the same device input; sizes beyond the static scratch fail with a clear error.
"""
import ctypes as C

import numpy as np
import pytest

from fenix_amd import abi
from fenix_amd.render import BloomPass
from fenix_amd.simulation import CombustionVolumeSimulation, make_desc

pytestmark = pytest.mark.gpu


def _hdr_image(h, w, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    img = np.zeros((h, w, 4), np.float32)
    for _ in range(8):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(4, 30)
        g = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
        img[..., :3] += g[..., None] * rng.uniform(0.2, 2.5, 3)
        img[..., 3] += g * rng.uniform(0.3, 1.0)
    img[..., 3] = np.clip(img[..., 3], 0, 1)
    return img.astype(np.float16)


class _Dev:
    """A device buffer allocated by the test (the entry point under test allocates nothing)."""

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes, self.ptr = hip, nbytes, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0

    def upload(self, arr):
        assert self.hip.hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0

    def download(self, arr):
        assert self.hip.hipMemcpy(arr.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(arr.nbytes), 2) == 0

    def free(self):
        self.hip.hipFree(self.ptr)


def test_noalloc_bloom_matches_allocating_bloom_on_two_calls():
    hip = C.CDLL("libamdhip64.so")
    lib = abi.load()
    sim = CombustionVolumeSimulation(make_desc((16, 16, 16)))
    params = BloomPass.params_struct(enabled=True, threshold=0.6, strength=0.9, radius=0.7)
    sizes = [(1280, 720), (333, 211)]  # the second call is smaller; its input differs too
    for k, (w, h) in enumerate(sizes):
        src = _hdr_image(h, w, seed=10 + k)
        dsrc, dref, dout = _Dev(hip, src.nbytes), _Dev(hip, w * h * 4), _Dev(hip, w * h * 4)
        dsrc.upload(src)
        abi.check(lib, sim._h, lib.fx_bloom_device(sim._h, C.byref(params), w, h, dsrc.ptr, dref.ptr))
        abi.check(lib, sim._h, lib.fx_bloom_device_noalloc(sim._h, C.byref(params), w, h, dsrc.ptr, dout.ptr))
        sim.synchronize()
        ref = np.empty((h, w, 4), np.uint8)
        out = np.empty((h, w, 4), np.uint8)
        dref.download(ref)
        dout.download(out)
        for b in (dsrc, dref, dout):
            b.free()
        np.testing.assert_array_equal(out, ref, err_msg=f"call {k} ({w}x{h})")
        assert out[..., 3].max() > 100 and np.any(out[..., :3] > 0)
    # beyond the static scratch: rejected on the host with a message before any launch
    w = h = 2050
    dsrc, dout = _Dev(hip, w * h * 8), _Dev(hip, w * h * 4)
    rc = lib.fx_bloom_device_noalloc(sim._h, C.byref(params), w, h, dsrc.ptr, dout.ptr)
    dsrc.free()
    dout.free()
    assert rc == abi.FX_ERR_INVALID_ARG
    assert b"static scratch" in lib.fx_last_error(sim._h)
    sim.dispose()
