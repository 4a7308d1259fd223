"""splat_loss_noalloc (the loss entry point that allocates nothing) against
splat_loss on the same images: dL/drgb and the loss value bit-identical, on
two calls in a row with the second image smaller, plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from parallax_amd import splat as S
from parallax_amd import synthetic as syn

H2D, D2H = 1, 2


class DeviceMemory:
    """hipMalloc'd buffers of the runtime libparallax_splat.so is linked to
    (the test allocates; the entry point under test does not)."""

    def __init__(self):
        S.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMalloc.restype = C.c_int
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipFree.restype = C.c_int
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemcpy.restype = C.c_int
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.hip.hipMemset.restype = C.c_int
        self.ptrs = []

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, dev: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(C.c_void_p(dev), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, dev: int, a: np.ndarray):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(dev), a.nbytes, D2H) == 0
        return a

    def poison(self, dev: int, nbytes: int):
        assert self.hip.hipMemset(C.c_void_p(dev), 0xFF, C.c_size_t(nbytes)) == 0  # NaN floats

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(C.c_void_p(p))
        self.ptrs = []


@pytest.mark.gpu
def test_loss_noalloc_equals_splat_loss(ctx):
    cloud = syn.random_cloud(4000, 11, 1, (-1, -1, 2), (1, 1, 6))
    big = syn.pinhole(200, 150, 160.0)
    small = syn.pinhole(120, 90, 100.0)
    rng = np.random.default_rng(5)
    targets = [rng.random(v.width * v.height * 3, dtype=np.float32) for v in (big, small)]
    ctx.set_capacity(cloud.n, 1)
    ctx.upload_cloud(cloud)
    ctx.upload_views([big, small], targets)
    mem = DeviceMemory()
    try:
        n_big = big.width * big.height * 3
        d_rgb, d_tgt, d_dl = (mem.alloc(4 * n_big) for _ in range(3))
        d_loss = mem.alloc(256)
        # two calls in a row, the second on a smaller image; then the L1-only form
        for slot, view, w in ((0, big, 0.2), (1, small, 0.2), (1, small, 0.0)):
            print(f"loss_noalloc: view {view.width}x{view.height}, ssim weight {w}", flush=True)
            ctx.render_forward([slot], S.options(sh_degree=1))
            rgb, _, _ = ctx.image(0)
            want_loss = ctx.loss(w)[0]
            want_dl = ctx.image_gradient(0)
            n = view.width * view.height * 3
            mem.put(d_rgb, rgb)
            mem.put(d_tgt, targets[slot])
            mem.poison(d_dl, 4 * n_big)
            mem.poison(d_loss, 4)
            ctx.loss_noalloc(d_rgb, d_tgt, view.width, view.height, w, d_dl, d_loss)
            ctx.synchronize()
            got_dl = mem.get(d_dl, np.zeros(n, np.float32))
            got_loss = mem.get(d_loss, np.zeros(1, np.float32))
            assert np.array_equal(got_dl.view(np.uint32), want_dl.view(np.uint32)), \
                f"dL/drgb differs at {view.width}x{view.height}, w={w}"
            assert got_loss.view(np.uint32)[0] == np.float32(want_loss).view(np.uint32), (got_loss[0], want_loss)
        # beyond the fixed scratch: refused before anything is enqueued
        with pytest.raises(S.SplatError) as err:
            ctx.loss_noalloc(d_rgb, d_tgt, 2048, 1025, 0.2, d_dl, d_loss)
        assert err.value.status == 4 and "scratch" in str(err.value)
        with pytest.raises(S.SplatError):
            ctx.loss_noalloc(d_rgb, d_tgt, 16384, 1, 0.2, d_dl, d_loss)
        ctx.synchronize()
    finally:
        mem.free()
