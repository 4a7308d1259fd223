"""splat_loss_noalloc (the loss entry point that allocates nothing) against
splat_loss on the same images: dL/drgb and the loss value bit-identical, on
two calls in a row with the second image smaller, plus the size limit."""
import numpy as np
import pytest

from parallax_amd import splat as S
from parallax_amd import synthetic as syn

from devmem import DeviceMemory


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
