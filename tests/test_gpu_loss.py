"""The photometric loss kernels on image shapes that do not fill their work
units (128 x 86 pixels, statistics in 64-column halves) and in uneven
batches.

* Through splat_loss_noalloc, on seeded images of ragged sizes (a single
  pixel, thin strips, one pixel either side of a unit edge, 1080p): the loss
  within 1e-5 of the float64 evaluation of the engine's loss
  (SplatRaster.cpp:580-738), dL/drgb within the bars the oracle's own float32
  evaluation meets (max error 5e-5 of max |dL|, relative L2 2e-5), and the
  L1-only form (w = 0) exact: dL = sign(x - y) / (3 W H).
* Identical images: SSIM = 1, the loss 0 and dL/drgb 0 up to float32
  rounding residue below 5e-5 of a pixel's gradient scale 1 / (3 W H).
* A batch of three views of different sizes through splat_loss: per-view
  losses and dL/drgb bit-identical to the same views run one at a time (each
  view's partial sums are reduced in a fixed order of its own units).
"""
import numpy as np
import pytest

import oracle as O
import ssim64
from devmem import DeviceMemory
from parallax_amd import splat as S
from parallax_amd import synthetic as syn

pytestmark = pytest.mark.gpu

DL_MAX = 5e-5
DL_REL_L2 = 2e-5
SHAPES = [(1, 1), (5, 3), (127, 63), (128, 64), (129, 65), (96, 85), (96, 86), (96, 87), (131, 173), (200, 150),
          (7, 300), (300, 7), (1920, 1080)]


def images(w, h, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(size=w * h * 3).astype(np.float32)
    # a target correlated with the image, so SSIM is neither ~0 nor 1
    y = np.clip(x + rng.normal(0.0, 0.1, size=x.shape), 0.0, 1.0).astype(np.float32)
    return x, y


@pytest.fixture(scope="module")
def mem():
    m = DeviceMemory()
    yield m
    m.free()


def run_noalloc(ctx, mem, x, y, w, h, weight):
    n = w * h * 3
    d_x, d_y, d_dl, d_loss = mem.alloc(4 * n), mem.alloc(4 * n), mem.alloc(4 * n), mem.alloc(256)
    mem.put(d_x, x)
    mem.put(d_y, y)
    mem.poison(d_dl, 4 * n)
    ctx.loss_noalloc(d_x, d_y, w, h, weight, d_dl, d_loss)
    ctx.synchronize()
    dl = mem.get(d_dl, np.zeros(n, np.float32))
    loss = float(mem.get(d_loss, np.zeros(1, np.float32))[0])
    return loss, dl


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_loss_on_ragged_shapes_against_float64(ctx, mem, shape):
    w, h = shape
    x, y = images(w, h, 7 * w + h)
    print(f"[loss] {w}x{h}", flush=True)
    loss, dl = run_noalloc(ctx, mem, x, y, w, h, 0.2)
    l64, dl64 = ssim64.photometric_loss(x, y, w, h, 0.2)
    e_max, e_l2 = ssim64.errors(dl, dl64)
    print(f"[loss] {w}x{h}: loss {loss:.8f} vs float64 {l64:.8f}; dL max {e_max:.3g} rel-L2 {e_l2:.3g}", flush=True)
    assert abs(loss - l64) <= 1e-5
    assert e_max <= DL_MAX and e_l2 <= DL_REL_L2
    # the L1-only form
    loss0, dl0 = run_noalloc(ctx, mem, x, y, w, h, 0.0)
    inverse = np.float32(1.0) / np.float32(w * h * 3)
    want = (np.sign(x - y).astype(np.float32) * inverse).astype(np.float32)
    assert np.array_equal(dl0, want)
    assert abs(loss0 - float(np.abs(x.astype(np.float64) - y).mean())) <= 1e-6


def test_loss_of_identical_images(ctx, mem):
    w, h = 150, 70
    x, _ = images(w, h, 3)
    loss, dl = run_noalloc(ctx, mem, x, x.copy(), w, h, 0.2)
    l64, dl64 = ssim64.photometric_loss(x, x, w, h, 0.2)
    assert abs(loss) <= 1e-6 and abs(l64) <= 1e-12
    # the SSIM term's partials cancel: float32 leaves rounding residue far
    # below the scale of a pixel's gradient, 1 / (3 W H)
    assert np.abs(dl.astype(np.float64) - dl64).max() <= DL_MAX / (3 * w * h)


def test_batched_losses_equal_single_views(ctx):
    cloud = syn.random_cloud(6000, 12, 1, (-1, -1, 2), (1, 1, 6))
    views = [syn.pinhole(200, 150, 160.0), syn.pinhole(127, 63, 90.0), syn.pinhole(64, 200, 120.0)]
    rng = np.random.default_rng(9)
    targets = [rng.uniform(size=v.pixels * 3).astype(np.float32) for v in views]
    ctx.set_capacity(cloud.n, 1)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, targets)
    opts = S.options(sh_degree=1)
    ctx.render_forward([0, 1, 2], opts)
    batch_losses = ctx.loss(0.2)
    batch_dl = [ctx.image_gradient(k) for k in range(3)]
    for k in range(3):
        ctx.render_forward([k], opts)
        single = ctx.loss(0.2)
        assert single[0] == batch_losses[k], k
        assert np.array_equal(ctx.image_gradient(0), batch_dl[k]), k
        rgb, _, _ = ctx.image(0)
        oloss, _ = O.photometric_loss(rgb, targets[k], views[k].width, views[k].height, 0.2)
        assert abs(single[0] - oloss) <= 1e-5
