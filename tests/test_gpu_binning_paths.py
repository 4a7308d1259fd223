"""The binning's two ways of reading each slot's tile rect in depth order,
each against the oracle's buildTiles (SplatRaster.cpp:114-151), bit for bit:

* packed 32-bit rects riding along with the depth sort, where every view of
  the batch is at most 256 x 256 tiles (4096 pixels) -- the benchmark's
  batches and every other test; here also a batch of 17 views of 2048 x 2048
  (4352 super-tiles, a 13-bit super-tile sort);
* the 8-byte rect gathers, for a view wider than 4096 pixels (260 tile
  columns) in a batch with a small view.

The same views binned one at a time and in the big batch must also give the
same lists, and deterministic mode must repeat bit for bit on each path.
"""
import numpy as np
import pytest

import oracle as O
from parallax_amd import splat as S
from parallax_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def lists_equal_checker(ctx, cloud, views, slot_views, sh):
    for k, v in slot_views:
        f = O.Frame(cloud, v, O.default_options(sh_degree=sh))
        cols, rows, offs, ents = ctx.tiles(k)
        ocols, orows, ooffs, oents = f.tiles()
        assert (cols, rows) == (ocols, orows), k
        assert np.array_equal(offs, ooffs), k
        assert np.array_equal(ents, oents), k
        del f


def test_batch_over_4096_super_tiles_matches_the_checker(ctx):
    cloud = syn.random_cloud(6000, 21, 1, (-2, -2, -2), (2, 2, 2), log_scale_mu=-2.5)
    views = syn.orbit_views(17, 6.0, [(2048, 2048)], seed=4)
    ctx.set_capacity(cloud.n, 1)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views)
    opts = S.options(sh_degree=1)
    print("[bin] 17 x 2048^2 in one batch (4352 super-tiles)", flush=True)
    ctx.render_forward(list(range(17)), opts)
    batch = [ctx.tiles(k) for k in range(17)]
    lists_equal_checker(ctx, cloud, views, [(k, views[k]) for k in (0, 8, 16)], 1)
    print("[bin] the same views one at a time", flush=True)
    for k in range(17):
        ctx.render_forward([k], opts)
        one = ctx.tiles(0)
        assert one[:2] == batch[k][:2]
        assert np.array_equal(one[2], batch[k][2]) and np.array_equal(one[3], batch[k][3]), k


def test_view_wider_than_4096_pixels_matches_the_checker(ctx):
    cloud = syn.random_cloud(8000, 22, 1, (-2, -1, -2), (2, 1, 2), log_scale_mu=-3.0)
    wide = syn.pinhole(4160, 96, 900.0)
    small = syn.pinhole(320, 200, 300.0)
    ctx.set_capacity(cloud.n, 1)
    ctx.upload_cloud(cloud)
    ctx.upload_views([wide, small])
    print("[bin] 4160x96 + 320x200 in one batch", flush=True)
    ctx.render_forward([0, 1], S.options(sh_degree=1))
    lists_equal_checker(ctx, cloud, [wide, small], [(0, wide), (1, small)], 1)


def test_deterministic_mode_on_every_binning_path(ctx):
    """Deterministic mode repeats bit for bit on each path: a view binned alone
    (packed rects), and with a >4096-pixel view (gathered rects). Across the two,
    the gradients agree to float rounding only: the fixed-order reduction of
    a splat's per-pair partials runs over the batch's record layout, so its
    summation tree depends on what else the batch holds."""
    cloud = syn.random_cloud(5000, 23, 1, (-2, -1, -2), (2, 1, 2), log_scale_mu=-3.0)
    view = syn.pinhole(320, 200, 300.0)
    wide = syn.pinhole(4160, 64, 900.0)
    tgt = np.random.default_rng(3).uniform(size=view.pixels * 3).astype(np.float32)
    wtgt = np.zeros(wide.pixels * 3, np.float32)
    ctx.set_capacity(cloud.n, 1)
    ctx.upload_cloud(cloud)
    ctx.upload_views([view, wide], [tgt, wtgt])
    opts = S.options(sh_degree=1, deterministic=True)

    def run(ids):
        ctx.zero_gradients()
        ctx.render_forward(ids, opts)
        ctx.loss(0.2)
        if len(ids) == 2:  # only view 0's image gradient
            ctx.set_image_gradient(1, np.zeros(wide.pixels * 3, np.float32))
        ctx.render_backward()
        g = ctx.gradients()
        return np.concatenate([np.asarray(g[k], np.float64).ravel() for k in ("pos", "opacity", "dc")])

    alone = [run([0]) for _ in range(2)]
    mixed = [run([0, 1]) for _ in range(2)]
    assert np.array_equal(alone[0], alone[1])
    assert np.array_equal(mixed[0], mixed[1])
    rel = np.linalg.norm(alone[0] - mixed[0]) / np.linalg.norm(alone[0])
    print(f"[bin] packed vs gathered-rect batch: gradients rel-L2 {rel:.3g}", flush=True)
    assert rel <= 1e-6
