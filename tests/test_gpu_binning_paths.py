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

The sorts' plans, each against the checker too:

* depths from 0.05 to 400 m (a 27-bit span of float keys) in 8 views: the
  view bits fit above the depth bits (30 of 31), and the depth sort runs all
  four 9-bit passes;
* the same depths in 64 views: 33 bits do not fit, so the slots are not
  grouped by view and the super-tile sort runs its second pass, by view;
* a view of 18 x 18 super-tiles: more than 256, so super-tiles are numbered
  across the batch and the super-tile sort takes the plain route;
* every splat behind the cameras: no kept key, empty lists (with a digit
  table of one scan chunk and of two).
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


def deep_cloud(n, seed, near=0.05, far=400.0):
    """Splats inside a +-0.45 (x/z, y/z) cone in front of the origin, depths
    log-uniform in [near, far], sizes growing with depth (the cameras below
    look along +z from the origin)."""
    rng = np.random.default_rng(seed)
    cloud = syn.random_cloud(n, seed, 1, (-1, -1, 1), (1, 1, 2))
    z = np.exp(rng.uniform(np.log(near), np.log(far), size=n))
    uv = rng.uniform(-0.45, 0.45, size=(n, 2))
    cloud.pos[:] = np.stack([uv[:, 0] * z, uv[:, 1] * z, z], axis=1).astype(np.float32)
    cloud.log_scale[:] = (np.log(0.004 * z)[:, None] + rng.normal(0.0, 0.3, size=(n, 3))).astype(np.float32)
    return cloud


def batch_against_checker(ctx, cloud, views, check, sh=1):
    ctx.set_capacity(cloud.n, sh)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views)
    opts = S.options(sh_degree=sh)
    ctx.render_forward(list(range(len(views))), opts)
    batch = {k: ctx.tiles(k) for k in check}
    lists_equal_checker(ctx, cloud, views, [(k, views[k]) for k in check], sh)
    return batch, opts


@pytest.mark.parametrize("n_views", [8, 64])
def test_depth_keys_over_27_bits(ctx, n_views):
    cloud = deep_cloud(3000, 31)
    keys = cloud.pos[:, 2].astype(np.float32).view(np.uint32)
    span = int(keys.max()) - int(keys.min())
    assert span.bit_length() == 27  # with 3 view bits: 30, four 9-bit passes; with 6: 33, not grouped
    views = [syn.pinhole(320, 200, 300.0)] * n_views
    check = [0, n_views // 2, n_views - 1]
    print(f"[bin] depths 0.05..400 m in {n_views} views", flush=True)
    batch, opts = batch_against_checker(ctx, cloud, views, check)
    ctx.render_forward([0], opts)
    one = ctx.tiles(0)
    assert np.array_equal(one[2], batch[0][2]) and np.array_equal(one[3], batch[0][3])


def test_super_tiles_numbered_across_the_batch(ctx):
    cloud = syn.random_cloud(6000, 32, 1, (-2, -2, -2), (2, 2, 2), log_scale_mu=-2.5)
    big = syn.orbit_views(1, 6.0, [(2304, 2304)], seed=5)[0]  # 18 x 18 super-tiles
    small = syn.orbit_views(1, 6.0, [(320, 200)], seed=6)[0]
    print("[bin] 2304^2 + 320x200 (more than 256 super-tiles in a view)", flush=True)
    batch_against_checker(ctx, cloud, [big, small], [0, 1])


@pytest.mark.parametrize("n", [2000, 30000])
def test_every_splat_culled(ctx, n):
    # 6,000 slots: the depth sort's digit table is one scan chunk; 90,000
    # slots (11 blocks of 8192): two chunks, whose zero total the scatter's
    # block 0 writes as the kept count
    cloud = syn.random_cloud(n, 33, 1, (-1, -1, -3), (1, 1, -2))  # behind the cameras
    views = [syn.pinhole(320, 200, 300.0)] * 3
    print(f"[bin] nothing in front of the cameras, {n} splats", flush=True)
    ctx.set_capacity(cloud.n, 1)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views)
    ctx.render_forward([0, 1, 2], S.options(sh_degree=1))
    for k in range(3):
        cols, rows, offs, ents = ctx.tiles(k)
        assert len(ents) == 0 and not offs.any(), k
