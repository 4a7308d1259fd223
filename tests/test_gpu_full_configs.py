"""The benchmark configurations at their stated sizes.

Headline (1M SH3 in a 4 m cube, 64-camera 1920x1080 orbit, a batch of 8
  views): one view of the batch against the CPU checker -- image,
  transmittance, last contributor and per-tile lists bit-exact -- and its
  loss within 1e-5 of the float64 evaluation, the bar of the smaller parity
  tests (the engine's serial float sum drifts by ~1e-3 relative at this
  size, so the checker's own loss is not the yardstick here).
C3 (500k SH3, 32 orbit views of four resolutions in one batch): the batched
  adjoint against the sum of the 32 single-view adjoints (float rounding
  only); graph-replayed, autotuned training steps against the plain launch
  sequence bit for bit (deterministic mode); the smallest view's adjoint
  within 1e-3 relative L2 of the FP64 checker.
C4 (300k -> 1M cap SH3, 64-camera 1080p orbit, 2,000 steps, densify every
  100): every densify call's source indices, counts, RNG state and surviving
  parameters equal the CPU densify on the same state; the run clones, splits
  and reaches the cap. At the engine's 2e-4 threshold this scene's mean
  view-space gradient never qualifies (its 1080p views spread the loss over
  2M pixels each), so the threshold is the 80th percentile of the first
  call's statistic, fixed for the run.
C5 (3M SH3 room, 8 interior 1080p views): deterministic gradients bitwise
  equal run to run, the atomic mode within float rounding of them, and one
  view's image and per-tile lists bit-exact against the checker.
"""
import numpy as np
import pytest

import oracle as O
from parallax_amd import splat as S
from parallax_amd import synthetic as syn

pytestmark = pytest.mark.gpu

KEYS = ("pos", "log_scale", "rot", "opacity", "dc", "rest")


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(1e-30, np.linalg.norm(b)))


def device_targets(ctx, truth, views, deg):
    ctx.set_capacity(truth.n, truth.sh_degree)
    ctx.upload_cloud(truth)
    ctx.upload_views(views)
    out = []
    for k in range(len(views)):
        ctx.render_forward([k], S.options(sh_degree=deg))
        out.append(ctx.image(0)[0])
    return out


def check_view_against_checker(ctx, slot, cloud, view, tag):
    rgb, T, last = ctx.image(slot)
    cols, rows, offs, ents = ctx.tiles(slot)
    print(f"[{tag}] checker forward of batch slot {slot} ({view.width}x{view.height}, {len(ents)} pairs)", flush=True)
    f = O.Frame(cloud, view, O.default_options(sh_degree=3))
    orgb, oT, olast = f.image()
    assert np.array_equal(rgb, orgb), f"{tag}: image"
    assert np.array_equal(T, oT), f"{tag}: transmittance"
    assert np.array_equal(last, olast), f"{tag}: last contributor"
    _, _, ooffs, oents = f.tiles()
    assert np.array_equal(offs, ooffs) and np.array_equal(ents, oents), f"{tag}: per-tile lists"
    return orgb


def test_headline_batch_view_matches_the_checker(ctx):
    cloud, truth, views = syn.config4()
    batch = views[:8]
    print("[headline] device targets", flush=True)
    targets = device_targets(ctx, truth, batch, 3)
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(batch, targets)
    ctx.render_forward(list(range(8)), S.options(sh_degree=3))
    losses = ctx.loss(0.2)
    print(f"[headline] batch of 8: {ctx.last_pair_count} pairs", flush=True)
    slot = 5
    orgb = check_view_against_checker(ctx, slot, cloud, batch[slot], "headline")
    # the engine sums the loss's 6.2M terms serially in one float
    # (SplatRaster.cpp:653-682), which drifts by ~1e-3 relative at 1080p; the
    # device sums per tile, then the tiles in a fixed tree. Both are measured
    # against the float64 evaluation.
    import ssim64
    v = batch[slot]
    l64, _ = ssim64.photometric_loss(orgb.astype(np.float64), targets[slot].astype(np.float64), v.width, v.height,
                                     0.2)
    oloss, _ = O.photometric_loss(orgb, targets[slot], v.width, v.height, 0.2)
    print(f"[headline] loss {losses[slot]:.8f}, float64 {l64:.8f}, checker's serial float sum {oloss:.8f}",
          flush=True)
    assert abs(float(losses[slot]) - l64) <= 1e-5


def test_c3_batched_adjoint_graph_and_autotune_full_scale(ctx):
    cloud, truth, views = syn.config3()
    print("[c3] device targets", flush=True)
    targets = device_targets(ctx, truth, views, 3)
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, targets)
    opts = S.options(sh_degree=3, deterministic=True)
    ids = list(range(len(views)))
    # batched adjoint vs the sum of single-view adjoints
    ctx.zero_gradients()
    ctx.render_forward(ids, opts)
    ctx.loss(0.2)
    ctx.render_backward()
    batched = ctx.gradients()
    print(f"[c3] batched adjoint over {len(views)} views, {ctx.last_pair_count} pairs", flush=True)
    ctx.zero_gradients()
    for k in ids:
        ctx.render_forward([k], opts)
        ctx.loss(0.2)
        ctx.render_backward()
    summed = ctx.gradients()
    for k in KEYS + ("screen_gradient",):
        r = rel_l2(batched[k], summed[k])
        print(f"[c3] batched vs summed single views, {k}: rel-L2 {r:.3g}", flush=True)
        assert r <= 1e-5, k
    assert np.array_equal(batched["visible_count"], summed["visible_count"])
    del batched, summed
    # the smallest view against the FP64 checker
    k = min(ids, key=lambda i: views[i].pixels)
    ctx.zero_gradients()
    ctx.render_forward([k], opts)
    ctx.loss(0.2)
    ctx.render_backward()
    g = ctx.gradients()
    dl = ctx.image_gradient(0)
    print(f"[c3] FP64 checker adjoint of view {k} ({views[k].width}x{views[k].height})", flush=True)
    f = O.Frame(cloud, views[k], O.default_options(sh_degree=3))
    og = f.backward_f64(dl, O.ccloud_d(cloud))
    got = np.concatenate([np.asarray(g[key], np.float64).ravel() for key in KEYS])
    r = rel_l2(got, og.flat())
    print(f"[c3] view {k} gradient rel-L2 vs FP64 {r:.3g}", flush=True)
    assert r <= 1e-3
    del g, og, f
    # graph-replayed, autotuned steps vs plain steps, from the same state
    results = {}
    for mode in ("plain", "graph+autotune"):
        ctx.set_graph_mode(False)
        ctx.set_autotune(False)
        ctx.set_raster_config(None)
        ctx.set_capacity(cloud.n, 3)
        ctx.upload_cloud(cloud)
        ctx.upload_views(views, targets)
        if mode != "plain":
            ctx.set_autotune(True)
            ctx.set_graph_mode(True)
        steps = []
        for it in range(3):
            losses = ctx.train_step(ids, opts, 0.2, S.adam_params(it, 1e-4))
            steps.append((losses, ctx.download_cloud()))
        results[mode] = steps
        print(f"[c3] {mode}: 3 steps, launch shape {ctx.raster_config()}", flush=True)
    ctx.set_graph_mode(False)
    ctx.set_autotune(False)
    ctx.set_raster_config(None)
    for it, (a, b) in enumerate(zip(results["plain"], results["graph+autotune"])):
        assert np.array_equal(a[0], b[0]), f"step {it}: losses"
        for key in KEYS:
            assert np.array_equal(getattr(a[1], key), getattr(b[1], key)), f"step {it}: {key}"


def densify_params(cfg):
    p = S.CDensify()
    for k in ("densify_grad_threshold", "percent_dense", "split_scale_divider", "opacity_reset_interval",
              "prune_opacity", "prune_scale_fraction", "max_gaussians"):
        setattr(p, k, getattr(cfg, k))
    return p


def test_c4_densifying_run_to_the_cap_full_scale(ctx):
    start = syn.random_cloud(300_000, 0, 3, (-2, -2, -2), (2, 2, 2))
    truth = syn.random_cloud(1_000_000, 1, 3, (-2, -2, -2), (2, 2, 2))
    views = syn.orbit_views(64, 6.0, [(1920, 1080)], 2)
    cap, steps = 1_000_000, 2000
    print("[c4] device targets", flush=True)
    targets = device_targets(ctx, truth, views, 3)
    del truth
    extent = O.camera_extent(views)
    ctx.set_capacity(cap, 3)
    ctx.upload_cloud(start)
    ctx.upload_views(views, targets)
    del targets
    cfg = O.default_config(iterations=steps, max_gaussians=cap, prune_opacity=0.005)
    order = np.random.default_rng(3).permutation(steps) % len(views)
    ratio = np.float32(cfg.position_lr_final) / np.float32(cfg.position_lr_initial)
    rng_state, totals, sizes, engine_qualified = 7, np.zeros(3, np.int64), [], None
    ctx.set_graph_mode(True)
    try:
        for it in range(steps):
            t = np.float32(it) / np.float32(steps - 1)
            lr = float(np.float32(cfg.position_lr_initial) * np.power(ratio, t, dtype=np.float32) * np.float32(extent))
            ctx.train_step([int(order[it])], S.options(sh_degree=min(3, it // 1000)), 0.2, S.adam_params(it, lr),
                           want_loss=False)
            if (it + 1) % 100 or it + 1 >= steps:
                continue
            before = ctx.download_cloud()
            sg, vc = ctx.stats()
            mean = sg / np.maximum(vc, 1)
            if engine_qualified is None:
                engine_qualified = int(np.count_nonzero((vc > 0) & (mean >= np.float32(2e-4))))
                cfg.densify_grad_threshold = float(np.quantile(mean, 0.8))
                print(f"[c4] first call: {engine_qualified} gaussians at the engine's 2e-4 (max mean gradient "
                      f"{mean.max():.3g}); threshold {cfg.densify_grad_threshold:.3g} (80th percentile)", flush=True)
            src, counts, state = ctx.densify(densify_params(cfg), extent, it + 1, rng_state)
            after = ctx.download_cloud()
            want, osrc, ocounts, ostate = O.densify(before, sg, vc, cfg, extent, it + 1, rng_state)
            assert np.array_equal(src, osrc), f"step {it + 1}: source indices"
            assert np.array_equal(counts, ocounts) and state == ostate, f"step {it + 1}: counts / RNG"
            kept = src >= 0
            for k in KEYS:
                assert np.array_equal(getattr(after, k)[kept], getattr(want, k)[kept]), f"step {it + 1}: {k}"
                np.testing.assert_allclose(getattr(after, k), getattr(want, k), rtol=1e-5, atol=1e-6)
            rng_state = state
            totals += counts
            sizes.append(after.n)
            print(f"[c4] step {it + 1}: {before.n} -> {after.n} (clone {counts[0]}, split {counts[1]}, prune "
                  f"{counts[2]}), indices equal", flush=True)
            del before, after, want
    finally:
        ctx.set_graph_mode(False)
    assert len(sizes) == 19
    assert totals[0] > 0 and totals[1] > 0, "the run must clone and split"
    assert max(sizes) == cap, f"the run must reach the cap: {sizes}"


def test_c5_room_full_scale(ctx):
    cloud, truth, views = syn.config5()
    print("[c5] device targets", flush=True)
    targets = device_targets(ctx, truth, views, 3)
    del truth
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, targets)
    ids = list(range(len(views)))
    runs = {}
    for det, reps in ((True, 2), (False, 1)):
        for r in range(reps):
            ctx.zero_gradients()
            ctx.render_forward(ids, S.options(sh_degree=3, deterministic=det))
            ctx.loss(0.2)
            ctx.render_backward()
            g = ctx.gradients()
            runs[(det, r)] = (np.concatenate([np.asarray(g[k]).ravel() for k in KEYS]), g["screen_gradient"],
                              g["visible_count"])
            del g
            print(f"[c5] deterministic={det} run {r}: {ctx.last_pair_count} pairs", flush=True)
    a, b, c = runs[(True, 0)], runs[(True, 1)], runs[(False, 0)]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    between = rel_l2(c[0], a[0])
    print(f"[c5] atomic vs deterministic gradients rel-L2 {between:.3g}", flush=True)
    assert between <= 1e-5 and np.array_equal(c[2], a[2])
    del runs, a, b, c
    check_view_against_checker(ctx, 2, cloud, views[2], "c5")
