"""Every compositing launch shape gives the same result.

The forward runs as 16x16, 8x8 or 32x8 workgroups, each with one workgroup
per unit or as a persistent grid over a work queue (tiles, longest list
first); the adjoint stages 64 or 128 splats per batch, with either dispatch.
These are scheduling choices only: images, transmittance and last
contributors must be identical, and so must deterministic-mode gradients.
The autotuner picks one of them per resolution mix.
"""
import numpy as np
import pytest

import oracle as O
from parallax_amd import splat as S
from parallax_amd import synthetic as syn

pytestmark = pytest.mark.gpu

FWD = [(shape, queue) for shape in (0, 1, 2) for queue in (False, True)]
BWD = [(batch, queue) for batch in (64, 128) for queue in (False, True)]


def grads_flat(g):
    return np.concatenate([np.asarray(g[k]).ravel() for k in ("pos", "log_scale", "rot", "opacity", "dc", "rest")])


@pytest.fixture(scope="module")
def scene():
    cloud = syn.random_cloud(30_000, 41, 3, (-2, -2, -2), (2, 2, 2))
    views = syn.orbit_views(3, 6.0, [(160, 120), (200, 100), (97, 61)], seed=42)
    truth = syn.random_cloud(30_000, 43, 3, (-2, -2, -2), (2, 2, 2))
    targets = [O.render(truth, v, O.default_options(sh_degree=3))[0] for v in views]
    return cloud, views, targets


def step(ctx, deterministic):
    ctx.zero_gradients()
    ctx.render_forward([0, 1, 2], S.options(sh_degree=3, deterministic=deterministic))
    images = [ctx.image(k) for k in range(3)]
    ctx.loss(0.2)
    ctx.render_backward()
    return images, ctx.gradients()


@pytest.mark.parametrize("deterministic", [True, False], ids=["deterministic", "atomic"])
def test_every_launch_shape_gives_the_same_result(ctx, scene, deterministic):
    cloud, views, targets = scene
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, targets)
    ctx.set_raster_config(0, False, 128, False)
    ref_images, ref_grads = step(ctx, deterministic)
    ref = grads_flat(ref_grads)
    for (shape, fq), (batch, bq) in zip(FWD, BWD + BWD[:2]):
        print(f"[variants] forward {shape}/{fq} backward {batch}/{bq} det={deterministic}", flush=True)
        ctx.set_raster_config(shape, fq, batch, bq)
        images, grads = step(ctx, deterministic)
        cfg = ctx.raster_config()
        assert cfg["backward_batch"] == batch and cfg["forward_queue"] == fq
        for k in range(3):
            for a, b in zip(images[k], ref_images[k]):
                assert np.array_equal(a, b), (shape, fq, k)
        got = grads_flat(grads)
        if deterministic:
            assert np.array_equal(got, ref), (batch, bq)
            assert np.array_equal(grads["screen_gradient"], ref_grads["screen_gradient"])
        else:
            assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref), (batch, bq)
    ctx.set_raster_config(None)


def test_autotune_keeps_a_valid_shape_with_the_same_result(ctx, scene):
    cloud, views, targets = scene
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, targets)
    ctx.set_raster_config(None)
    ctx.set_autotune(False)
    ref_images, ref_grads = step(ctx, True)
    ctx.set_autotune(True)
    images, grads = step(ctx, True)  # tunes this resolution mix
    cfg = ctx.raster_config()
    print(f"[variants] tuned {cfg}", flush=True)
    assert cfg["forward_shape"] in ("16x16", "8x8", "32x8") and cfg["backward_batch"] in (64, 128)
    for k in range(3):
        for a, b in zip(images[k], ref_images[k]):
            assert np.array_equal(a, b)
    assert np.array_equal(grads_flat(grads), grads_flat(ref_grads))
    again_images, _ = step(ctx, True)  # the tuned shape, without tuning again
    assert ctx.raster_config() == cfg
    for a, b in zip(again_images[0], ref_images[0]):
        assert np.array_equal(a, b)
    ctx.set_autotune(False)
