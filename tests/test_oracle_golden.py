"""The oracle against the golden vectors written by the reference's own
sources (tests/golden/make_golden.py). The reference build uses
-ffast-math and libm's expf; the oracle is strict IEEE with the deterministic
exponential, so values agree to float rounding and discrete outputs (which
splats project, their order, visible counts, densify indices) agree exactly.
"""
import os

import numpy as np
import pytest

import oracle as O
from parallax_amd import synthetic as syn

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def cloud_from(d, prefix="in_"):
    return syn.Cloud(d[prefix + "pos"], d[prefix + "log_scale"], d[prefix + "rot"], d[prefix + "opacity"],
                     d[prefix + "dc"], d[prefix + "rest"], int(d[prefix + "sh_degree"])).contiguous()


def view_from(d):
    fx, fy, cx, cy = d["view_K"]
    w, h = d["view_size"]
    return syn.View(d["view_R"], d["view_t"], fx, fy, cx, cy, int(w), int(h))


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(1e-30, np.linalg.norm(np.asarray(b, np.float64))))


@pytest.mark.parametrize("name", ["fd16_sh2", "c1_small_sh0", "sh3_bg"])
def test_raster_golden(name):
    d = load(f"raster_{name}.npz")
    cloud, view = cloud_from(d), view_from(d)
    opt = O.default_options(sh_degree=int(d["opt_sh_degree"]), background=tuple(d["opt_background"]),
                            min_alpha=float(d["opt_min_alpha"]), transmittance_cutoff=float(d["opt_cutoff"]))
    f = O.Frame(cloud, view, opt)
    p = f.projected()
    assert np.array_equal(p["index"], d["proj_index"])
    for key, tol in (("mean", 1e-5), ("conic", 1e-5), ("cov2d", 1e-5), ("depth", 1e-6), ("radius", 1e-5),
                     ("color", 1e-5), ("opacity", 1e-6)):
        ref = d["proj_" + key]
        np.testing.assert_allclose(p[key], ref, rtol=tol, atol=tol * max(1.0, np.abs(ref).max()), err_msg=key)
    assert np.array_equal(p["clamped"], d["proj_clamped"])
    rgb, T, _ = f.image()
    assert np.abs(rgb - d["rgb"]).max() <= 1e-5
    assert np.abs(T - d["T"]).max() <= 1e-5
    loss, dl = O.photometric_loss(rgb, d["target"], view.width, view.height, float(d["ssim_weight"]))
    assert abs(loss - float(d["loss"])) <= 1e-5
    assert np.abs(dl - d["dL"]).max() <= 1e-5 * np.abs(d["dL"]).max()
    g = f.backward(d["dL"])
    ours = g.flat()
    theirs = np.concatenate([d["grad_" + k].ravel() for k in ("pos", "log_scale", "rot", "opacity", "dc", "rest")])
    assert rel_l2(ours, theirs) <= 1e-4
    assert np.array_equal(g.visible_count, d["grad_visible_count"])
    assert rel_l2(g.screen_gradient, d["grad_screen_gradient"]) <= 1e-4


@pytest.mark.parametrize("name", ["fd16_sh2", "c1_small_sh0", "sh3_bg"])
def test_float64_loss_yardstick_matches_the_reference(name):
    """The float64 loss restatement that measures dL/drgb accuracy reproduces
    the reference build's loss and dL/drgb on its own rendered image, and the
    oracle's float32 gradient stays within the accuracy bar the device is held
    to."""
    import ssim64
    d = load(f"raster_{name}.npz")
    w, h = (int(v) for v in d["view_size"])
    loss64, dl64 = ssim64.photometric_loss(d["rgb"], d["target"], w, h, float(d["ssim_weight"]))
    assert abs(loss64 - float(d["loss"])) <= 1e-5
    assert ssim64.errors(d["dL"], dl64)[0] <= 2e-5
    _, dl = O.photometric_loss(d["rgb"], d["target"], w, h, float(d["ssim_weight"]))
    worst, l2 = ssim64.errors(dl, dl64)
    assert worst <= 5e-5 and l2 <= 2e-5
    # a gradient with any one term dropped is far outside the bar
    for term in ("mu", "var", "var_mu", "cov", "cov_mu"):
        _, broken = ssim64.photometric_loss(d["rgb"], d["target"], w, h, float(d["ssim_weight"]), omit=term)
        assert ssim64.errors(broken, dl64)[0] > 1e-3, term


@pytest.mark.parametrize("name", ["prune", "cap"])
def test_densify_golden(name):
    d = load(f"densify_{name}.npz")
    cloud = cloud_from(d)
    cfg = O.default_config(max_gaussians=int(d["max_gaussians"]), opacity_reset_interval=3000)
    out, src, counts, rng = O.densify(cloud, d["screen_gradient"], d["visible_count"], cfg, float(d["extent"]),
                                      int(d["iteration"]), int(d["rng_in"]))
    assert np.array_equal(src, d["source_index"])
    assert np.array_equal(counts, d["counts"])
    assert rng == int(d["rng_out"])
    for k in ("pos", "log_scale", "rot", "opacity", "dc", "rest"):
        np.testing.assert_allclose(getattr(out, k), d["out_" + k], rtol=1e-5, atol=1e-6, err_msg=k)


@pytest.mark.parametrize("tag", ["plain", "densify"])
def test_trainer_golden(tag):
    d = load("trainer.npz")
    steps = int(d[f"{tag}_steps"])
    if tag == "plain":
        views, start, extra = syn.ring_of_views(6, 1.6, 32, 32), syn.rough_start(), {}
    else:
        views, start = syn.ring_of_views(4, 1.6, 16, 16), syn.rough_start(8)
        extra = dict(densify_from_iteration=20, densify_until_iteration=60, densify_interval=20,
                     densify_grad_threshold=0.0)
    base = dict(iterations=steps, sh_degree=0, sh_increase_interval=0, densify_from_iteration=steps + 1,
                densify_until_iteration=steps + 1, opacity_reset_interval=0, ssim_weight=0.2)
    base.update(extra)
    cfg = O.default_config(**base)
    t = O.Trainer(start, views, list(d[f"{tag}_images"]), cfg)
    losses = []
    while t.step():
        losses.append(t.last_loss)
    assert len(losses) == steps
    np.testing.assert_allclose(losses, d[f"{tag}_losses"], rtol=2e-3, atol=1e-5)
    assert t.count == int(d[f"{tag}_count"])
    c = t.cloud()
    np.testing.assert_allclose(c.pos, d[f"{tag}_out_pos"], atol=2e-3)
    np.testing.assert_allclose(c.opacity, d[f"{tag}_out_opacity"], atol=2e-2)
