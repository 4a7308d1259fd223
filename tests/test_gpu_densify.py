"""Device densify / prune / opacity reset against the oracle: source
indices, counts and the RNG state bit-exact; parameters of survivors and
clones bit-exact; split children to float tolerance (their normals use the
device's log and cos); moments carried by source index, new ones zero."""
import os

import numpy as np
import pytest

import oracle as O
from parallax_amd import splat as S
from parallax_amd import synthetic as syn

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def densify_params(cfg):
    p = S.CDensify()
    p.densify_grad_threshold = cfg.densify_grad_threshold
    p.percent_dense = cfg.percent_dense
    p.split_scale_divider = cfg.split_scale_divider
    p.opacity_reset_interval = cfg.opacity_reset_interval
    p.prune_opacity = cfg.prune_opacity
    p.prune_scale_fraction = cfg.prune_scale_fraction
    p.max_gaussians = cfg.max_gaussians
    return p


def run_case(ctx, cloud, sg, vc, cfg, extent, iteration, rng):
    ctx.set_capacity(max(1, 3 * cloud.n), cloud.sh_degree)
    ctx.upload_cloud(cloud)
    ctx.upload_stats(sg, vc)
    src, counts, state = ctx.densify(densify_params(cfg), extent, iteration, rng)
    return ctx.download_cloud(), src, counts, state


@pytest.mark.parametrize("name", ["prune", "cap"])
def test_densify_matches_oracle_on_golden_inputs(ctx, name):
    d = np.load(os.path.join(GOLDEN, f"densify_{name}.npz"))
    cloud = syn.Cloud(d["in_pos"], d["in_log_scale"], d["in_rot"], d["in_opacity"], d["in_dc"], d["in_rest"],
                      int(d["in_sh_degree"])).contiguous()
    cfg = O.default_config(max_gaussians=int(d["max_gaussians"]), opacity_reset_interval=3000)
    got, src, counts, state = run_case(ctx, cloud, d["screen_gradient"], d["visible_count"], cfg, float(d["extent"]),
                                       int(d["iteration"]), int(d["rng_in"]))
    assert np.array_equal(src, d["source_index"])
    assert np.array_equal(counts, d["counts"])
    assert state == int(d["rng_out"])
    exp_cloud, osrc, ocounts, ostate = O.densify(cloud, d["screen_gradient"], d["visible_count"], cfg,
                                                 float(d["extent"]), int(d["iteration"]), int(d["rng_in"]))
    assert np.array_equal(src, osrc) and state == ostate
    kept = src >= 0
    for k in ("pos", "log_scale", "rot", "opacity", "dc", "rest"):
        a, b = getattr(got, k), getattr(exp_cloud, k)
        assert np.array_equal(a[kept], b[kept]), k
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6, err_msg=k)


def test_densify_carries_moments_by_source_index(ctx):
    cloud = syn.random_cloud(3000, 31, 1, (-1, -1, 2), (1, 1, 6), log_scale_mu=-2.5, log_scale_sigma=1.0)
    view = syn.pinhole(160, 120, 150.0)
    tgt, _ = O.render(syn.random_cloud(3000, 32, 1, (-1, -1, 2), (1, 1, 6)), view, O.default_options(sh_degree=1))
    ctx.set_capacity(12000, 1)
    ctx.upload_cloud(cloud)
    ctx.upload_views([view], [tgt])
    ad = S.adam_params(0, 1e-3)
    for _ in range(3):
        ctx.train_step([0], S.options(sh_degree=1), 0.2, ad)
    m0, v0 = ctx.moments()
    before = ctx.download_cloud()
    sg, vc = ctx.stats()
    cfg = O.default_config(densify_grad_threshold=float(np.quantile(sg / np.maximum(vc, 1), 0.7)),
                           max_gaussians=12000)
    src, counts, state = ctx.densify(densify_params(cfg), 1.0, 100, 7)
    after = ctx.download_cloud()
    exp_cloud, osrc, ocounts, ostate = O.densify(before, sg, vc, cfg, 1.0, 100, 7)
    assert np.array_equal(src, osrc) and np.array_equal(counts, ocounts) and state == ostate
    assert counts[0] + counts[1] > 0
    m1, v1 = ctx.moments()
    n0, n1 = before.n, after.n
    widths = [3, 3, 4, 1, 3, before.rest_stride]

    def group(a, n, q):
        off = sum(w * n for w in widths[:q])
        return a[off:off + widths[q] * n].reshape(n, widths[q]) if widths[q] else np.zeros((n, 0), np.float32)

    for q in range(6):
        old, new = group(m0, n0, q), group(m1, n1, q)
        assert np.array_equal(new[src >= 0], old[src[src >= 0]])
        assert np.all(new[src < 0] == 0)
    assert np.all(ctx.stats()[1] == 0)


def test_reset_opacity_and_prune_and_sort(ctx):
    cloud = syn.random_cloud(5000, 41, 0, (-1, -1, 2), (1, 1, 6))
    cloud.opacity[:] = np.random.default_rng(4).normal(0.0, 4.0, size=cloud.n).astype(np.float32)
    ctx.set_capacity(cloud.n, 0)
    ctx.upload_cloud(cloud)
    ceiling = float(np.float32(np.log(np.float32(0.01) / (np.float32(1) - np.float32(0.01)))))
    ctx.reset_opacity(ceiling)
    got = ctx.download_cloud()
    assert np.array_equal(got.opacity, np.minimum(cloud.opacity, np.float32(ceiling)))
    ctx.upload_cloud(cloud)
    ctx.prune_and_sort(0.005)
    got = ctx.download_cloud()
    exp = cloud.copy()
    k = O.ccloud(exp)
    order = np.zeros(cloud.n, np.int32)
    n1 = O.lib().orc_prune_and_sort(O.C.byref(k.struct), 0.005, O.ptr(order, O.i32p))
    assert got.n == n1
    assert np.array_equal(got.opacity, k.arrays[3][:n1])
    assert np.array_equal(got.pos, k.arrays[0].reshape(-1, 3)[:n1])
