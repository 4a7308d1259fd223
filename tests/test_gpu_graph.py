"""The captured-graph training step against the plain launch sequence.

splat_set_graph_mode makes splat_train_step replay one captured HIP graph per
batch shape. The replays must give what the plain launch sequence gives from
the same state: bit for bit in deterministic mode (same kernels, same
operands), to float-atomic reordering in atomic mode. Every replay is checked
against the plain run, with downloads and synchronisations between replays,
including a batch whose pair count outgrows the capacities the graphs run
with (detected on the device, the update skipped, the step redone) and
changes of batch shape (more graphs).

A training step applies Adam inside the per-gaussian adjoint. The separate
public calls (splat_render_forward, splat_loss, splat_render_backward,
splat_adam_step: the plain adjoint, then k_adam) must give the same
parameters, moments and statistics bit for bit in deterministic mode, with
every SH band in use and with fewer bands than the cloud stores.
"""
import numpy as np
import pytest

from parallax_amd import splat as S
from parallax_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(1e-30, np.linalg.norm(b)))


def scene():
    cloud = syn.random_cloud(20_000, 21, 3, (-2, -2, -2), (2, 2, 2))
    far = syn.orbit_views(3, 6.0, [(200, 150)], seed=22)
    big = syn.orbit_views(1, 6.0, [(640, 480)], seed=23)  # ~2x the pairs of a small view
    tall = syn.orbit_views(1, 6.0, [(120, 180)], seed=24)
    views = far + big + tall
    truth = syn.random_cloud(20_000, 25, 3, (-2, -2, -2), (2, 2, 2))
    import oracle as O
    targets = [O.render(truth, v, O.default_options(sh_degree=3))[0] for v in views]
    # 0-2 small views, 3 a large one, 4 another resolution; the first step sets
    # the capacities from one small view, the fourth needs four times as many
    batches = [[0], [1], [2], [3, 1, 2], [0], [4, 2], [1, 3], [0, 2], [2]]
    return cloud, views, targets, batches


def run(c, cloud, views, targets, batches, graph, deterministic):
    c.set_graph_mode(False)
    c.set_capacity(cloud.n, 3)
    c.upload_cloud(cloud)
    c.upload_views(views, targets)
    c.set_graph_mode(graph)
    out = []
    for it, b in enumerate(batches):
        opts = S.options(sh_degree=3, deterministic=deterministic)
        losses = c.train_step(b, opts, 0.2, S.adam_params(it, 1e-4))
        cl = c.download_cloud()
        sg, vc = c.stats()
        out.append((losses, cl, sg, vc, c.last_pair_count))
    c.set_graph_mode(False)
    return out


@pytest.mark.parametrize("deterministic", [True, False], ids=["deterministic", "atomic"])
def test_graph_steps_match_plain_steps(ctx, deterministic):
    cloud, views, targets, batches = scene()
    print(f"[graph] plain run, deterministic={deterministic}", flush=True)
    plain = run(ctx, cloud, views, targets, batches, False, deterministic)
    print(f"[graph] graph run, deterministic={deterministic}", flush=True)
    graph = run(ctx, cloud, views, targets, batches, True, deterministic)
    pairs = [p[4] for p in plain]
    assert max(pairs) > 1.25 * pairs[0] + 65536, "the large view must outgrow the first capacities"
    for it, (a, b) in enumerate(zip(plain, graph)):
        assert a[4] == b[4], f"step {it}: pair counts differ"
        np.testing.assert_array_equal(a[3], b[3], err_msg=f"step {it}: visible counts")
        for k in ("pos", "log_scale", "rot", "opacity", "dc", "rest"):
            if deterministic:
                assert np.array_equal(getattr(a[1], k), getattr(b[1], k)), f"step {it}: {k}"
            else:
                assert rel_l2(getattr(b[1], k), getattr(a[1], k)) <= 1e-4, f"step {it}: {k}"
        if deterministic:
            assert np.array_equal(a[0], b[0]), f"step {it}: losses"
            assert np.array_equal(a[2], b[2]), f"step {it}: screen gradients"
        else:
            assert np.allclose(a[0], b[0], rtol=1e-6), f"step {it}: losses"
            assert rel_l2(b[2], a[2]) <= 1e-4, f"step {it}: screen gradients"


def run_separate(c, cloud, views, targets, batches, sh):
    c.set_graph_mode(False)
    c.set_capacity(cloud.n, 3)
    c.upload_cloud(cloud)
    c.upload_views(views, targets)
    out = []
    for it, b in enumerate(batches):
        c.render_forward(b, S.options(sh_degree=sh, deterministic=True))
        c.loss(0.2)
        c.render_backward()
        c.adam_step(S.adam_params(it, 1e-4))
        sg, vc = c.stats()
        m, v = c.moments()
        out.append((c.download_cloud(), sg, vc, m, v))
    return out


def run_steps(c, cloud, views, targets, batches, sh):
    c.set_graph_mode(False)
    c.set_capacity(cloud.n, 3)
    c.upload_cloud(cloud)
    c.upload_views(views, targets)
    out = []
    for it, b in enumerate(batches):
        c.train_step(b, S.options(sh_degree=sh, deterministic=True), 0.2, S.adam_params(it, 1e-4))
        sg, vc = c.stats()
        m, v = c.moments()
        out.append((c.download_cloud(), sg, vc, m, v))
    return out


@pytest.mark.parametrize("sh", [3, 1], ids=["sh3", "sh1_of_3"])
def test_fused_adam_matches_separate_calls(ctx, sh):
    cloud, views, targets, batches = scene()
    batches = batches[:5]
    print(f"[fused] separate calls, sh {sh}", flush=True)
    sep = run_separate(ctx, cloud, views, targets, batches, sh)
    print(f"[fused] training steps, sh {sh}", flush=True)
    fused = run_steps(ctx, cloud, views, targets, batches, sh)
    for it, (a, b) in enumerate(zip(sep, fused)):
        for k in ("pos", "log_scale", "rot", "opacity", "dc", "rest"):
            assert np.array_equal(getattr(a[0], k), getattr(b[0], k)), f"step {it}: {k}"
        for j, name in ((1, "screen gradient"), (2, "visible count"), (3, "m"), (4, "v")):
            assert np.array_equal(np.asarray(a[j]), np.asarray(b[j])), f"step {it}: {name}"


@pytest.mark.parametrize("graph", [False, True], ids=["plain", "graph"])
def test_step_adds_a_stored_gradient(ctx, graph):
    """A standalone adjoint leaves its gradient stored; the next training step
    adds it to its own before Adam (the fused update's general form), exactly
    as the separate calls do, and the step after that starts from a clear
    gradient again (the form that neither reads nor clears it)."""
    cloud, views, targets, batches = scene()
    opts = S.options(sh_degree=3, deterministic=True)

    def start():
        ctx.set_graph_mode(False)
        ctx.set_capacity(cloud.n, 3)
        ctx.upload_cloud(cloud)
        ctx.upload_views(views, targets)

    def snapshot():
        sg, vc = ctx.stats()
        m, v = ctx.moments()
        return ctx.download_cloud(), sg, vc, m, v

    print("[stored] separate calls", flush=True)
    start()
    ctx.render_forward([0], opts)
    ctx.loss(0.2)
    ctx.render_backward()  # stored: view 0's gradient
    sep = []
    for it, b in enumerate(([1], [2], [1], [2])):
        ctx.render_forward(b, opts)
        ctx.loss(0.2)
        ctx.render_backward()
        ctx.adam_step(S.adam_params(it, 1e-4))
        sep.append(snapshot())
    print(f"[stored] training steps, graph={graph}", flush=True)
    start()
    ctx.render_forward([0], opts)
    ctx.loss(0.2)
    ctx.render_backward()
    ctx.set_graph_mode(graph)
    fused = []
    for it, b in enumerate(([1], [2], [1], [2])):
        ctx.train_step(b, opts, 0.2, S.adam_params(it, 1e-4))
        fused.append(snapshot())
    ctx.set_graph_mode(False)
    for it, (a, b) in enumerate(zip(sep, fused)):
        for k in ("pos", "log_scale", "rot", "opacity", "dc", "rest"):
            assert np.array_equal(getattr(a[0], k), getattr(b[0], k)), f"step {it}: {k}"
        for j, name in ((1, "screen gradient"), (2, "visible count"), (3, "m"), (4, "v")):
            assert np.array_equal(np.asarray(a[j]), np.asarray(b[j])), f"step {it}: {name}"
