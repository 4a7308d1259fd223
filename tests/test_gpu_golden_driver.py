"""The facade's training schedule and the splat stage's device half against
the reference build's golden fixtures, through tests/cpp/golden_driver.

* HipSplatTrainer (batch of one view, deterministic, graph-replayed steps) on
  the two SplatTrainer runs of trainer.npz (SplatTrainer.cpp:326-383): 40
  plain steps over 6 views, and 60 steps over 4 views that densify every 20
  from step 20. Per-step losses, the final count, positions and opacities
  are held to the tolerances the CPU oracle is held to on the same fixture;
  the view order (mt19937 + std::shuffle per epoch), the SH unlock, the
  position-rate decay, the EMA and the densify timing all feed them.
* The stage's seed subset and seeded cloud (Pipeline.cpp:1860-1876) on
  stage.npz's point sets: with partial colours, without a subset, without
  colours. Positions and rotations exact; log sigma and DC within 2 ulp, as
  for initializeFromPointCloud alone (the reference build's -ffast-math).
* The stage's derived configuration (Pipeline.cpp:1921-1926): every field
  and validate()'s verdict equal, over nine iteration counts, degrees and
  budgets.
"""
import os
import subprocess

import numpy as np
import pytest

from parallax_amd import synthetic as syn

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "golden_driver")
GOLDEN = os.path.join(HERE, "golden")


def nums(a):
    return " ".join(repr(float(x)) for x in np.asarray(a, np.float64).ravel())


def cloud_text(c):
    rows = np.concatenate([c.pos, c.log_scale, c.rot, c.opacity[:, None], c.dc, c.rest], axis=1)
    return f"cloud {c.n} {c.sh_degree} {c.rest_stride}\n" + "\n".join(nums(r) for r in rows) + "\n"


def parse_cloud(words, at):
    assert words[at] == "cloud"
    n, deg, rs = int(words[at + 1]), int(words[at + 2]), int(words[at + 3])
    width = 14 + rs
    vals = np.array(words[at + 4: at + 4 + n * width], np.float64).astype(np.float32).reshape(n, width)
    return dict(n=n, sh_degree=deg, pos=vals[:, 0:3], log_scale=vals[:, 3:6], rot=vals[:, 6:10],
                opacity=vals[:, 10], dc=vals[:, 11:14], rest=vals[:, 14:])


def run_text(mode, text, tmp_path, timeout=150):
    """The driver's output for `text` as input."""
    assert os.path.exists(BIN), "build it with __graft_entry__.build() (make -C tests/cpp)"
    src, dst = tmp_path / f"{mode}.in", tmp_path / f"{mode}.out"
    src.write_text(text)
    print(f"[golden_driver] {mode}", flush=True)
    r = subprocess.run([BIN, mode, str(src), str(dst)], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return dst.read_text()


def run(mode, text, tmp_path):
    return run_text(mode, text, tmp_path).split()


def ulp(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("tag", ["plain", "densify"])
def test_trainer_schedule_matches_the_reference_trainer(tag, tmp_path):
    d = np.load(os.path.join(GOLDEN, "trainer.npz"))
    steps = int(d[f"{tag}_steps"])
    if tag == "plain":
        views, start = syn.ring_of_views(6, 1.6, 32, 32), syn.rough_start()
        dens = (steps + 1, steps + 1, 100, 2e-4)
    else:
        views, start = syn.ring_of_views(4, 1.6, 16, 16), syn.rough_start(8)
        dens = (20, 60, 20, 0.0)
    # iterations sh_degree sh_increase_interval densify_from densify_until densify_interval threshold
    # opacity_reset_interval ssim_weight max_gaussians deterministic graph_step batch_views
    text = f"config {steps} 0 0 {dens[0]} {dens[1]} {dens[2]} {dens[3]!r} 0 0.2 600000 1 1 1\n"
    text += cloud_text(start)
    text += f"views {len(views)}\n"
    for v, img in zip(views, d[f"{tag}_images"]):
        text += nums(np.asarray(v.R).ravel()) + " " + nums(v.t) + f" {v.fx!r} {v.fy!r} {v.cx!r} {v.cy!r} "
        text += f"{v.width} {v.height}\n" + nums(img) + "\n"
    w = run("trainer", text, tmp_path)
    assert w[0] == "losses"
    k = int(w[1])
    losses = np.array(w[2:2 + k], np.float64)
    at = 2 + k
    assert w[at] == "smoothed" and w[at + 2] == "extent"
    cloud = parse_cloud(w, at + 4)
    print(f"[golden_driver] {tag}: losses {losses[0]:.5f} -> {losses[-1]:.5f} (reference "
          f"{d[f'{tag}_losses'][0]:.5f} -> {d[f'{tag}_losses'][-1]:.5f}), {cloud['n']} gaussians", flush=True)
    assert k == steps
    np.testing.assert_allclose(losses, d[f"{tag}_losses"], rtol=2e-3, atol=1e-5)
    assert cloud["n"] == int(d[f"{tag}_count"])
    np.testing.assert_allclose(cloud["pos"], d[f"{tag}_out_pos"], atol=2e-3)
    np.testing.assert_allclose(cloud["opacity"], d[f"{tag}_out_opacity"], atol=2e-2)


SEEDS = ("subset_partial_colours", "no_subset", "subset_no_colours")


@pytest.mark.parametrize("tag", SEEDS)
def test_stage_seed_matches_the_reference_stage(tag, tmp_path):
    g = np.load(os.path.join(GOLDEN, "stage.npz"))
    xyz, rgb = g[f"seed_{tag}_xyz"], g[f"seed_{tag}_rgb"]
    max_g, deg = (int(x) for x in g[f"seed_{tag}_args"])
    text = f"{max_g} {deg} {len(xyz)} {len(rgb)}\n{nums(xyz)}\n" + " ".join(str(int(x)) for x in rgb.ravel()) + "\n"
    c = parse_cloud(run("stage-seed", text, tmp_path), 0)
    want_pos = g[f"seed_{tag}_out_pos"].reshape(-1, 3)
    assert c["n"] == len(want_pos) and c["sh_degree"] == deg
    assert np.array_equal(c["pos"], want_pos)
    assert np.array_equal(c["rot"], g[f"seed_{tag}_out_rot"].reshape(-1, 4))
    assert np.array_equal(c["opacity"], g[f"seed_{tag}_out_opacity"])
    assert ulp(c["log_scale"], g[f"seed_{tag}_out_log_scale"].reshape(-1, 3)).max() <= 2
    want_dc = g[f"seed_{tag}_out_dc"].reshape(-1, 3)
    assert np.abs(c["dc"] - want_dc).max() <= 2.4e-7  # 2 ulp of the DC range
    assert c["rest"].size == 0 or not c["rest"].any()


def test_stage_config_matches_the_reference_stage(tmp_path):
    g = np.load(os.path.join(GOLDEN, "stage.npz"))
    args = g["config_args"]
    text = f"{len(args)}\n" + "\n".join(" ".join(str(int(x)) for x in a) for a in args) + "\n"
    lines = [ln for ln in run_text("stage-config", text, tmp_path).splitlines() if ln.strip()]
    assert len(lines) == len(args)
    for line, want, verdict, a in zip(lines, g["config_values"], g["config_verdicts"], args):
        head, _, tail = line.partition("|")
        got = np.array(head.split(), np.float64)
        assert np.array_equal(got.astype(np.float32), want.astype(np.float32)), (a, got, want)
        assert tail.rstrip("|") == str(verdict), (a, tail, verdict)
