"""Benchmark: Gaussian-splat training steps on one MI355X.

Workload: 1,000,000 SH-degree-3 Gaussians uniform in a 4 m cube, seen by a
64-camera 1920x1080 orbit (radius 6 m, elevation U(+-30 deg), f = 1500), the
end state of the densifying training run. Targets are renders of a second
seeded 1M set. A step is fwd + L1/D-SSIM loss + adjoint + Adam:
  batched: one step over a batch of B views (one launch per stage for the
           whole batch, gradients summed, one Adam);
  plain:   the per-view launch sequence, one view per step.
`value` is training views per second of the batched path (B views per step);
the plain path's rate and the CPU reference trainer's rate on the same
workload go beside it. Inputs are resident in HBM before timing starts.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
PEAK_FP32_TFLOPS = 157.3   # MI355X vector FP32
# Algorithmic work per (pixel, list-entry) evaluation of the compositing
# loops (the engine's SplatRaster.cpp:294-303 and :390-428): FP32 operations
# with the exponential counted as one.
FLOPS_FWD_EVAL = 24
FLOPS_BWD_EVAL = 70


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--deterministic", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="batched leg as a plain launch sequence")
    ap.add_argument("--no-autotune", action="store_true", help="default compositing launch shape")
    ap.add_argument("--profile", action="store_true",
                    help="for rocprofv3: only the batched step, targets uploaded instead of rendered")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (data-parallel sharding by view is "
              "described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from parallax_amd import splat as S
    from parallax_amd import synthetic as syn

    ok, name, reason = S.availability()
    if not ok:
        raise SystemExit(f"no usable device: {reason}")
    log("device", name)
    n = args.gaussians
    B = args.batch
    t0 = time.time()
    cloud, target, views = syn.config4(n=n, views=64)
    log(f"scene generated in {time.time() - t0:.1f}s")
    opts = S.options(sh_degree=3, deterministic=args.deterministic)
    ctx = S.Context(0)
    ctx.set_capacity(n, 3)
    ctx.set_autotune(not args.no_autotune)
    if args.profile:
        # no extra forward launches in the profile: fixed-seed noise targets
        rng = np.random.default_rng(5)
        images = [rng.uniform(size=v.pixels * 3).astype(np.float32) for v in views]
    else:
        # targets: renders of the second seeded set, on the device
        ctx.upload_cloud(target)
        ctx.upload_views(views)
        images = []
        for k in range(len(views)):
            ctx.render_forward([k], S.options(sh_degree=3))
            images.append(ctx.image(0)[0])
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, images)
    del images
    log("targets ready; cloud resident")

    def adam(it):
        return S.adam_params(it, 1.6e-4 * 1.1 * 6.0)  # position lr x extent of a radius-6 rig

    batches = [list(range(i, i + B)) for i in range(0, len(views), B)]
    it = 0

    def run(ids, want_loss=False):
        nonlocal it
        out = ctx.train_step(ids, opts, 0.2, adam(it), want_loss=want_loss)
        it += 1
        return out

    # ---- batched: the whole step replayed from a captured HIP graph ----
    ctx.set_graph_mode(not args.no_graph)
    for w in range(args.warmup):
        run(batches[w % len(batches)])
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        run(batches[s % len(batches)])
    ctx.synchronize()
    batched_s = (time.perf_counter() - t0) / args.steps
    cfg_batched = ctx.raster_config()
    log(f"batched: {batched_s * 1e3:.2f} ms per step of {B} views ({cfg_batched})")
    if args.profile:
        ctx.close()
        print(json.dumps({"profile": True, "batched_ms_per_step": round(batched_s * 1e3, 3), "batch_views": B}))
        return

    # ---- plain per-view launch sequence ----
    ctx.set_graph_mode(False)
    plain_steps = max(args.steps, 2 * B)
    for w in range(args.warmup):
        run([w % len(views)])
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(plain_steps):
        run([s % len(views)])
    ctx.synchronize()
    plain_s = (time.perf_counter() - t0) / plain_steps
    cfg_plain = ctx.raster_config()
    log(f"plain: {plain_s * 1e3:.2f} ms per single-view step ({cfg_plain})")

    # ---- per-kernel times (HIP events on the context's stream) ----
    ctx.set_timing(True)
    ctx.reset_timing()
    prof_steps = 3
    for s in range(prof_steps):
        run(batches[s % len(batches)])
    ctx.synchronize()
    kernels = ctx.timings()
    ctx.set_timing(False)
    for k, (ms, cnt) in sorted(kernels.items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:14s} {ms / cnt:8.3f} ms/launch x{cnt}")

    # algorithmic evaluation counts of the last batch
    ctx.render_forward(batches[0], opts)
    pairs = ctx.last_pair_count
    e_fwd = e_bwd = 0
    for slot in range(B):
        _, T, last = ctx.image(slot)
        e_bwd += int(last.astype(np.int64).sum())
        cols, rows, offsets, _ = ctx.tiles(slot)
        v = views[batches[0][slot]]
        tl = np.diff(offsets).astype(np.int64)
        tile_of_px = (np.arange(v.height)[:, None] // 16) * cols + (np.arange(v.width)[None, :] // 16)
        full = tl[tile_of_px.ravel()]
        stopped = T < 1e-4
        e_fwd += int(np.where(stopped, last, full).sum())
    fwd_ms = kernels["forward"][0] / kernels["forward"][1]
    bwd_ms = kernels["backward"][0] / kernels["backward"][1]
    dominant = max(kernels.items(), key=lambda kv: kv[1][0] / kv[1][1])[0]
    roof_bwd = e_bwd * FLOPS_BWD_EVAL / (bwd_ms * 1e-3) / 1e12
    roof_fwd = e_fwd * FLOPS_FWD_EVAL / (fwd_ms * 1e-3) / 1e12
    roofline = {"kernel": "backward" if bwd_ms >= fwd_ms else "forward", "bound": "valu",
                "achieved": round(roof_bwd if bwd_ms >= fwd_ms else roof_fwd, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": None, "traffic": None,
                "evals_per_launch": e_bwd if bwd_ms >= fwd_ms else e_fwd,
                "flops_per_eval": FLOPS_BWD_EVAL if bwd_ms >= fwd_ms else FLOPS_FWD_EVAL,
                "ms_per_launch": round(bwd_ms if bwd_ms >= fwd_ms else fwd_ms, 4),
                "dominant_by_time": dominant}
    roofline["frac"] = round(roofline["achieved"] / PEAK_FP32_TFLOPS, 4)
    # HBM-bound kernels beside it
    rs = 45
    adam_bytes = 32.0 * n * (14 + rs)
    proj_bytes = n * (56 + 4 * rs) + n * B * 76.0
    hbm = {}
    for k, b in (("adam", adam_bytes), ("project", proj_bytes)):
        if k in kernels:
            ms = kernels[k][0] / kernels[k][1]
            hbm[k] = {"GB/s": round(b / (ms * 1e-3) / 1e9, 1), "frac": round(b / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                      "bytes": int(b), "ms": round(ms, 4)}

    # ---- CPU baseline: the reference trainer, one step on one view ----
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cloud, views[0])
    ctx.close()

    px = sum(v.width * v.height for v in views[:B])
    value = B / batched_s
    line = {
        "metric": "splat train steps/s (fwd+bwd+Adam), 1M Gaussians @1080p; batched vs plain GPU vs CPU",
        "value": round(value, 3),
        "unit": "view-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n} SH3 gaussians, 1920x1080, batch of {B} views from a 64-camera orbit",
                   "gaussians": n, "width": 1920, "height": 1080, "batch_views": B, "sh_degree": 3,
                   "deterministic": bool(args.deterministic), "graph_captured_step": not args.no_graph,
                   "parallelism": "single GPU"},
        "plain_steps_per_s": round(1.0 / plain_s, 3),
        "raster_config": {"batched": cfg_batched, "plain": cfg_plain},
        "batched_vs_plain_per_view": round(plain_s / (batched_s / B), 3),
        "gaussian_pixels_per_s": float(f"{n * px / batched_s:.4g}"),
        "pairs_per_view": int(pairs // B),
        "roofline": roofline,
        "hbm_kernels": hbm,
        "kernel_ms": {k: round(ms / c, 4) for k, (ms, c) in kernels.items()},
        "cpu_baseline": cpu,
    }
    if cpu and cpu.get("value"):
        line["vs_cpu_per_view"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


def cpu_baseline(cloud, view):
    """One step of the reference CPU trainer (oracle/_ref, the engine's own
    SplatRaster/SplatTrainer sources) on view 0 of the same workload,
    single-threaded as the engine runs it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import oracle as O

    kind = "reference" if O.ref_available() else "port"
    from parallax_amd import synthetic as syn
    _, target, _ = syn.config4(n=cloud.n, views=1)
    log("cpu baseline: rendering its target on the host")
    t0 = time.time()
    tgt, _ = O.render(target, view, O.default_options(sh_degree=3))
    log(f"  target {time.time() - t0:.1f}s; timing one {kind} step")
    cfg = O.default_config(iterations=1000, sh_degree=3, sh_increase_interval=0, densify_from_iteration=10**6,
                           densify_until_iteration=10**6, opacity_reset_interval=0)
    if kind == "reference":
        R = O.ref()
        k = O.ccloud(cloud)
        varr = (O.CView * 1)(O.cview(view))
        img = np.ascontiguousarray(tgt, np.float32)
        ip = (O.f32p * 1)(O.ptr(img))
        loss = C.c_float()
        secs = R.ref_time_steps(C.byref(k.struct), 1, varr, ip, C.byref(cfg), 1, C.byref(loss))
    else:
        t = O.Trainer(cloud, [view], [tgt], cfg)
        t0 = time.perf_counter()
        t.step()
        secs = time.perf_counter() - t0
    log(f"  cpu step {secs:.2f}s")
    cpu_name = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(1.0 / secs, 5), "unit": "steps/s", "cores": 1, "kind": kind,
            "sample": f"1 training step (fwd + L1/D-SSIM + bwd + Adam) on one 1920x1080 view of the same "
                      f"{cloud.n}-gaussian SH3 scene, single-threaded; {secs:.2f} s",
            "host_cpu": cpu_name, "host_threads": os.cpu_count()}


if __name__ == "__main__":
    main()
