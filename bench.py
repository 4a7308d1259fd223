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
# HBM bytes per pixel the two loss passes move by design: the statistics pass
# reads rgb and target (24 B) and writes three SSIM partial planes per channel
# (36 B); the gradient pass reads the planes, rgb and target (60 B) and
# writes dL/drgb (12 B).
LOSS_DESIGN_BYTES_PER_PX = 132.0


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
    ap.add_argument("--workload", default="headline", choices=["headline", "c1", "c2", "c3", "c4", "c5", "formats"],
                    help="headline: the benchmark line; c1/c3/c4/c5: the other configurations' measurements; "
                         "formats: seeding from a point cloud and the PLY, .spz and .splat files, against the "
                         "reference")
    ap.add_argument("--train-steps", type=int, default=2000, help="c4: steps of the densifying run")
    ap.add_argument("--grad-quantile", type=float, default=0.8,
                    help="c4: densify threshold = this quantile of the first call's mean view-space gradients "
                         "(negative: the configuration's 2e-4)")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a batch's views would shard data-parallel "
              "across GPUs, with one gradient all-reduce per step); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from parallax_amd import splat as S
    from parallax_amd import synthetic as syn

    ok, name, reason = S.availability()
    if not ok:
        raise SystemExit(f"no usable device: {reason}")
    log("device", name)
    if args.workload != "headline":
        out = {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5,
               "formats": bench_formats}[args.workload](args, S, syn)
        print(json.dumps(out), flush=True)
        return
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
        calibration_copy()
        print(json.dumps({"profile": True, "batched_ms_per_step": round(batched_s * 1e3, 3), "batch_views": B,
                          "calibration_copy_bytes": CALIB_BYTES}))
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
    # HBM bytes per launch of the variant the autotuner kept, from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    rc = cfg_batched
    # the compositing kernels keep the falloff's power <= -80 test only for
    # min_alpha below 1e-30 (third template argument); the adjoint's fourth
    # is deterministic mode
    low_cut = str(not (opts.min_alpha >= 1e-30)).lower()
    det = str(bool(opts.deterministic)).lower()
    variant = (f"k_backward<{rc['backward_batch']},{str(rc['backward_queue']).lower()},{low_cut},{det}>"
               if roofline["kernel"] == "backward" else
               f"k_forward<{['16x16', '8x8', '32x8'].index(rc['forward_shape'])},{str(rc['forward_queue']).lower()},"
               f"{low_cut}>")
    t = profile_traffic(variant)
    if t is not None:
        roofline["traffic"] = t
        roofline["traffic_source"] = f"{PROFILE_FILE} ({variant}, FETCH_SIZE + WRITE_SIZE)"
    # the other compositing kernel's VALU roofline beside it
    valu = {}
    for k, evals, flops, ms in (("backward", e_bwd, FLOPS_BWD_EVAL, bwd_ms), ("forward", e_fwd, FLOPS_FWD_EVAL, fwd_ms)):
        tf = evals * flops / (ms * 1e-3) / 1e12
        valu[k] = {"TFLOP/s": round(tf, 3), "frac": round(tf / PEAK_FP32_TFLOPS, 4), "evals": evals,
                   "flops_per_eval": flops, "ms": round(ms, 4)}
    # HBM-bound kernels: `bytes` is the algorithmic traffic per launch (each
    # input read once, each output written once, the per-unit figures below);
    # `design_bytes` is what the kernel moves by design (intermediates,
    # padding and bookkeeping included); GB/s and frac use `bytes`
    rs = 45
    param_bytes = 56 + 4 * rs  # one SH3 gaussian's parameters
    px_batch = sum(views[i].pixels for i in batches[0])
    alg = {
        "project": n * param_bytes + n * B * 56.0,  # parameters; 52 B record + 4 B tile count per slot
        # the per-gaussian adjoint with Adam fused (the training step's form): record + 10 screen terms per
        # slot; read p, g, m, v and write p, m, v per parameter (the gradient stays in registers)
        "adjoint_adam": n * B * (52 + 40) + 28.0 * n * (14 + rs),
        "loss": px_batch * 36.0,  # rgb and target in, dL/drgb out
        "tile_write": 4.0 * pairs,  # the per-tile lists
    }
    design = {
        # 64 B record + depth key + packed tile rect (views up to 4096 px; 84 B a slot beyond)
        "project": n * param_bytes + n * B * 72.0,
        # 64 B record, 48 B screen; read p, m, v and write them (a training step's stored gradient is known
        # to be zero and neither read nor cleared)
        "adjoint_adam": n * B * (64 + 48) + 24.0 * n * (14 + rs),
        "loss": px_batch * LOSS_DESIGN_BYTES_PER_PX,
        "tile_write": 4.0 * pairs,
    }
    hbm = {}
    for k in alg:
        if k in kernels:
            ms = kernels[k][0] / kernels[k][1]
            gbs = alg[k] / (ms * 1e-3) / 1e9
            hbm[k] = {"GB/s": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes": int(alg[k]),
                      "design_bytes": int(design[k]), "ms": round(ms, 4)}

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
        "valu_kernels": valu,
        "hbm_kernels": hbm,
        "kernel_ms": {k: round(ms / c, 4) for k, (ms, c) in kernels.items()},
        "cpu_baseline": cpu,
    }
    if cpu and cpu.get("value"):
        line["vs_cpu_per_view"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


CALIB_BYTES = 256 << 20
PROFILE_FILE = "profiles/train_step_1M_sh3_1080p_batch8.txt"


def profile_traffic(kernel):
    """HBM bytes per dispatch (calibrated fetch + write) of `kernel` in the
    committed rocprof summary of this workload, or None."""
    try:
        for line in open(os.path.join(ROOT, PROFILE_FILE)):
            f = line.split()
            if len(f) == 6 and f[0] == kernel:
                return int((float(f[4]) + float(f[5])) * 1e6)
    except (OSError, ValueError):
        pass
    return None


def calibration_copy():
    """One plain device-to-device copy of CALIB_BYTES (hipMemcpy between two
    hipMalloc buffers: the runtime's __amd_rocclr_copyBuffer kernel), so a
    profiler run of --profile holds a kernel of known HBM traffic to scale
    FETCH_SIZE and WRITE_SIZE by (parallax_amd/rocprof_summary.py)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(CALIB_BYTES)) != 0:
            raise RuntimeError("hipMalloc failed for the calibration copy")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(CALIB_BYTES)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed for the calibration copy")
        if hip.hipMemcpy(dst, src, C.c_size_t(CALIB_BYTES), 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemcpy (device to device) failed for the calibration copy")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


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




def bench_formats(args, S, syn):
    """The stages either side of training at 1M SH3: seeding from 1M coloured
    points (initializeFromPointCloud), writing / reading the 3DGS PLY
    (236 MB), the .spz and the .splat, on the device against the reference
    engine's own functions on one host core. Times include the host-device
    copies, the .spz deflate / inflate on the host and the file I/O."""
    import ctypes as C
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    n = 1_000_000
    rng = np.random.default_rng(3)
    xyz = (rng.normal(size=(n, 3)) * np.array([2.0, 0.6, 1.5])).astype(np.float32)
    rgb = rng.integers(0, 256, size=(n, 3), dtype=np.uint8)
    cloud = syn.random_cloud(n, 9, 3, (-2, -2, -2), (2, 2, 2))
    ctx = S.Context(0)
    ctx.set_capacity(n, 3)
    out = {"workload": "seeding from 1M coloured points, and the 1M SH3 field as 3DGS PLY (236 MB), .spz and "
                       ".splat; device vs the reference engine on one host core; times include copies, the .spz "
                       "deflate / inflate and file I/O"}

    def best(fn, reps=3):
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            t.append(time.perf_counter() - t0)
        return min(t)

    out["init_ms"] = round(1e3 * best(lambda: ctx.init_from_points(xyz, rgb, 3, 0.1)), 2)
    ctx.upload_cloud(cloud)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "field.ply")
        out["write_ply_ms"] = round(1e3 * best(lambda: ctx.write_ply(path)), 2)
        out["read_ply_ms"] = round(1e3 * best(lambda: ctx.read_ply(path)), 2)
        out["ply_bytes"] = os.path.getsize(path)
        if not args.no_cpu and O.ref_available():
            R = O.ref()
            k = O.ccloud(cloud)
            rs = cloud.rest_stride
            dst = [np.zeros(m * n, np.float32) for m in (3, 3, 4, 1, 3, rs)]
            kc = O.CCloud(n, 3, rs, *[O.ptr(a) for a in dst])
            t0 = time.perf_counter()
            R.ref_init_from_points(n, O.ptr(xyz), rgb.ctypes.data_as(C.POINTER(C.c_ubyte)), 3, 0.1, C.byref(kc))
            out["cpu_reference_init_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            ref_path = os.path.join(d, "ref.ply")
            t0 = time.perf_counter()
            R.ref_write_ply(C.byref(k.struct), ref_path.encode())
            out["cpu_reference_write_ply_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            cnt, deg = C.c_int(), C.c_int()
            err = C.create_string_buffer(256)
            t0 = time.perf_counter()
            R.ref_read_ply(ref_path.encode(), C.byref(cnt), C.byref(deg), C.byref(kc), err, 256)
            out["cpu_reference_read_ply_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            out["files_identical"] = open(ref_path, "rb").read() == open(path, "rb").read()
        # the compact exporters: .spz (quantised on the device, deflated on
        # the host) and .splat (32-byte records, read back at SH degree 0)
        spz = os.path.join(d, "field.spz")
        ctx.upload_cloud(cloud)  # the PLY read renormalised the rotations
        out["write_spz_ms"] = round(1e3 * best(lambda: ctx.write_spz(spz, 12, False)), 2)
        out["read_spz_ms"] = round(1e3 * best(lambda: ctx.read_spz(spz)), 2)
        out["spz_bytes"] = os.path.getsize(spz)
        ctx.upload_cloud(cloud)
        spl = os.path.join(d, "field.splat")
        out["write_splat_ms"] = round(1e3 * best(lambda: ctx.write_splat(spl)), 2)
        ctx0 = S.Context(0)
        ctx0.set_capacity(n, 0)
        out["read_splat_ms"] = round(1e3 * best(lambda: ctx0.read_splat(spl)), 2)
        ctx0.close()
        if not args.no_cpu and O.ref_available():
            ref_spz = os.path.join(d, "ref.spz")
            t0 = time.perf_counter()
            R.ref_write_spz(C.byref(k.struct), ref_spz.encode(), 12, 0)
            out["cpu_reference_write_spz_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            aa = C.c_int()
            t0 = time.perf_counter()
            R.ref_read_spz(ref_spz.encode(), C.byref(cnt), C.byref(deg), C.byref(aa), C.byref(kc), err, 256)
            out["cpu_reference_read_spz_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            out["spz_identical"] = open(ref_spz, "rb").read() == open(spz, "rb").read()
            ref_spl = os.path.join(d, "ref.splat")
            t0 = time.perf_counter()
            R.ref_write_splat(C.byref(k.struct), ref_spl.encode())
            out["cpu_reference_write_splat_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            k0 = O.CCloud(n, 0, 0, *[O.ptr(a) for a in dst])
            t0 = time.perf_counter()
            R.ref_read_splat(ref_spl.encode(), C.byref(cnt), C.byref(deg), C.byref(k0), err, 256)
            out["cpu_reference_read_splat_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            a = np.fromfile(ref_spl, np.uint8).reshape(-1, 32)
            b = np.fromfile(spl, np.uint8).reshape(-1, 32)
            sig = np.abs(a[:, 12:24].copy().view(np.int32).astype(np.int64) -
                         b[:, 12:24].copy().view(np.int32).astype(np.int64))
            out["splat_records_identical_but_sigma"] = bool(np.array_equal(a[:, :12], b[:, :12]) and
                                                            np.array_equal(a[:, 24:], b[:, 24:]))
            out["splat_sigma_max_ulp"] = int(sig.max())
    ctx.close()
    return out


# --------------------------------------------------- the other workloads --

def device_targets(ctx, S, truth, views, deg):
    ctx.set_capacity(truth.n, truth.sh_degree)
    ctx.upload_cloud(truth)
    ctx.upload_views(views)
    out = []
    for k in range(len(views)):
        ctx.render_forward([k], S.options(sh_degree=deg))
        out.append(ctx.image(0)[0])
    return out


def timed_steps(ctx, S, batches, opts, steps, warmup, it0=0):
    for w in range(warmup):
        ctx.train_step(batches[w % len(batches)], opts, 0.2, S.adam_params(it0 + w, 1e-4), want_loss=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.train_step(batches[k % len(batches)], opts, 0.2, S.adam_params(it0 + warmup + k, 1e-4), want_loss=False)
    ctx.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_c1(args, S, syn):
    """The reference's CPU-runnable case: 10k SH0, 256x256, one view; GPU
    step (plain and graph-captured) against the reference trainer's step."""
    cloud, truth, views = syn.config1()
    ctx = S.Context(0)
    tgt = device_targets(ctx, S, truth, views, 0)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, tgt)
    opts = S.options(sh_degree=0)
    plain = timed_steps(ctx, S, [[0]], opts, 200, 10)
    ctx.set_graph_mode(True)
    graph = timed_steps(ctx, S, [[0]], opts, 200, 10)
    ctx.close()
    cpu = None if args.no_cpu else cpu_step(cloud, views[0], tgt[0], 0, samples=20)
    return {"workload": "c1: 10k SH0 gaussians, 256x256, one view per step", "plain_ms": round(plain * 1e3, 4),
            "graph_ms": round(graph * 1e3, 4), "cpu_reference_ms": cpu,
            "graph_vs_cpu": round(cpu / (graph * 1e3), 1) if cpu else None}


def bench_c2(args, S, syn):
    """The single-view gradient-check configuration: 200k SH3 at 1280x720;
    GPU step (plain and graph-captured) against the reference trainer's step
    on the same view and target."""
    cloud, truth, views = syn.config2()
    ctx = S.Context(0)
    tgt = device_targets(ctx, S, truth, views, 3)
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, tgt)
    opts = S.options(sh_degree=3)
    plain = timed_steps(ctx, S, [[0]], opts, 50, 5)
    ctx.set_graph_mode(True)
    graph = timed_steps(ctx, S, [[0]], opts, 50, 5)
    ctx.close()
    cpu = None if args.no_cpu else cpu_step(cloud, views[0], tgt[0], 3, samples=2)
    return {"workload": "c2: 200k SH3 gaussians, 1280x720, one view per step", "plain_ms": round(plain * 1e3, 4),
            "graph_ms": round(graph * 1e3, 4), "cpu_reference_ms": cpu, "cpu_sample": "2 steps of the view",
            "graph_vs_cpu": round(cpu / (graph * 1e3), 1) if cpu else None}


def bench_c3(args, S, syn):
    """500k SH3, 32 orbit views at four resolutions: all 32 in one
    graph-captured step against the per-view launch loop."""
    cloud, truth, views = syn.config3()
    ctx = S.Context(0)
    ctx.set_autotune(not args.no_autotune)
    tgt = device_targets(ctx, S, truth, views, 3)
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, tgt)
    # the CPU leg: one reference step on the first view of each resolution
    # the batch holds, weighted by how many views have it
    sizes = {}
    for k, v in enumerate(views):
        sizes.setdefault((v.width, v.height), []).append(k)
    cpu_views = {wh: (ks[0], len(ks)) for wh, ks in sizes.items()}
    cpu_tgt = {wh: tgt[k] for wh, (k, _) in cpu_views.items()}
    del tgt
    opts = S.options(sh_degree=3)
    everything = [list(range(len(views)))]
    ctx.set_graph_mode(True)
    batched = timed_steps(ctx, S, everything, opts, max(3, args.steps // 3), 2)
    cfg = ctx.raster_config()
    pairs = ctx.last_pair_count
    ctx.set_graph_mode(False)
    singles = [[k] for k in range(len(views))]
    plain = timed_steps(ctx, S, singles, opts, len(views), 4)
    ctx.close()
    px = sum(v.pixels for v in views)
    cpu_view_ms = None
    if not args.no_cpu:
        per_size = {wh: cpu_step(cloud, views[k], cpu_tgt[wh], 3) for wh, (k, _) in cpu_views.items()}
        if all(t is not None for t in per_size.values()):
            cpu_view_ms = round(sum(per_size[wh] * count for wh, (_, count) in cpu_views.items()) / len(views), 1)
    return {"workload": "c3: 500k SH3 gaussians, 32 orbit views of 640x480..1920x1080 in one batch",
            "batched_ms_per_step": round(batched * 1e3, 3), "batched_ms_per_view": round(batched * 1e3 / len(views), 4),
            "plain_ms_per_view": round(plain * 1e3, 4), "batched_vs_plain_per_view": round(plain * len(views) / batched, 3),
            "pixels_per_step": px, "pairs_per_step": pairs, "raster_config": cfg,
            "gaussian_pixels_per_s": float(f"{cloud.n * px / batched:.4g}"),
            "cpu_reference_ms_per_view": cpu_view_ms,
            "cpu_sample": "one reference step per resolution in the batch, weighted by its views",
            "batched_vs_cpu_per_view": round(cpu_view_ms / (batched * 1e3 / len(views)), 1) if cpu_view_ms else None}


def orbit_extent(views):
    """cameraExtent (SplatTrainer.cpp:74-85) in float: 1.1 x the largest
    camera distance from the cameras' centroid, at least 1e-3."""
    f = np.float32
    centres = [(-np.asarray(v.R, np.float64).T @ np.asarray(v.t, np.float64)).astype(np.float32) for v in views]
    centroid = np.zeros(3, np.float32)
    for c in centres:
        centroid = centroid + c
    centroid = centroid / f(len(views))
    furthest = f(0)
    for c in centres:
        d = c - centroid
        furthest = max(furthest, np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
    return float(max(furthest * f(1.1), f(1e-3)))


def bench_c4(args, S, syn):
    """The densifying training run: 300k -> 1M cap, 64-camera 1080p orbit,
    densify every 100 steps (prune opacity 0.005, the engine's other densify
    settings). The engine's 2e-4 view-space gradient threshold admits no
    gaussian of this scene (1080p views spread the loss over 2M pixels), so
    by default the threshold is the given quantile of the first call's
    statistic, fixed for the run (--grad-quantile < 0: the engine's 2e-4).
    Every call's indices are checked against the CPU densify at full size by
    tests/test_gpu_full_configs.py; this times the run."""
    cloud = syn.random_cloud(300_000, 0, 3, (-2, -2, -2), (2, 2, 2))
    truth = syn.random_cloud(1_000_000, 1, 3, (-2, -2, -2), (2, 2, 2))
    views = syn.orbit_views(64, 6.0, [(1920, 1080)], 2)
    cap = 1_000_000
    ctx = S.Context(0)
    ctx.set_autotune(not args.no_autotune)
    tgt = device_targets(ctx, S, truth, views, 3)
    ctx.set_capacity(cap, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, tgt)
    del tgt
    ctx.set_graph_mode(True)
    # SplatTrainingConfig defaults (SplatTrainer.h:25-94), prune opacity 0.005
    dp = S.CDensify()
    dp.densify_grad_threshold = 2e-4
    dp.percent_dense = 0.01
    dp.split_scale_divider = 1.6
    dp.opacity_reset_interval = 3000
    dp.prune_opacity = 0.005
    dp.prune_scale_fraction = 0.1
    dp.max_gaussians = cap
    lr0, lr1, sh_interval = np.float32(1.6e-4), np.float32(1.6e-6), 1000
    extent = orbit_extent(views)
    order = np.random.default_rng(3).permutation(args.train_steps) % len(views)
    rng_state, gpu_s, sizes, counts_all = 7, 0.0, [], []
    quantile = args.grad_quantile if args.grad_quantile is not None and args.grad_quantile >= 0 else None
    ratio = lr1 / lr0
    for it in range(args.train_steps):
        t = np.float32(it) / np.float32(max(1, args.train_steps - 1))
        lr = float(lr0 * np.power(ratio, t, dtype=np.float32) * np.float32(extent))
        t0 = time.perf_counter()
        ctx.train_step([int(order[it])], S.options(sh_degree=min(3, it // sh_interval)), 0.2, S.adam_params(it, lr),
                       want_loss=False)
        if (it + 1) % 100 == 0 and it + 1 < args.train_steps:
            if quantile is not None and it + 1 == 100:
                ctx.synchronize()
                gpu_s += time.perf_counter() - t0
                sg, vc = ctx.stats()
                mean = sg / np.maximum(vc, 1)
                dp.densify_grad_threshold = float(np.quantile(mean, quantile))
                log(f"c4: {int(np.count_nonzero((vc > 0) & (mean >= np.float32(2e-4))))} gaussians at the engine's "
                    f"2e-4; threshold {dp.densify_grad_threshold:.3g} (quantile {quantile})")
                t0 = time.perf_counter()
            _, counts, state = ctx.densify(dp, extent, it + 1, rng_state)
            rng_state = state
            sizes.append(ctx.count)
            counts_all.append([int(c) for c in counts])
            log(f"c4 step {it + 1}: {ctx.count} gaussians (clone {counts[0]}, split {counts[1]}, prune {counts[2]})")
        ctx.synchronize()
        gpu_s += time.perf_counter() - t0
    ctx.close()
    return {"workload": "c4: densifying run, 300k -> 1M SH3, 64-camera 1920x1080 orbit, densify every 100 steps",
            "steps": args.train_steps, "gpu_seconds": round(gpu_s, 3), "steps_per_s": round(args.train_steps / gpu_s, 2),
            "densify_grad_threshold": float(dp.densify_grad_threshold), "threshold_quantile": quantile,
            "gaussians_after_each_densify": sizes, "clone_split_prune_each_densify": counts_all}


def bench_c5(args, S, syn):
    """3M SH3 room, 8 interior 1080p views per batch: deterministic and atomic
    gradient accumulation -- step time, device memory, and run-to-run spread
    of the gradients."""
    cloud, truth, views = syn.config5()
    ctx = S.Context(0)
    ctx.set_autotune(not args.no_autotune)
    tgt = device_targets(ctx, S, truth, views, 3)
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    ctx.upload_views(views, tgt)
    tgt0 = tgt[0]  # the CPU leg's target
    del tgt
    out = {"workload": "c5: 3M SH3 gaussians in a 10x10x3 m room, batch of 8 interior 1920x1080 views"}
    for det in (True, False):
        opts = S.options(sh_degree=3, deterministic=det)
        ctx.set_graph_mode(True)
        ms = timed_steps(ctx, S, [list(range(len(views)))], opts, max(3, args.steps // 2), 2) * 1e3
        ctx.set_graph_mode(False)
        grads = []
        for r in range(2):
            ctx.zero_gradients()
            ctx.render_forward(list(range(len(views))), opts)
            ctx.loss(0.2)
            ctx.render_backward()
            g = ctx.gradients()
            grads.append(np.concatenate([np.asarray(g[k], np.float64).ravel() for k in
                                         ("pos", "log_scale", "rot", "opacity", "dc", "rest")]))
            del g
        spread = float(np.linalg.norm(grads[0] - grads[1]) / max(1e-30, np.linalg.norm(grads[0])))
        mine, used = ctx.memory()
        key = "deterministic" if det else "atomic"
        out[key] = {"ms_per_step": round(ms, 3), "views_per_s": round(len(views) / ms * 1e3, 2),
                    "pairs_per_step": ctx.last_pair_count, "context_hbm_gb": round(mine / 1e9, 3),
                    "device_used_gb": round(used / 1e9, 3), "gradient_rel_l2_run_to_run": spread}
        del grads
        log(f"c5 {key}: {out[key]}")
    ctx.close()
    if not args.no_cpu:
        cpu = cpu_step(cloud, views[0], tgt0, 3)
        out["cpu_reference_ms_per_view"] = cpu
        out["cpu_sample"] = "one reference step on view 0"
        if cpu:
            out["atomic_vs_cpu_per_view"] = round(cpu / (out["atomic"]["ms_per_step"] / len(views)), 1)
        log(f"c5 cpu: {cpu} ms per view")
    return out


def cpu_step(cloud, view, tgt, sh, samples=1):
    """Mean seconds of one reference CPU trainer step (ms)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import oracle as O
    if not O.ref_available():
        return None
    cfg = O.default_config(iterations=1000, sh_degree=sh, sh_increase_interval=0, densify_from_iteration=10**6,
                           densify_until_iteration=10**6, opacity_reset_interval=0)
    R = O.ref()
    k = O.ccloud(cloud)
    varr = (O.CView * 1)(O.cview(view))
    img = np.ascontiguousarray(tgt, np.float32)
    ip = (O.f32p * 1)(O.ptr(img))
    loss = C.c_float()
    secs = R.ref_time_steps(C.byref(k.struct), 1, varr, ip, C.byref(cfg), samples, C.byref(loss))
    return round(secs / samples * 1e3, 3)


if __name__ == "__main__":
    main()
