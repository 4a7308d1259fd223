"""Writes the golden vectors under tests/golden/ from the reference engine's
own splat sources, compiled by `make -C oracle ref` into
oracle/_ref/libparallax_ref.so (its Release flags, -ffast-math included).

Each fixture holds the inputs and the reference's outputs:
  raster_*.npz   projectSplats, renderSplats, photometricLoss (+ dL/drgb) and
                 rasterizeSplatsBackward for one seeded scene and view;
  densify_*.npz  one densifyAndPrune call each (cloud, source_index, counts,
                 RNG): one dominated by pruning, one that reaches the cap;
  trainer.npz    SplatTrainer runs: per-step losses and the final cloud,
                 including one with densification on;
  init_points.npz  initializeFromPointCloud on seeded point sets (with and
                 without colours);
  stage.npz      the splat stage's seed subset and seeded cloud, and its
                 derived training configuration with validate()'s verdict;
  cloud_sh2.ply + ply_sh2.npz  writeGaussianPly's bytes for a seeded SH2
                 cloud, and readGaussianPly's cloud for a hand-laid PLY
                 (other property order, no normals or f_dc, unnormalised and
                 zero quaternions).
Run from the repository root: python tests/golden/make_golden.py
"""
import ctypes as C
import gzip
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from parallax_amd import synthetic as syn  # noqa: E402


def cloud_arrays(prefix, c):
    return {f"{prefix}pos": c.pos, f"{prefix}log_scale": c.log_scale, f"{prefix}rot": c.rot,
            f"{prefix}opacity": c.opacity, f"{prefix}dc": c.dc, f"{prefix}rest": c.rest,
            f"{prefix}sh_degree": np.int32(c.sh_degree)}


def view_arrays(prefix, v):
    return {f"{prefix}R": v.R, f"{prefix}t": v.t,
            f"{prefix}K": np.array([v.fx, v.fy, v.cx, v.cy]), f"{prefix}size": np.array([v.width, v.height], np.int32)}


def ref_raster(name, cloud, view, opt_kw, target, ssim_weight=0.2):
    R = O.ref()
    opt = O.default_options(**opt_kw)
    k = O.ccloud(cloud)
    v = O.cview(view)
    n = max(1, cloud.n)
    out = dict(index=np.zeros(n, np.int32), mean=np.zeros((n, 2), np.float32), cov2d=np.zeros((n, 3), np.float32),
               conic=np.zeros((n, 3), np.float32), depth=np.zeros(n, np.float32), radius=np.zeros(n, np.float32),
               color=np.zeros((n, 3), np.float32), opacity=np.zeros(n, np.float32), clamped=np.zeros((n, 3), np.int32))
    m = R.ref_project(C.byref(k.struct), C.byref(v), C.byref(opt), O.ptr(out["index"], O.i32p), O.ptr(out["mean"]),
                      O.ptr(out["cov2d"]), O.ptr(out["conic"]), O.ptr(out["depth"]), O.ptr(out["radius"]),
                      O.ptr(out["color"]), O.ptr(out["opacity"]), O.ptr(out["clamped"], O.i32p))
    proj = {f"proj_{kk}": vv[:m] for kk, vv in out.items()}
    rgb = np.zeros(view.pixels * 3, np.float32)
    T = np.zeros(view.pixels, np.float32)
    R.ref_render(C.byref(k.struct), C.byref(v), C.byref(opt), O.ptr(rgb), O.ptr(T))
    dl = np.zeros_like(rgb)
    loss = R.ref_photometric_loss(O.ptr(rgb), O.ptr(target), view.width, view.height, float(ssim_weight), O.ptr(dl))
    g = O.Grads(cloud.n, cloud.rest_stride)
    R.ref_backward(C.byref(k.struct), C.byref(v), C.byref(opt), O.ptr(dl), O.ptr(g.pos), O.ptr(g.log_scale),
                   O.ptr(g.rot), O.ptr(g.opacity), O.ptr(g.dc), O.ptr(g.rest), O.ptr(g.screen_gradient),
                   O.ptr(g.visible_count, O.i32p))
    data = {}
    data.update(cloud_arrays("in_", cloud))
    data.update(view_arrays("view_", view))
    data.update({"opt_background": np.array(opt_kw.get("background", (0, 0, 0)), np.float32),
                 "opt_sh_degree": np.int32(opt_kw.get("sh_degree", 0)),
                 "opt_min_alpha": np.float32(opt_kw.get("min_alpha", 1.0 / 255.0)),
                 "opt_cutoff": np.float32(opt_kw.get("transmittance_cutoff", 1e-4)),
                 "target": target, "ssim_weight": np.float32(ssim_weight)})
    data.update(proj)
    data.update({"rgb": rgb, "T": T, "loss": np.float32(loss), "dL": dl})
    data.update({f"grad_{kk}": getattr(g, kk) for kk in ("pos", "log_scale", "rot", "opacity", "dc", "rest",
                                                          "screen_gradient", "visible_count")})
    np.savez_compressed(os.path.join(HERE, f"raster_{name}.npz"), **data)
    print(f"raster_{name}: {m} projected, loss {loss:.6f}")


def ref_image(cloud, view, opt_kw=None):
    """renderSplats of the reference build: the fixtures' input images come
    from the reference itself, so regenerating them never depends on the
    oracle restatement."""
    R = O.ref()
    opt = O.default_options(**(opt_kw or {}))
    k = O.ccloud(cloud)
    v = O.cview(view)
    rgb = np.zeros(view.pixels * 3, np.float32)
    T = np.zeros(view.pixels, np.float32)
    R.ref_render(C.byref(k.struct), C.byref(v), C.byref(opt), O.ptr(rgb), O.ptr(T))
    return rgb


def ref_densify(name, seed, opacity_mu, max_gaussians, iteration):
    R = O.ref()
    rng = np.random.default_rng(seed)
    cloud = syn.random_cloud(400, seed, 1, (-1, -1, -1), (1, 1, 1), log_scale_mu=-3.0, log_scale_sigma=1.0)
    cloud.opacity[:] = rng.normal(opacity_mu, 2.0, size=cloud.n).astype(np.float32)
    sg = rng.exponential(3e-4, size=cloud.n).astype(np.float32)
    vc = rng.integers(0, 4, size=cloud.n).astype(np.int32)
    cfg = O.default_config(max_gaussians=max_gaussians, opacity_reset_interval=3000)
    extent, state0 = 2.0, 7
    cap = 3 * cloud.n
    out = syn.Cloud(np.zeros((cap, 3), np.float32), np.zeros((cap, 3), np.float32), np.zeros((cap, 4), np.float32),
                    np.zeros(cap, np.float32), np.zeros((cap, 3), np.float32),
                    np.zeros((cap, cloud.rest_stride), np.float32), 1)
    ko = O.ccloud(out)
    ki = O.ccloud(cloud)
    src = np.zeros(cap, np.int32)
    counts = np.zeros(3, np.int32)
    st = C.c_uint32(state0)
    n1 = R.ref_densify(C.byref(ki.struct), O.ptr(sg), O.ptr(vc, O.i32p), C.byref(cfg), extent, iteration, C.byref(st),
                       C.byref(ko.struct), O.ptr(src, O.i32p), O.ptr(counts, O.i32p))
    res = {k: a.reshape(b.shape)[:n1] for k, a, b in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"),
                                                       ko.arrays, (out.pos, out.log_scale, out.rot, out.opacity,
                                                                   out.dc, out.rest))}
    data = cloud_arrays("in_", cloud)
    data.update({"screen_gradient": sg, "visible_count": vc, "extent": np.float32(extent),
                 "iteration": np.int32(iteration), "rng_in": np.uint32(state0),
                 "max_gaussians": np.int32(max_gaussians),
                 "rng_out": np.uint32(st.value), "source_index": src[:n1], "counts": counts})
    data.update({f"out_{k}": v for k, v in res.items()})
    np.savez_compressed(os.path.join(HERE, f"densify_{name}.npz"), **data)
    print(f"densify_{name}: {cloud.n} -> {n1}, counts {counts.tolist()}")


def ref_trainer():
    R = O.ref()
    data = {}
    for tag, views, start, cfg_kw, steps in (
        ("plain", syn.ring_of_views(6, 1.6, 32, 32), syn.rough_start(), dict(), 40),
        ("densify", syn.ring_of_views(4, 1.6, 16, 16), syn.rough_start(8),
         dict(densify_from_iteration=20, densify_until_iteration=60, densify_interval=20,
              densify_grad_threshold=0.0), 60),
    ):
        truth = syn.truth_cloud()
        images = [ref_image(truth, v) for v in views]
        cfg = O.default_config(iterations=steps, sh_degree=0, sh_increase_interval=0,
                               densify_from_iteration=steps + 1, densify_until_iteration=steps + 1,
                               opacity_reset_interval=0, ssim_weight=0.2)
        for kk, vv in cfg_kw.items():
            setattr(cfg, kk, vv)
        cap = 4096
        out = syn.Cloud(np.zeros((cap, 3), np.float32), np.zeros((cap, 3), np.float32),
                        np.zeros((cap, 4), np.float32), np.zeros(cap, np.float32), np.zeros((cap, 3), np.float32),
                        np.zeros((cap, 0), np.float32), 0)
        ko = O.ccloud(out)
        ki = O.ccloud(start)
        varr = (O.CView * len(views))(*[O.cview(v) for v in views])
        imgs = [np.ascontiguousarray(im, np.float32) for im in images]
        ip = (O.f32p * len(imgs))(*[O.ptr(im) for im in imgs])
        losses = np.zeros(steps, np.float32)
        n1 = R.ref_train(C.byref(ki.struct), len(views), varr, ip, C.byref(cfg), steps, O.ptr(losses),
                         C.byref(ko.struct), 0)
        data[f"{tag}_losses"] = losses
        data[f"{tag}_images"] = np.stack(imgs)
        data[f"{tag}_count"] = np.int32(n1)
        for k, a, b in zip(("pos", "log_scale", "rot", "opacity", "dc"), ko.arrays,
                           (out.pos, out.log_scale, out.rot, out.opacity, out.dc)):
            data[f"{tag}_out_{k}"] = a.reshape(b.shape)[:n1]
        data[f"{tag}_steps"] = np.int32(steps)
        print(f"trainer {tag}: {steps} steps, {n1} gaussians, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    np.savez_compressed(os.path.join(HERE, "trainer.npz"), **data)


def empty_ccloud(n, rest_stride, sh_degree=0):
    arrs = [np.zeros(k * max(1, n), np.float32) for k in (3, 3, 4, 1, 3, max(1, rest_stride))]
    return O._Keep(O.CCloud(n, sh_degree, rest_stride, *[O.ptr(a) for a in arrs]), arrs)


def ref_init_points():
    R = O.ref()
    data = {}
    for tag, n, seed, deg, opacity, colours in (("rgb_sh1", 3000, 31, 1, 0.1, True),
                                                 ("plain_sh0", 700, 32, 0, 0.35, False),
                                                 ("big_sh3", 8000, 33, 3, 0.1, True)):
        rng = np.random.default_rng(seed)
        xyz = (rng.normal(size=(n, 3)) * np.array([1.5, 0.4, 1.0])).astype(np.float32)
        rgb = rng.integers(0, 256, size=(n, 3), dtype=np.uint8) if colours else None
        rs = 3 * [0, 3, 8, 15][deg]
        out = empty_ccloud(n, rs)
        R.ref_init_from_points(n, O.ptr(xyz), None if rgb is None else rgb.ctypes.data_as(C.POINTER(C.c_ubyte)),
                               deg, opacity, C.byref(out.struct))
        data[f"{tag}_xyz"] = xyz
        if rgb is not None:
            data[f"{tag}_rgb"] = rgb
        data[f"{tag}_args"] = np.array([deg, opacity], np.float64)
        for k, a in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"), out.arrays):
            data[f"{tag}_out_{k}"] = a[: a.size if k != "rest" or rs else 0]
        print(f"init_points {tag}: log sigma {out.arrays[1][0]:.6f}")
    np.savez_compressed(os.path.join(HERE, "init_points.npz"), **data)


STAGE_SEEDS = (("subset_partial_colours", 5000, 3000, 4000, 3, 51),
               ("no_subset", 900, 900, 8000, 1, 52),
               ("subset_no_colours", 2001, 0, 2000, 0, 53))
STAGE_CONFIGS = ((0, 3, 600000), (1, 3, 600000), (7, 3, 600000), (8, 2, 600000), (100, 3, 1000),
                 (7000, 3, 600000), (30001, 1, 250000), (500, 4, 600000), (500, 3, 0))


def ref_stage():
    """The splat stage's seeding (subset of a quarter of the budget, grey
    fill, then initializeFromPointCloud) and its derived training
    configuration with validate()'s verdict (Pipeline.cpp:1860-1876,
    1921-1928)."""
    R = O.ref()
    R.ref_stage_seed.argtypes = [C.c_int, O.f32p, C.c_int, C.POINTER(C.c_ubyte), C.c_int, C.c_int,
                                 C.POINTER(O.CCloud)]
    R.ref_stage_config.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(O.CConfig), C.c_char_p, C.c_int]
    data = {}
    for tag, n, n_colours, max_g, deg, seed in STAGE_SEEDS:
        rng = np.random.default_rng(seed)
        xyz = (rng.normal(size=(n, 3)) * np.array([2.0, 0.5, 1.5])).astype(np.float32)
        rgb = rng.integers(0, 256, size=(max(1, n_colours), 3), dtype=np.uint8)
        rs = 3 * [0, 3, 8, 15][deg]
        out = empty_ccloud(n, rs)
        m = R.ref_stage_seed(n, O.ptr(xyz), n_colours, rgb.ctypes.data_as(C.POINTER(C.c_ubyte)), max_g, deg,
                             C.byref(out.struct))
        data[f"seed_{tag}_xyz"] = xyz
        data[f"seed_{tag}_rgb"] = rgb[:n_colours]
        data[f"seed_{tag}_args"] = np.array([max_g, deg], np.int32)
        for k, a, width in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"), out.arrays,
                               (3, 3, 4, 1, 3, rs)):
            data[f"seed_{tag}_out_{k}"] = a[: m * width]
        print(f"stage seed {tag}: {n} points -> {m} gaussians")
    fields = [f for f, _ in O.CConfig._fields_]
    rows, verdicts = [], []
    for it, deg, max_g in STAGE_CONFIGS:
        cfg = O.CConfig()
        err = C.create_string_buffer(128)
        R.ref_stage_config(it, deg, max_g, C.byref(cfg), err, 128)
        rows.append([float(getattr(cfg, f)) for f in fields])
        verdicts.append(err.value.decode())
    data["config_args"] = np.array(STAGE_CONFIGS, np.int64)
    data["config_fields"] = np.array(fields)
    data["config_values"] = np.array(rows, np.float64)
    data["config_verdicts"] = np.array(verdicts)
    np.savez_compressed(os.path.join(HERE, "stage.npz"), **data)
    print(f"stage configs: {len(rows)}, verdicts {verdicts}")


def ref_ply():
    R = O.ref()
    cloud = syn.random_cloud(500, 37, 2, (-2, -2, -2), (2, 2, 2))
    k = O.ccloud(cloud)
    path = os.path.join(HERE, "cloud_sh2.ply")
    assert R.ref_write_ply(C.byref(k.struct), path.encode()) == 1
    data = {f"in_{n}": a for n, a in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"), k.arrays)}
    # a hand-laid file: other property order, no normals, no f_dc, rotations
    # that need renormalising (one zero)
    rng = np.random.default_rng(38)
    n = 64
    props = ["opacity", "z", "y", "x"] + [f"f_rest_{i}" for i in range(9)] + \
            ["rot_0", "rot_1", "rot_2", "rot_3", "scale_2", "scale_0", "scale_1"]
    rows = rng.normal(size=(n, len(props))).astype(np.float32)
    rows[5, props.index("rot_0"):props.index("rot_3") + 1] = 0.0
    header = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n + \
             "".join(f"property float {p}\n" for p in props) + "end_header\n"
    laid = os.path.join(HERE, "laid_sh1.ply")
    with open(laid, "wb") as f:
        f.write(header.encode() + rows.tobytes())
    cnt, deg = C.c_int(), C.c_int()
    err = C.create_string_buffer(256)
    assert R.ref_read_ply(laid.encode(), C.byref(cnt), C.byref(deg), None, err, 256) == 1
    out = empty_ccloud(cnt.value, 3 * [0, 3, 8, 15][deg.value], deg.value)
    assert R.ref_read_ply(laid.encode(), C.byref(cnt), C.byref(deg), C.byref(out.struct), err, 256) == 1
    data["laid_degree"] = np.int32(deg.value)
    for nm, a in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"), out.arrays):
        data[f"laid_out_{nm}"] = a
    np.savez_compressed(os.path.join(HERE, "ply_sh2.npz"), **data)
    print(f"ply: wrote {os.path.getsize(path)} bytes; read {cnt.value} gaussians of degree {deg.value}")


def compact_cloud(n, seed, deg):
    """A cloud that reaches every branch of the compact writers: positions past
    the 24-bit range and on rounding ties, log scales past both ends of the
    byte, unnormalised quaternions with w < 0 and one zero quaternion, colours
    and SH outside [0, 1] and [-1, 1]."""
    c = syn.random_cloud(n, seed, deg, (-3, -3, -3), (3, 3, 3))
    rng = np.random.default_rng(seed + 1000)
    c.pos[: n // 8] *= np.float32(1000.0)
    c.pos[n // 8: n // 4] = (rng.integers(-20000, 20000, size=(n // 4 - n // 8, 3)) + 0.5).astype(np.float32) / 4096
    c.log_scale[:] = rng.uniform(-12, 7, size=c.log_scale.shape).astype(np.float32)
    c.rot[:] = (rng.normal(size=c.rot.shape) * rng.uniform(0.3, 3, size=(n, 1))).astype(np.float32)
    c.rot[7] = 0.0
    c.opacity[:] = (rng.normal(size=n) * 5).astype(np.float32)
    c.dc[:] = (rng.normal(size=c.dc.shape) * 2.5).astype(np.float32)
    c.rest[:] = (rng.normal(size=c.rest.shape) * 0.7).astype(np.float32)
    return c


def tie_cloud():
    """Gaussians of a 4M-gaussian seeded cloud where the literal quantiser
    expressions and the forms the engine's optimised build evaluates round to
    different bytes: round(255 * (1 / (1 + e))) against round(255 / (1 + e)),
    and round(v * 127.5 + 127.5) against round((v + 1) * 127.5) for the .spz
    rotation; plus 16 ordinary gaussians."""
    f32 = np.float32
    big = syn.random_cloud(4_000_000, 9, 3, (-2, -2, -2), (2, 2, 2))

    def rnd(v):
        v = np.asarray(v, np.float64)
        return np.sign(v) * np.floor(np.abs(v) + 0.5)

    e = np.exp(-big.opacity.astype(np.float64)).astype(f32)
    alpha = rnd((f32(1) / (f32(1) + e)) * f32(255)) != rnd(f32(255) / (f32(1) + e))
    w, x, y, z = (big.rot[:, k] for k in range(4))
    n = np.sqrt((x * x + z * z) + (y * y + w * w)).astype(f32)
    q = big.rot / n[:, None]
    q = np.where(q[:, :1] < 0, -q, q)[:, 1:]
    rot = (rnd(q * f32(127.5) + f32(127.5)) != rnd((q + f32(1)) * f32(127.5))).any(1)
    idx = np.concatenate([np.arange(16), np.nonzero(alpha | rot)[0]])
    print(f"ties: {int(alpha.sum())} opacity and {int(rot.sum())} rotation cases in 4M gaussians")
    return syn.Cloud(big.pos[idx], big.log_scale[idx], big.rot[idx], big.opacity[idx], big.dc[idx],
                     big.rest[idx], 3).contiguous()


def read_ref(fn, path, spz=False):
    cnt, deg, aa = C.c_int(), C.c_int(), C.c_int()
    err = C.create_string_buffer(256)
    head = (path.encode(), C.byref(cnt), C.byref(deg)) + ((C.byref(aa),) if spz else ())
    assert fn(*head, None, err, 256) == 1, err.value
    out = empty_ccloud(cnt.value, 3 * [0, 3, 8, 15][deg.value], deg.value)
    assert fn(*head, C.byref(out.struct), err, 256) == 1
    return out, deg.value, aa.value


def ref_compact():
    """writeSplat / writeSpz of two seeded clouds, and readSplat / readSpz of
    those files and of hand-laid files covering every byte value."""
    R = O.ref()
    data = {}
    for tag, c, bits, aa in (("sh3", compact_cloud(600, 41, 3), 12, 1), ("sh1", compact_cloud(200, 42, 1), 16, 0),
                             ("ties", tie_cloud(), 12, 0)):
        deg = c.sh_degree
        k = O.ccloud(c)
        for nm, a in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"), k.arrays):
            data[f"{tag}_in_{nm}"] = a
        data[f"{tag}_args"] = np.array([deg, bits, aa], np.int32)
        spz = os.path.join(HERE, f"compact_{tag}.spz")
        assert R.ref_write_spz(C.byref(k.struct), spz.encode(), bits, aa) == 1
        out, d, a = read_ref(R.ref_read_spz, spz, True)
        assert (d, a) == (deg, aa)
        for nm, arr in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"), out.arrays):
            data[f"{tag}_spz_out_{nm}"] = arr
        if tag in ("sh3", "ties"):
            sp = os.path.join(HERE, f"compact_{tag}.splat")
            assert R.ref_write_splat(C.byref(k.struct), sp.encode()) == 1
            out, d, _ = read_ref(R.ref_read_splat, sp)
            for nm, arr in zip(("pos", "log_scale", "rot", "opacity", "dc"), out.arrays):
                data[f"{tag}_splat_out_{nm}"] = arr
    # hand-laid .splat: every byte value in every byte field, a zero
    # quaternion, sigmas of 0, denormal, 1e-20 and ordinary sizes
    rng = np.random.default_rng(43)
    n = 512
    rec = np.zeros((n, 32), np.uint8)
    rec[:, :12] = rng.normal(size=(n, 3)).astype(np.float32).view(np.uint8).reshape(n, 12)
    sig = np.exp(rng.uniform(-14, 3, size=(n, 3))).astype(np.float32)
    sig[:4, 0] = np.array([0.0, 1e-40, 1e-20, 1.0], np.float32)
    rec[:, 12:24] = sig.view(np.uint8).reshape(n, 12)
    rec[:, 24:32] = rng.integers(0, 256, size=(n, 8), dtype=np.uint8)
    rec[:256, 24:32] = np.arange(256, dtype=np.uint8)[:, None]
    rec[300, 28:32] = 128
    laid = os.path.join(HERE, "laid.splat")
    rec.tofile(laid)
    out, d, _ = read_ref(R.ref_read_splat, laid)
    for nm, arr in zip(("pos", "log_scale", "rot", "opacity", "dc"), out.arrays):
        data[f"laid_splat_out_{nm}"] = arr
    # hand-laid .spz (SH2, 11 fractional bits, antialiased): every byte value
    # in every plane and a block of rotation bytes around the identity
    n, rpc = 512, 8
    posb = rng.integers(0, 256, size=(n, 9), dtype=np.uint8)
    posb[:256] = np.arange(256, dtype=np.uint8)[:, None]
    planes = [rng.integers(0, 256, size=(n, w), dtype=np.uint8) for w in (3, 3, 1, 3, 3 * rpc)]
    for p in planes:
        p[:256] = np.arange(256, dtype=np.uint8)[:, None]
    g = np.stack(np.meshgrid(*[np.arange(125, 131)] * 3, indexing="ij"), -1).reshape(-1, 3)
    planes[1][256:256 + len(g)] = g
    raw = np.array([0x5053474e, 2, n], np.uint32).tobytes() + bytes([2, 11, 1, 0]) + posb.tobytes() + \
        b"".join(p.tobytes() for p in planes)
    laid = os.path.join(HERE, "laid_sh2.spz")
    with open(laid, "wb") as f:
        f.write(gzip.compress(raw, mtime=0))
    out, d, a = read_ref(R.ref_read_spz, laid, True)
    assert (d, a) == (2, 1)
    for nm, arr in zip(("pos", "log_scale", "rot", "opacity", "dc", "rest"), out.arrays):
        data[f"laid_spz_out_{nm}"] = arr
    np.savez_compressed(os.path.join(HERE, "compact.npz"), **data)
    print("compact: .splat and .spz written and read by the reference")


def main():
    if not O.ref_available():
        raise SystemExit("build the reference first: make -C oracle ref")
    fd_target = np.random.default_rng(41).uniform(size=16 * 16 * 3).astype(np.float32)
    ref_raster("fd16_sh2", syn.overlapping_cloud(2), syn.test_view(16, 16),
               dict(sh_degree=2, background=(0.2, 0.3, 0.4)), fd_target)
    c1 = syn.random_cloud(2000, 0, 0, (-1, -1, 2), (1, 1, 6))
    v1 = syn.pinhole(128, 128, 150.0)
    t1 = ref_image(syn.random_cloud(2000, 1, 0, (-1, -1, 2), (1, 1, 6)), v1)
    ref_raster("c1_small_sh0", c1, v1, dict(), t1)
    c3 = syn.random_cloud(1500, 5, 3, (-1, -1, 2), (1, 1, 6))
    v3 = syn.pinhole(128, 80, 160.0)
    t3 = ref_image(syn.random_cloud(1500, 6, 3, (-1, -1, 2), (1, 1, 6)), v3, dict(sh_degree=3))
    ref_raster("sh3_bg", c3, v3, dict(sh_degree=3, background=(0.1, 0.2, 0.3)), t3)
    ref_densify("prune", 17, -2.0, 560, 3500)
    ref_densify("cap", 23, 1.0, 470, 200)
    ref_trainer()
    ref_init_points()
    ref_stage()
    ref_ply()
    ref_compact()


if __name__ == "__main__":
    main()
