"""Python binding of the splat trainer's C ABI (include/splat_hip.h).

The library is libparallax_splat.so next to this file, built by
`__graft_entry__.build()` (or `make -C parallax_amd/csrc`). There is no
fallback: if the library is missing or no gfx950 device is visible, the
constructor raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libparallax_splat.so")

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)

OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "no device", 3: "out of memory", 4: "capacity", 5: "device error",
          6: "state"}


class SplatError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class CView(C.Structure):
    _fields_ = [("R", C.c_double * 9), ("t", C.c_double * 3), ("fx", C.c_double), ("fy", C.c_double),
                ("cx", C.c_double), ("cy", C.c_double), ("width", C.c_int32), ("height", C.c_int32)]


class COptions(C.Structure):
    _fields_ = [("background", C.c_float * 3), ("sh_degree", C.c_int32), ("min_alpha", C.c_float),
                ("transmittance_cutoff", C.c_float), ("screen_dilation", C.c_float), ("deterministic", C.c_int32)]


class CCloud(C.Structure):
    _fields_ = [("n", C.c_int32), ("sh_degree", C.c_int32), ("rest_stride", C.c_int32), ("pos", f32p),
                ("log_scale", f32p), ("rot", f32p), ("opacity", f32p), ("dc", f32p), ("rest", f32p)]


class CGrads(C.Structure):
    _fields_ = [("pos", f32p), ("log_scale", f32p), ("rot", f32p), ("opacity", f32p), ("dc", f32p),
                ("rest", f32p), ("screen_gradient", f32p), ("visible_count", i32p)]


class CRasterConfig(C.Structure):
    _fields_ = [("forward_shape", C.c_int32), ("forward_queue", C.c_int32), ("backward_batch", C.c_int32),
                ("backward_queue", C.c_int32)]


class CAdam(C.Structure):
    _fields_ = [("lr_position", C.c_float), ("lr_log_scale", C.c_float), ("lr_rotation", C.c_float),
                ("lr_opacity", C.c_float), ("lr_dc", C.c_float), ("lr_rest", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("epsilon", C.c_float), ("correction1", C.c_float),
                ("correction2", C.c_float)]


class CDensify(C.Structure):
    _fields_ = [("densify_grad_threshold", C.c_float), ("percent_dense", C.c_float),
                ("split_scale_divider", C.c_float), ("opacity_reset_interval", C.c_int32),
                ("prune_opacity", C.c_float), ("prune_scale_fraction", C.c_float), ("max_gaussians", C.c_int32)]


_LIB = None

# Every entry point the header declares, with its ctypes signature.
SIGNATURES = {
    "splat_availability": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]),
    "splat_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "splat_destroy": (None, [C.c_void_p]),
    "splat_last_error": (C.c_char_p, [C.c_void_p]),
    "splat_default_options": (None, [C.POINTER(COptions)]),
    "splat_set_capacity": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "splat_upload_cloud": (C.c_int, [C.c_void_p, C.POINTER(CCloud)]),
    "splat_download_cloud": (C.c_int, [C.c_void_p, C.POINTER(CCloud)]),
    "splat_count": (C.c_int32, [C.c_void_p]),
    "splat_upload_views": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(CView), C.POINTER(f32p)]),
    "splat_render_forward": (C.c_int, [C.c_void_p, C.c_int32, i32p, C.POINTER(COptions)]),
    "splat_download_image": (C.c_int, [C.c_void_p, C.c_int32, f32p, f32p, i32p]),
    "splat_download_projection": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                            i32p, i32p]),
    "splat_download_tiles": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, i32p]),
    "splat_loss": (C.c_int, [C.c_void_p, C.c_float, f32p]),
    "splat_loss_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p,
                                     C.c_void_p]),
    "splat_set_image_gradient": (C.c_int, [C.c_void_p, C.c_int32, f32p]),
    "splat_download_image_gradient": (C.c_int, [C.c_void_p, C.c_int32, f32p]),
    "splat_render_backward": (C.c_int, [C.c_void_p]),
    "splat_download_gradients": (C.c_int, [C.c_void_p, C.POINTER(CGrads)]),
    "splat_zero_gradients": (C.c_int, [C.c_void_p]),
    "splat_adam_step": (C.c_int, [C.c_void_p, C.POINTER(CAdam)]),
    "splat_train_step": (C.c_int, [C.c_void_p, C.c_int32, i32p, C.POINTER(COptions), C.c_float, C.POINTER(CAdam),
                                   f32p]),
    "splat_densify": (C.c_int, [C.c_void_p, C.POINTER(CDensify), C.c_float, C.c_int32, C.POINTER(C.c_uint32), i32p,
                                i32p]),
    "splat_set_graph_mode": (C.c_int, [C.c_void_p, C.c_int32]),
    "splat_set_autotune": (C.c_int, [C.c_void_p, C.c_int32]),
    "splat_set_raster_config": (C.c_int, [C.c_void_p, C.c_void_p]),
    "splat_get_raster_config": (C.c_int, [C.c_void_p, C.c_void_p]),
    "splat_reset_opacity": (C.c_int, [C.c_void_p, C.c_float]),
    "splat_init_from_points": (C.c_int, [C.c_void_p, C.c_int32, f32p, C.POINTER(C.c_uint8), C.c_int32, C.c_float,
                                         C.POINTER(C.c_float)]),
    "splat_write_ply": (C.c_int, [C.c_void_p, C.c_char_p]),
    "splat_ply_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]),
    "splat_read_ply": (C.c_int, [C.c_void_p, C.c_char_p]),
    "splat_write_splat": (C.c_int, [C.c_void_p, C.c_char_p]),
    "splat_splat_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]),
    "splat_read_splat": (C.c_int, [C.c_void_p, C.c_char_p]),
    "splat_write_spz": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32]),
    "splat_spz_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.c_char_p, C.c_size_t]),
    "splat_read_spz": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int32)]),
    "splat_download_stats": (C.c_int, [C.c_void_p, f32p, i32p]),
    "splat_prune_and_sort": (C.c_int, [C.c_void_p, C.c_float]),
    "splat_upload_stats": (C.c_int, [C.c_void_p, f32p, i32p]),
    "splat_download_moments": (C.c_int, [C.c_void_p, f32p, f32p]),
    "splat_upload_moments": (C.c_int, [C.c_void_p, f32p, f32p]),
    "splat_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "splat_timing_count": (C.c_int32, [C.c_void_p]),
    "splat_timing_entry": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int64)]),
    "splat_reset_timing": (C.c_int, [C.c_void_p]),
    "splat_last_pair_count": (C.c_int64, [C.c_void_p]),
    "splat_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "splat_synchronize": (C.c_int, [C.c_void_p]),
}


def lib():
    """Load libparallax_splat.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or make -C parallax_amd/csrc")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def availability():
    name = C.create_string_buffer(256)
    reason = C.create_string_buffer(256)
    st = lib().splat_availability(name, 256, reason, 256)
    return st == OK, name.value.decode(), reason.value.decode()


def _p(a, t=f32p):
    return None if a is None else a.ctypes.data_as(t)


def cview(v) -> CView:
    s = CView()
    s.R[:] = [float(x) for x in np.asarray(v.R, dtype=np.float64).ravel()]
    s.t[:] = [float(x) for x in np.asarray(v.t, dtype=np.float64).ravel()]
    s.fx, s.fy, s.cx, s.cy = float(v.fx), float(v.fy), float(v.cx), float(v.cy)
    s.width, s.height = int(v.width), int(v.height)
    return s


def options(sh_degree=0, background=(0.0, 0.0, 0.0), min_alpha=1.0 / 255.0, transmittance_cutoff=1e-4,
            screen_dilation=0.3, deterministic=False) -> COptions:
    o = COptions()
    o.background[:] = [float(b) for b in background]
    o.sh_degree = int(sh_degree)
    o.min_alpha = float(np.float32(min_alpha))
    o.transmittance_cutoff = float(transmittance_cutoff)
    o.screen_dilation = float(screen_dilation)
    o.deterministic = 1 if deterministic else 0
    return o


def adam_params(iteration: int, lr_position: float, scale_lr=5e-3, rotation_lr=1e-3, opacity_lr=0.05,
                dc_lr=2.5e-3, rest_lr=None, beta1=0.9, beta2=0.999, eps=1e-15) -> CAdam:
    """Adam scalars for 0-based step `iteration`, bias corrections as the
    engine computes them in float (SplatTrainer.cpp:219-222)."""
    f = np.float32
    t = f(iteration + 1)
    a = CAdam()
    a.lr_position = float(f(lr_position))
    a.lr_log_scale = float(f(scale_lr))
    a.lr_rotation = float(f(rotation_lr))
    a.lr_opacity = float(f(opacity_lr))
    a.lr_dc = float(f(dc_lr))
    a.lr_rest = float(f(dc_lr) / f(20.0)) if rest_lr is None else float(f(rest_lr))
    a.beta1, a.beta2, a.epsilon = float(f(beta1)), float(f(beta2)), float(f(eps))
    a.correction1 = float(f(1.0) - np.power(f(beta1), t, dtype=np.float32))
    a.correction2 = float(f(1.0) - np.power(f(beta2), t, dtype=np.float32))
    return a


class Context:
    """One device context: the cloud, the cameras and the current batch."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        st = L.splat_create(int(device), C.byref(h))
        if st != OK:
            raise SplatError(st, L.splat_last_error(None).decode())
        self.h = h
        self.views = []
        self.rest_stride = 0
        self.sh_degree = 0
        self._batch = []

    def close(self):
        if getattr(self, "h", None):
            lib().splat_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, st):
        if st != OK:
            raise SplatError(st, lib().splat_last_error(self.h).decode())

    # ---- scene ----
    def set_capacity(self, max_gaussians: int, sh_degree: int):
        self._ck(lib().splat_set_capacity(self.h, int(max_gaussians), int(sh_degree)))
        self.sh_degree = int(sh_degree)
        self.rest_stride = 3 * {0: 0, 1: 3, 2: 8, 3: 15}[int(sh_degree)]

    def upload_cloud(self, cloud):
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in
                (cloud.pos, cloud.log_scale, cloud.rot, cloud.opacity, cloud.dc, cloud.rest)]
        s = CCloud(cloud.n, cloud.sh_degree, cloud.rest_stride, *[_p(a) for a in arrs])
        self._ck(lib().splat_upload_cloud(self.h, C.byref(s)))

    @property
    def count(self) -> int:
        return int(lib().splat_count(self.h))

    def download_cloud(self):
        from parallax_amd.synthetic import Cloud
        n, rs = self.count, self.rest_stride
        c = Cloud(np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32),
                  np.zeros(n, np.float32), np.zeros((n, 3), np.float32), np.zeros((n, rs), np.float32), 0)
        s = CCloud(n, 0, rs, _p(c.pos), _p(c.log_scale), _p(c.rot), _p(c.opacity), _p(c.dc), _p(c.rest))
        self._ck(lib().splat_download_cloud(self.h, C.byref(s)))
        c.sh_degree = int(s.sh_degree)
        return c

    def upload_views(self, views, images=None):
        self.views = list(views)
        arr = (CView * max(1, len(views)))(*[cview(v) for v in views])
        self._imgs = None
        ip = None
        if images is not None:
            self._imgs = [None if im is None else np.ascontiguousarray(im, dtype=np.float32).ravel() for im in images]
            ip = (f32p * len(self._imgs))(*[_p(im) if im is not None else None for im in self._imgs])
        self._ck(lib().splat_upload_views(self.h, len(views), arr, ip))

    # ---- forward / backward ----
    def render_forward(self, view_ids, opts: COptions):
        ids = np.ascontiguousarray(view_ids, dtype=np.int32)
        self._batch = [int(i) for i in ids]
        self._ck(lib().splat_render_forward(self.h, len(ids), _p(ids, i32p), C.byref(opts)))

    def image(self, slot: int = 0):
        v = self.views[self._batch[slot]]
        px = v.width * v.height
        rgb = np.zeros(px * 3, np.float32)
        T = np.zeros(px, np.float32)
        last = np.zeros(px, np.int32)
        self._ck(lib().splat_download_image(self.h, int(slot), _p(rgb), _p(T), _p(last, i32p)))
        return rgb, T, last

    def projection(self, slot: int = 0) -> dict:
        n = max(1, self.count)
        out = dict(index=np.zeros(n, np.int32), mean=np.zeros((n, 2), np.float32), conic=np.zeros((n, 3), np.float32),
                   depth=np.zeros(n, np.float32), radius=np.zeros(n, np.float32), color=np.zeros((n, 3), np.float32),
                   opacity=np.zeros(n, np.float32), clamped=np.zeros((n, 3), np.int32),
                   rect=np.zeros((n, 4), np.int32))
        cnt = C.c_int32()
        self._ck(lib().splat_download_projection(
            self.h, int(slot), C.byref(cnt), _p(out["index"], i32p), _p(out["mean"]), _p(out["conic"]),
            _p(out["depth"]), _p(out["radius"]), _p(out["color"]), _p(out["opacity"]), _p(out["clamped"], i32p),
            _p(out["rect"], i32p)))
        k = cnt.value
        return {key: val[:k].copy() for key, val in out.items()}

    def tiles(self, slot: int = 0):
        v = self.views[self._batch[slot]]
        cols, rows = (v.width + 15) // 16, (v.height + 15) // 16
        cnt = C.c_int32()
        offsets = np.zeros(cols * rows + 1, np.int32)
        self._ck(lib().splat_download_tiles(self.h, int(slot), _p(offsets, i32p), None, C.byref(cnt)))
        entries = np.zeros(max(1, cnt.value), np.int32)
        self._ck(lib().splat_download_tiles(self.h, int(slot), _p(offsets, i32p), _p(entries, i32p),
                                            C.byref(cnt)))
        return cols, rows, offsets, entries[:cnt.value].copy()

    def loss(self, ssim_weight: float):
        out = np.zeros(len(self._batch), np.float32)
        self._ck(lib().splat_loss(self.h, float(ssim_weight), _p(out)))
        return out

    def loss_noalloc(self, rgb_dev: int, target_dev: int, width: int, height: int, ssim_weight: float,
                     dL_dev: int, loss_dev: int):
        """The loss of one image in device memory (addresses as ints), with
        the library's fixed scratch: no allocation; asynchronous on the
        context's stream (call synchronize() before reading the outputs)."""
        self._ck(lib().splat_loss_noalloc(self.h, C.c_void_p(rgb_dev), C.c_void_p(target_dev), int(width),
                                          int(height), float(ssim_weight), C.c_void_p(dL_dev),
                                          C.c_void_p(loss_dev)))

    def set_image_gradient(self, slot: int, dL_drgb):
        d = np.ascontiguousarray(dL_drgb, dtype=np.float32)
        self._ck(lib().splat_set_image_gradient(self.h, int(slot), _p(d)))

    def image_gradient(self, slot: int = 0):
        v = self.views[self._batch[slot]]
        d = np.zeros(v.width * v.height * 3, np.float32)
        self._ck(lib().splat_download_image_gradient(self.h, int(slot), _p(d)))
        return d

    def render_backward(self):
        self._ck(lib().splat_render_backward(self.h))

    def gradients(self):
        n, rs = self.count, self.rest_stride
        g = dict(pos=np.zeros((n, 3), np.float32), log_scale=np.zeros((n, 3), np.float32),
                 rot=np.zeros((n, 4), np.float32), opacity=np.zeros(n, np.float32), dc=np.zeros((n, 3), np.float32),
                 rest=np.zeros((n, rs), np.float32), screen_gradient=np.zeros(n, np.float32),
                 visible_count=np.zeros(n, np.int32))
        s = CGrads(_p(g["pos"]), _p(g["log_scale"]), _p(g["rot"]), _p(g["opacity"]), _p(g["dc"]), _p(g["rest"]),
                   _p(g["screen_gradient"]), _p(g["visible_count"], i32p))
        self._ck(lib().splat_download_gradients(self.h, C.byref(s)))
        return g

    def zero_gradients(self):
        self._ck(lib().splat_zero_gradients(self.h))

    # ---- optimiser / fused ----
    def adam_step(self, params: CAdam):
        self._ck(lib().splat_adam_step(self.h, C.byref(params)))

    def train_step(self, view_ids, opts: COptions, ssim_weight: float, adam: CAdam, want_loss=True):
        ids = np.ascontiguousarray(view_ids, dtype=np.int32)
        self._batch = [int(i) for i in ids]
        out = np.zeros(len(ids), np.float32) if want_loss else None
        self._ck(lib().splat_train_step(self.h, len(ids), _p(ids, i32p), C.byref(opts), float(ssim_weight),
                                        C.byref(adam), _p(out)))
        return out

    def densify(self, params: CDensify, extent: float, iteration: int, rng_state: int):
        n_max = max(1, self.count * 3 + 8)
        src = np.zeros(n_max, np.int32)
        counts = np.zeros(3, np.int32)
        st = C.c_uint32(rng_state)
        self._ck(lib().splat_densify(self.h, C.byref(params), float(extent), int(iteration), C.byref(st),
                                     _p(src, i32p), _p(counts, i32p)))
        return src[:self.count].copy(), counts, int(st.value)

    def reset_opacity(self, ceiling: float):
        self._ck(lib().splat_reset_opacity(self.h, float(ceiling)))

    def init_from_points(self, xyz, rgb=None, sh_degree: int = 0, initial_opacity: float = 0.1) -> float:
        """initializeFromPointCloud on the device (capacity already set);
        returns the shared log sigma."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        col = None if rgb is None else np.ascontiguousarray(rgb, np.uint8).reshape(-1, 3)
        ls = C.c_float()
        self._ck(lib().splat_init_from_points(self.h, len(xyz), _p(xyz),
                                              None if col is None else col.ctypes.data_as(C.POINTER(C.c_uint8)),
                                              int(sh_degree), float(initial_opacity), C.byref(ls)))
        return ls.value

    def write_ply(self, path: str):
        self._ck(lib().splat_write_ply(self.h, os.fsencode(path)))

    def read_ply(self, path: str):
        self._ck(lib().splat_read_ply(self.h, os.fsencode(path)))

    def write_splat(self, path: str):
        self._ck(lib().splat_write_splat(self.h, os.fsencode(path)))

    def read_splat(self, path: str):
        self._ck(lib().splat_read_splat(self.h, os.fsencode(path)))

    def write_spz(self, path: str, fractional_bits: int = 12, antialiased: bool = False):
        self._ck(lib().splat_write_spz(self.h, os.fsencode(path), int(fractional_bits), int(bool(antialiased))))

    def read_spz(self, path: str) -> bool:
        """Reads the file into the context; returns its antialiased flag."""
        aa = C.c_int32()
        self._ck(lib().splat_read_spz(self.h, os.fsencode(path), C.byref(aa)))
        return bool(aa.value)


    def stats(self):
        n = self.count
        sg = np.zeros(n, np.float32)
        vc = np.zeros(n, np.int32)
        self._ck(lib().splat_download_stats(self.h, _p(sg), _p(vc, i32p)))
        return sg, vc

    def upload_stats(self, screen_gradient, visible_count):
        sg = np.ascontiguousarray(screen_gradient, dtype=np.float32)
        vc = np.ascontiguousarray(visible_count, dtype=np.int32)
        self._ck(lib().splat_upload_stats(self.h, _p(sg), _p(vc, i32p)))

    def moments(self):
        k = self.count * (14 + self.rest_stride)
        m = np.zeros(k, np.float32)
        v = np.zeros(k, np.float32)
        self._ck(lib().splat_download_moments(self.h, _p(m), _p(v)))
        return m, v

    def upload_moments(self, m, v):
        m = np.ascontiguousarray(m, dtype=np.float32)
        v = np.ascontiguousarray(v, dtype=np.float32)
        self._ck(lib().splat_upload_moments(self.h, _p(m), _p(v)))

    def prune_and_sort(self, min_opacity: float):
        self._ck(lib().splat_prune_and_sort(self.h, float(min_opacity)))

    # ---- measurement ----
    def set_timing(self, on: bool):
        self._ck(lib().splat_set_timing(self.h, 1 if on else 0))

    def set_autotune(self, on: bool):
        self._ck(lib().splat_set_autotune(self.h, 1 if on else 0))

    def set_raster_config(self, forward_shape=None, forward_queue=False, backward_batch=128, backward_queue=False):
        """Force a compositing launch shape; forward_shape None: back to default/tuned."""
        if forward_shape is None:
            self._ck(lib().splat_set_raster_config(self.h, None))
            return
        c = CRasterConfig(int(forward_shape), 1 if forward_queue else 0, int(backward_batch),
                          1 if backward_queue else 0)
        self._ck(lib().splat_set_raster_config(self.h, C.byref(c)))

    def raster_config(self) -> dict:
        c = CRasterConfig()
        self._ck(lib().splat_get_raster_config(self.h, C.byref(c)))
        return {"forward_shape": ("16x16", "8x8", "32x8")[c.forward_shape], "forward_queue": bool(c.forward_queue),
                "backward_batch": int(c.backward_batch), "backward_queue": bool(c.backward_queue)}

    def set_graph_mode(self, on: bool):
        """train_step replays a captured HIP graph per batch shape."""
        self._ck(lib().splat_set_graph_mode(self.h, 1 if on else 0))

    def timings(self) -> dict:
        out = {}
        for i in range(int(lib().splat_timing_count(self.h))):
            name = C.c_char_p()
            ms = C.c_double()
            n = C.c_int64()
            self._ck(lib().splat_timing_entry(self.h, i, C.byref(name), C.byref(ms), C.byref(n)))
            out[name.value.decode()] = (ms.value, n.value)
        return out

    def reset_timing(self):
        self._ck(lib().splat_reset_timing(self.h))

    def memory(self):
        """(bytes the context holds, bytes used on the device)."""
        mine, used = C.c_uint64(), C.c_uint64()
        self._ck(lib().splat_memory(self.h, C.byref(mine), C.byref(used)))
        return int(mine.value), int(used.value)

    @property
    def last_pair_count(self) -> int:
        return int(lib().splat_last_pair_count(self.h))

    def synchronize(self):
        self._ck(lib().splat_synchronize(self.h))


def ply_info(path: str):
    """(count, sh_degree) of a 3DGS PLY file; raises SplatError with the
    engine's message when the file is not one."""
    n, d = C.c_int32(), C.c_int32()
    err = C.create_string_buffer(256)
    st = lib().splat_ply_info(os.fsencode(path), C.byref(n), C.byref(d), err, 256)
    if st != OK:
        raise SplatError(st, err.value.decode())
    return n.value, d.value


def splat_info(path: str) -> int:
    """The record count of a .splat file (SH degree 0); raises SplatError
    with the engine's message when the file is not one."""
    n = C.c_int32()
    err = C.create_string_buffer(256)
    st = lib().splat_splat_info(os.fsencode(path), C.byref(n), err, 256)
    if st != OK:
        raise SplatError(st, err.value.decode())
    return n.value


def spz_info(path: str):
    """(count, sh_degree, antialiased) of a .spz file; raises SplatError with
    the engine's message when the file is not one."""
    n, d, aa = C.c_int32(), C.c_int32(), C.c_int32()
    err = C.create_string_buffer(256)
    st = lib().splat_spz_info(os.fsencode(path), C.byref(n), C.byref(d), C.byref(aa), err, 256)
    if st != OK:
        raise SplatError(st, err.value.decode())
    return n.value, d.value, bool(aa.value)
