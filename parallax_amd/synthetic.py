"""Seeded synthetic scenes and cameras for the splat trainer.

Camera conventions follow the engine: a view holds the camera-from-world pose
(p_cam = R p + t), camera +X right, +Y down, +Z forward, pixel centres at
x + 0.5 (engine/native/cpp/src/testing/SyntheticScene.cpp:235-259,
engine/native/cpp/src/splat/SplatRaster.cpp:288). Gaussians use the INRIA
storage convention of `Gaussian` (engine/native/cpp/include/parallax/splat/
GaussianCloud.h:41-51): log-scales, opacity logits, unit quaternions stored
(w, x, y, z), a DC coefficient and channel-major higher-order SH.

`Rng` is the engine's deterministic xorshift64 test generator
(engine/native/cpp/include/parallax/testing/SyntheticScene.h:80-107), used for
the small fixtures that mirror the engine's own tests. Large scenes use
numpy's PCG64 with a fixed seed.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SH_C0 = 0.28209479177387814


def sh_rest_per_channel(degree: int) -> int:
    """shRestPerChannel (GaussianCloud.h:16-27)."""
    return {0: 0, 1: 3, 2: 8, 3: 15}.get(int(degree), 0)


@dataclasses.dataclass
class Cloud:
    """A Gaussian cloud in structure-of-arrays form (float32)."""

    pos: np.ndarray        # (n, 3)
    log_scale: np.ndarray  # (n, 3)
    rot: np.ndarray        # (n, 4) w, x, y, z
    opacity: np.ndarray    # (n,) logits
    dc: np.ndarray         # (n, 3)
    rest: np.ndarray       # (n, rest_stride), channel-major
    sh_degree: int = 0

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def rest_stride(self) -> int:
        return int(self.rest.shape[1])

    def copy(self) -> "Cloud":
        return Cloud(self.pos.copy(), self.log_scale.copy(), self.rot.copy(), self.opacity.copy(),
                     self.dc.copy(), self.rest.copy(), self.sh_degree)

    def contiguous(self) -> "Cloud":
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        return Cloud(f(self.pos), f(self.log_scale), f(self.rot), f(self.opacity), f(self.dc),
                     f(self.rest), self.sh_degree)

    def subset(self, idx) -> "Cloud":
        return Cloud(self.pos[idx], self.log_scale[idx], self.rot[idx], self.opacity[idx],
                     self.dc[idx], self.rest[idx], self.sh_degree).contiguous()

    def flat(self) -> np.ndarray:
        """All parameters, group-concatenated [pos|log_scale|rot|opacity|dc|rest]."""
        return np.concatenate([self.pos.ravel(), self.log_scale.ravel(), self.rot.ravel(),
                               self.opacity.ravel(), self.dc.ravel(), self.rest.ravel()])


@dataclasses.dataclass
class View:
    """SplatView: pinhole intrinsics plus the camera-from-world pose."""

    R: np.ndarray  # (3, 3) float64, row-major
    t: np.ndarray  # (3,) float64
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int

    def center(self) -> np.ndarray:
        return -self.R.T @ self.t

    @property
    def pixels(self) -> int:
        return self.width * self.height


class Rng:
    """The engine's xorshift64 test generator (SyntheticScene.h:80-107)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int = 88172645463325252):
        self.state = seed if seed else 1

    def next(self) -> int:
        s = self.state
        s ^= (s << 13) & self.MASK
        s ^= s >> 7
        s ^= (s << 17) & self.MASK
        self.state = s
        return s

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self) -> float:
        u1 = max(1e-12, self.uniform())
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


# ------------------------------------------------------------------ cameras --

def pinhole(width: int, height: int, f: float, R=None, t=None) -> View:
    R = np.eye(3) if R is None else np.asarray(R, dtype=np.float64)
    t = np.zeros(3) if t is None else np.asarray(t, dtype=np.float64)
    return View(R, t, float(f), float(f), 0.5 * width, 0.5 * height, int(width), int(height))


def test_view(width: int = 48, height: int = 36, distance: float = 2.0) -> View:
    """makeView of the engine's rasteriser tests (SplatRasterTest.cpp:15-25)."""
    return View(np.eye(3), np.array([0.0, 0.0, distance]), 0.9 * width, 0.9 * width,
                0.5 * width, 0.5 * height, width, height)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """SyntheticScene::lookAt (SyntheticScene.cpp:235-259): rows are the
    camera's right, down and forward axes in world coordinates."""
    eye = np.asarray(eye, dtype=np.float64)
    forward = np.asarray(target, dtype=np.float64) - eye
    if np.linalg.norm(forward) < 1e-9:
        forward = np.array([0.0, 0.0, 1.0])
    forward = forward / np.linalg.norm(forward)
    world_up = np.asarray(up, dtype=np.float64)
    if abs(np.dot(world_up / np.linalg.norm(world_up), forward)) > 0.999:
        world_up = np.array([1.0, 0.0, 0.0]) if abs(forward[0]) < 0.9 else np.array([0.0, 0.0, 1.0])
    right = np.cross(forward, world_up / np.linalg.norm(world_up))
    if np.linalg.norm(right) < 1e-9:
        right = np.array([1.0, 0.0, 0.0])
    right = right / np.linalg.norm(right)
    down = np.cross(forward, right)
    R = np.stack([right, down, forward])
    u, _, vt = np.linalg.svd(R)  # orthonormalize
    Ro = u @ vt
    if np.linalg.det(Ro) < 0:
        u[:, -1] *= -1
        Ro = u @ vt
    return Ro, -R @ eye


def ring_of_views(count: int, radius: float, width: int, height: int):
    """ringOfViews of the engine's trainer tests (SplatTrainerTest.cpp:15-38)."""
    views = []
    for i in range(count):
        a = 2.0 * math.pi * i / count
        eye = np.array([radius * math.sin(a), 0.0, -radius * math.cos(a)])
        forward = -eye / np.linalg.norm(eye)
        right = np.cross(np.array([0.0, 1.0, 0.0]), forward)
        right /= np.linalg.norm(right)
        down = np.cross(forward, right)
        R = np.stack([right, down, forward])
        views.append(View(R, -R @ eye, 1.1 * width, 1.1 * width, 0.5 * width, 0.5 * height, width, height))
    return views


def orbit_views(count: int, radius: float, sizes, seed: int, f_rule=None, elevation_deg=30.0):
    """Cameras on a seeded orbit about the origin, each looking at it, with
    elevation uniform in +-elevation_deg and a resolution drawn from `sizes`.
    f = 1000 * W / 1280 unless `f_rule(W, H)` says otherwise."""
    rng = np.random.default_rng(seed)
    views = []
    for i in range(count):
        az = 2.0 * math.pi * (i + rng.uniform()) / count
        el = math.radians(rng.uniform(-elevation_deg, elevation_deg))
        eye = radius * np.array([math.cos(el) * math.cos(az), -math.sin(el), math.cos(el) * math.sin(az)])
        R, t = look_at(eye, np.zeros(3))
        w, h = sizes[int(rng.integers(len(sizes)))]
        f = f_rule(w, h) if f_rule else 1000.0 * w / 1280.0
        views.append(View(R, t, f, f, 0.5 * w, 0.5 * h, int(w), int(h)))
    return views


# ------------------------------------------------------------------- clouds --

def random_cloud(n: int, seed: int, sh_degree: int, mean_lo, mean_hi, log_scale_mu: float = -4.0,
                 log_scale_sigma: float = 0.5, rest_sigma: float = 0.2) -> Cloud:
    """The benchmark's parameter distributions: means uniform in a box,
    log-scales N(mu, sigma), unit quaternions from normalised N(0,1)^4,
    opacity logits N(0,1), DC from linear RGB U(0,1) (dc = (rgb-0.5)/C0) and
    higher SH bands N(0, rest_sigma)."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(mean_lo, dtype=np.float64)
    hi = np.asarray(mean_hi, dtype=np.float64)
    pos = (lo + (hi - lo) * rng.uniform(size=(n, 3))).astype(np.float32)
    ls = rng.normal(log_scale_mu, log_scale_sigma, size=(n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    op = rng.normal(size=n).astype(np.float32)
    rgb = rng.uniform(size=(n, 3))
    dc = ((rgb - 0.5) / SH_C0).astype(np.float32)
    rs = 3 * sh_rest_per_channel(sh_degree)
    rest = rng.normal(0.0, rest_sigma, size=(n, rs)).astype(np.float32)
    return Cloud(pos, ls, q.astype(np.float32), op, dc, rest, sh_degree).contiguous()


def overlapping_cloud(sh_degree: int, count: int = 6, seed: int = 77) -> Cloud:
    """makeOverlappingCloud / makeGaussian of the engine's rasteriser tests
    (SplatRasterTest.cpp:27-59), drawn with the engine's Rng."""
    rng = Rng(seed)
    per = sh_rest_per_channel(sh_degree)
    pos, ls, rot, op, dc, rest = [], [], [], [], [], []
    for _ in range(count):
        p = [rng.uniform(-0.4, 0.4), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3)]
        s = rng.uniform(-2.6, -2.0)
        logit = rng.uniform(-0.5, 1.5)
        # Eigen::Quaternionf(w, x, y, z) from four normals, then normalised
        q = np.array([rng.normal(), rng.normal(), rng.normal(), rng.normal()], dtype=np.float32)
        if np.linalg.norm(q) < 1e-3:
            q = np.array([1, 0, 0, 0], dtype=np.float32)
        q = q / np.float32(np.sqrt(np.sum(q.astype(np.float32) ** 2)))
        d = [rng.uniform(0.2, 1.2) for _ in range(3)]
        r = [rng.uniform(-0.3, 0.3) for _ in range(per * 3)]
        pos.append(p)
        ls.append([s, s - 0.2, s + 0.3])
        rot.append(q)
        op.append(logit)
        dc.append(d)
        rest.append(r)
    f32 = lambda a: np.asarray(a, dtype=np.float32)
    return Cloud(f32(pos), f32(ls), f32(rot), f32(op), f32(dc),
                 f32(rest).reshape(count, 3 * per), sh_degree).contiguous()


def truth_cloud() -> Cloud:
    """truthCloud of the engine's trainer tests (SplatTrainerTest.cpp:41-57)."""
    positions = [(-0.25, 0.0, 0.0), (0.25, 0.0, 0.0), (0.0, 0.2, 0.2)]
    colors = [(0.9, 0.2, 0.2), (0.2, 0.8, 0.3), (0.3, 0.3, 0.9)]
    pos = np.array(positions, dtype=np.float32)
    ls = np.full((3, 3), np.log(np.float32(0.12)), dtype=np.float32)
    rot = np.tile(np.array([1, 0, 0, 0], dtype=np.float32), (3, 1))
    op = np.full(3, 4.0, dtype=np.float32)
    dc = ((np.array(colors, dtype=np.float32) - np.float32(0.5)) / np.float32(SH_C0)).astype(np.float32)
    return Cloud(pos, ls, rot, op, dc, np.zeros((3, 0), np.float32), 0).contiguous()


def rough_start(count: int = 24, seed: int = 3) -> Cloud:
    """roughStart of the engine's trainer tests (SplatTrainerTest.cpp:72-87)."""
    rng = Rng(seed)
    pos = []
    for _ in range(count):
        pos.append([rng.uniform(-0.4, 0.4), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3)])
    pos = np.array(pos, dtype=np.float32)
    ls = np.full((count, 3), np.log(np.float32(0.1)), dtype=np.float32)
    rot = np.tile(np.array([1, 0, 0, 0], dtype=np.float32), (count, 1))
    return Cloud(pos, ls, rot, np.zeros(count, np.float32), np.zeros((count, 3), np.float32),
                 np.zeros((count, 0), np.float32), 0).contiguous()


# ------------------------------------------------------ benchmark workloads --

def config1(seed: int = 0):
    """Plain single-view step: 10k SH0 gaussians in camera space
    (x, y ~ U(-1,1), z ~ U(2,6)), 256x256, f = 300, identity pose. The
    target is a render of a second set drawn with seed + 1."""
    cloud = random_cloud(10_000, seed, 0, (-1, -1, 2), (1, 1, 6))
    target = random_cloud(10_000, seed + 1, 0, (-1, -1, 2), (1, 1, 6))
    return cloud, target, [pinhole(256, 256, 300.0)]


def config2(seed: int = 0, n: int = 200_000):
    """Single-view gradient check: 200k SH3, 1280x720, f = 1000."""
    cloud = random_cloud(n, seed, 3, (-1, -1, 2), (1, 1, 6))
    target = random_cloud(n, seed + 1, 3, (-1, -1, 2), (1, 1, 6))
    return cloud, target, [pinhole(1280, 720, 1000.0)]


UNEVEN_SIZES = [(640, 480), (1280, 720), (1600, 1200), (1920, 1080)]


def config3(seed: int = 0, n: int = 500_000, views: int = 32):
    """Batched uneven views: 500k SH3 in a 4 m cube, 32 orbit cameras at
    radius 6 m with resolutions drawn from four sizes."""
    cloud = random_cloud(n, seed, 3, (-2, -2, -2), (2, 2, 2))
    target = random_cloud(n, seed + 1, 3, (-2, -2, -2), (2, 2, 2))
    return cloud, target, orbit_views(views, 6.0, UNEVEN_SIZES, seed + 2)


def config4(seed: int = 0, n: int = 1_000_000, views: int = 64):
    """1M SH3 gaussians in a 4 m cube seen by a 64-camera 1920x1080 orbit
    (radius 6 m, elevation +-30 deg, f = 1500): the end state of the
    densifying training run, and the benchmark's workload."""
    cloud = random_cloud(n, seed, 3, (-2, -2, -2), (2, 2, 2))
    target = random_cloud(n, seed + 1, 3, (-2, -2, -2), (2, 2, 2))
    return cloud, target, orbit_views(views, 6.0, [(1920, 1080)], seed + 2)


def config5(seed: int = 0, n: int = 3_000_000, views: int = 8):
    """Room scale: 3M SH3 in a 10 x 10 x 3 m box, log-scales N(-3.5, 0.5),
    8 interior 1920x1080 views at 1.6 m height with random yaw, f = 1100."""
    rng = np.random.default_rng(seed + 2)
    cloud = random_cloud(n, seed, 3, (-5, -3, -5), (5, 0, 5), log_scale_mu=-3.5)
    target = random_cloud(n, seed + 1, 3, (-5, -3, -5), (5, 0, 5), log_scale_mu=-3.5)
    cams = []
    for _ in range(views):
        yaw = rng.uniform(0, 2 * math.pi)
        eye = np.array([rng.uniform(-3, 3), -1.6, rng.uniform(-3, 3)])
        tgt = eye + np.array([math.cos(yaw), 0.0, math.sin(yaw)])
        R, t = look_at(eye, tgt, up=(0.0, -1.0, 0.0))
        cams.append(View(R, t, 1100.0, 1100.0, 960.0, 540.0, 1920, 1080))
    return cloud, target, cams
