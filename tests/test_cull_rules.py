"""The compositing kernels' quadrant cull is conservative.

Both compositing kernels (parallax_amd/csrc/raster.hip) skip a splat for an
8x8 quadrant when no point of the quadrant's pixel-centre rectangle lies in
the splat's alpha >= min_alpha ellipse (`ellipse_meets_rect`). That is only
allowed if every pixel of the quadrant would skip the splat anyway, as the
engine's compositing loop decides it (SplatRaster.cpp:294-303: power in
float, alpha = min(0.99, opacity * exp(power)), skip alpha < 1/255). This
test restates the rule in float32 (hardware log2 perturbed by +-2 ulp) and
checks, by brute force over the 64 pixel centres, that no rejected quadrant
holds a passing pixel -- for random splats, and for splats whose threshold
ellipse ends just beyond the quadrant's nearest corner.
"""
import numpy as np

f32 = np.float32
MIN_ALPHA = 1.0 / 255.0


def ellipse_meets_rect(a, b, c, op, x0, x1, y0, y1, log_err):
    """Vectorised float32 restatement of raster.hip's ellipse_meets_rect."""
    det = (a * c - b * b).astype(f32)
    elongated = ~(det > 0) | ((a * c).astype(f32) > f32(1000) * det)
    la = (np.log2(op).astype(f32) * f32(1 + log_err)).astype(f32)
    lm = f32(np.log2(f32(MIN_ALPHA))) * f32(1 - log_err)
    q = (f32(2 * 0.69314718) * (la - lm)).astype(f32)
    q_hi = (q * f32(1.01) + f32(1e-3)).astype(f32)
    inside = (x0 <= 0) & (x1 >= 0) & (y0 <= 0) & (y1 >= 0)
    ia, ic = (f32(1) / a).astype(f32), (f32(1) / c).astype(f32)
    best = np.full(a.shape, np.inf, f32)
    for xe, ye in ((x0, y0), (x1, y1)):
        yc = np.minimum(np.maximum((-b * xe * ic).astype(f32), y0), y1)
        xc = np.minimum(np.maximum((-b * ye * ia).astype(f32), x0), x1)
        best = np.minimum(best, (a * xe * xe + f32(2) * b * xe * yc + c * yc * yc).astype(f32))
        best = np.minimum(best, (a * xc * xc + f32(2) * b * xc * ye + c * ye * ye).astype(f32))
    return elongated | np.where(inside, q_hi >= 0, best <= q_hi)


def conics(rng, n):
    s = np.exp(rng.uniform(-1, 3.5, (n, 2)))
    th = rng.uniform(0, np.pi, n)
    cs, sn = np.cos(th), np.sin(th)
    cxx = cs * cs * s[:, 0] ** 2 + sn * sn * s[:, 1] ** 2 + 0.3
    cyy = sn * sn * s[:, 0] ** 2 + cs * cs * s[:, 1] ** 2 + 0.3
    cxy = cs * sn * (s[:, 0] ** 2 - s[:, 1] ** 2)
    det = cxx * cyy - cxy * cxy
    return cyy / det, -cxy / det, cxx / det  # inverse covariance: a, b, c


def passing_pixels(a, b, c, op, mx, my):
    """Pixels of the quadrant (centres 0.5..7.5) that take the splat, with the
    engine's float power and a 1e-6 slack towards passing."""
    px = (np.arange(8, dtype=f32) + f32(0.5))
    dx = (px[None, None, :] - mx[:, None, None]).astype(f32)
    dy = (px[None, :, None] - my[:, None, None]).astype(f32)
    A, B, C = a[:, None, None], b[:, None, None], c[:, None, None]
    power = (f32(-0.5) * ((A * dx * dx).astype(f32) + (C * dy * dy).astype(f32)) - (B * dx * dy).astype(f32))
    power = power.astype(f32)
    alpha = np.minimum(0.99, op[:, None, None].astype(np.float64) * np.exp(np.minimum(power, 0).astype(np.float64)))
    return ((power <= 0) & (alpha >= MIN_ALPHA * (1 - 1e-6))).reshape(len(a), -1).any(axis=1)


def check(a, b, c, op, mx, my, rng):
    a, b, c, op, mx, my = (v.astype(f32) for v in (a, b, c, op, mx, my))
    x0, x1 = (f32(0.5) - mx).astype(f32), (f32(7.5) - mx).astype(f32)
    y0, y1 = (f32(0.5) - my).astype(f32), (f32(7.5) - my).astype(f32)
    keep = np.ones(len(a), bool)
    for err in (-2.4e-7, 2.4e-7):
        keep &= ellipse_meets_rect(a, b, c, op, x0, x1, y0, y1, err)
    rejected = ~keep
    wrong = rejected & passing_pixels(a, b, c, op, mx, my)
    return int(rejected.sum()), int(wrong.sum())


def test_random_splats_are_never_wrongly_rejected():
    rng = np.random.default_rng(11)
    n = 40_000
    a, b, c = conics(rng, n)
    op = rng.uniform(MIN_ALPHA, 1, n)
    mx, my = rng.uniform(-40, 48, n), rng.uniform(-40, 48, n)
    rejected, wrong = check(a, b, c, op, mx, my, rng)
    assert rejected > n // 2
    assert wrong == 0


def test_splats_just_beyond_the_corner_are_never_wrongly_rejected():
    rng = np.random.default_rng(12)
    n = 40_000
    a, b, c = conics(rng, n)
    op = rng.uniform(MIN_ALPHA, 1, n)
    q = 2 * np.log(op / MIN_ALPHA)
    ang = rng.uniform(0, 2 * np.pi, n)
    dx, dy = np.cos(ang), np.sin(ang)
    Qd = a * dx * dx + 2 * b * dx * dy + c * dy * dy
    r = np.sqrt(q / Qd) * (1 + rng.uniform(0.0, 0.03, n))  # threshold ellipse ends 0-3% short of the corner
    mx, my = 0.5 - r * dx, 0.5 - r * dy
    rejected, wrong = check(a, b, c, op, mx, my, rng)
    assert rejected > n // 8
    assert wrong == 0
