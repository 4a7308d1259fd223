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


# ---- the projection's cull box (project.hip: cull_q, cull_extents) ----

def cull_box_double(a, b, c, op, min_alpha):
    """The box in double precision: the exact half extents widened by 1 + 1e-6
    and 1e-3 px (+inf where the conic is not positive definite or too
    elongated for the float error bound, ac/det > 187500)."""
    A, B, C = (np.asarray(v, np.float64) for v in (a, b, c))
    det = A * C - B * B
    ex = np.full(A.shape, np.inf)
    ey = np.full(A.shape, np.inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        kappa = A * C / det
        shrink = 1.0 - 4e-6 * kappa
        ok = (det > 0) & (shrink > 0.25)
        q = np.maximum(0.0, 2.0 * (np.log(np.float64(op) / np.float64(min_alpha)) + 1e-5))
        ex[ok] = np.sqrt(q * C / (det * shrink))[ok] * (1 + 1e-6) + 1e-3
        ey[ok] = np.sqrt(q * A / (det * shrink))[ok] * (1 + 1e-6) + 1e-3
    return ex, ey


def cull_box_float(a, b, c, op, min_alpha, rcp_err, sqrt_err):
    """float32 restatement of the kernel's box: q per gaussian in double,
    rounded up; Kahan's fused determinant; the hardware reciprocal and square
    root modelled as the correctly rounded ones times (1 + err)."""
    a, b, c = (np.asarray(v, f32) for v in (a, b, c))
    qd = np.maximum(0.0, 2.0 * (np.log(np.float64(op) / np.float64(min_alpha)) + 1e-5))
    q = (qd * (1.0 + 1e-7)).astype(f32)

    def fma(x, y, z):  # one rounding of x y + z (the double product of two floats is exact)
        return (x.astype(np.float64) * y.astype(np.float64) + z.astype(np.float64)).astype(f32)

    def rcp(x):
        return ((f32(1) / x).astype(f32) * f32(1 + rcp_err)).astype(f32)

    def hsqrt(x):
        return (np.sqrt(x).astype(f32) * f32(1 + sqrt_err)).astype(f32)

    bb = (b * b).astype(f32)
    det = (fma(a, c, -bb) + fma(-b, b, bb)).astype(f32)
    ex = np.full(a.shape, np.inf, f32)
    ey = np.full(a.shape, np.inf, f32)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        rdet = rcp(det)
        shrink = (f32(1) - (f32(4e-6) * ((a * c).astype(f32) * rdet).astype(f32)).astype(f32)).astype(f32)
        ok = (det > 0) & (shrink > f32(0.25))
        s = ((q * rdet).astype(f32) * rcp(shrink)).astype(f32)
        bx = (hsqrt((s * c).astype(f32)) * f32(1.00001) + f32(1e-3)).astype(f32)
        by = (hsqrt((s * a).astype(f32)) * f32(1.00001) + f32(1e-3)).astype(f32)
    ex[ok], ey[ok] = bx[ok], by[ok]
    return ex, ey, ok, (a.astype(np.float64) * c) / np.where(det > 0, det, 1)


def test_the_float_cull_box_holds_the_double_box():
    rng = np.random.default_rng(13)
    n = 200_000
    a, b, c = conics(rng, n)
    # near-singular conics too: the minor axis up to 2000 times the major
    s = np.exp(rng.uniform(-1, 3.5, n))
    ratio = np.exp(rng.uniform(0, np.log(2000.0), n))
    th = rng.uniform(0, np.pi, n)
    cs, sn = np.cos(th), np.sin(th)
    cxx = cs * cs * s ** 2 + sn * sn * (s * ratio) ** 2
    cyy = sn * sn * s ** 2 + cs * cs * (s * ratio) ** 2
    cxy = cs * sn * (s ** 2 - (s * ratio) ** 2)
    det = cxx * cyy - cxy * cxy
    a = np.concatenate([a, cyy / det]).astype(f32)
    b = np.concatenate([b, -cxy / det]).astype(f32)
    c = np.concatenate([c, cxx / det]).astype(f32)
    op = rng.uniform(1e-3, 1, len(a))
    for min_alpha in (MIN_ALPHA, 1e-3, 1e-20):
        dex, dey = cull_box_double(a, b, c, op, min_alpha)
        for rcp_err in (-2.4e-7, 2.4e-7):
            for sqrt_err in (-2.4e-7, 2.4e-7):
                fex, fey, ok, kappa = cull_box_float(a, b, c, op, min_alpha, rcp_err, sqrt_err)
                # wherever the double box is finite the float box holds it,
                # or is +inf (no cull); a finite float box where the double
                # gave none is allowed only at the elongation limit
                finite = np.isfinite(dex)
                assert np.all((fex[finite] >= dex[finite]) & (fey[finite] >= dey[finite])), min_alpha
                extra = ok & ~finite
                assert np.all(np.abs(kappa[extra] / 187500.0 - 1) < 1e-4), min_alpha
                assert finite.sum() > len(a) // 2
