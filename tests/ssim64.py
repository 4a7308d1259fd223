"""Float64 restatement of the engine's photometric loss and its gradient.

photometricLoss / meanSsim / blur (engine/native/cpp/src/splat/
SplatRaster.cpp:580-738): L = (1-w) mean|x-y| + w (1 - mean SSIM), the SSIM
statistics from an 11-tap sigma 1.5 Gaussian blur, separable and zero
padded, C1 = 0.01^2, C2 = 0.03^2, means over 3 * pixels. The gradient is the
engine's five blurred partial planes,
    dSSIM/dx = B(t_mu) + 2x B(t_var) - 2 B(t_var mu_x) + y B(t_cov) - B(t_cov mu_y),
evaluated in float64 so that float32 evaluations in any association can be
measured against it. Test infrastructure only: the accuracy yardstick for
dL/drgb, checked against the reference build's golden vectors.
"""
import numpy as np

RADIUS = 5
SIGMA = 1.5


def kernel():
    i = np.arange(-RADIUS, RADIUS + 1, dtype=np.float64)
    k = np.exp(-0.5 * i * i / (SIGMA * SIGMA))
    return k / k.sum()


def blur(img):
    """Separable zero-padded blur of an (h, w) float64 plane, rows first."""
    k = kernel()
    h, w = img.shape
    pad = np.zeros((h, w + 2 * RADIUS))
    pad[:, RADIUS:RADIUS + w] = img
    rows = sum(k[i] * pad[:, i:i + w] for i in range(2 * RADIUS + 1))
    pad = np.zeros((h + 2 * RADIUS, w))
    pad[RADIUS:RADIUS + h] = rows
    return sum(k[i] * pad[i:i + h] for i in range(2 * RADIUS + 1))


def photometric_loss(rendered, target, width, height, ssim_weight, omit=None):
    """(loss, dL/drgb) in float64 for interleaved RGB images of width x height.
    omit names one of the five gradient terms ("mu", "var", "var_mu", "cov",
    "cov_mu") to leave out: a planted mistake the accuracy bar must catch."""
    x_all = np.asarray(rendered, dtype=np.float64).reshape(height, width, 3)
    y_all = np.asarray(target, dtype=np.float64).reshape(height, width, 3)
    n = width * height * 3
    inverse = 1.0 / n
    diff = x_all - y_all
    l1 = np.abs(diff).sum() * inverse
    dl = (1.0 - ssim_weight) * np.sign(diff) * inverse
    if ssim_weight <= 0.0:
        return l1, dl.ravel()
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    total = 0.0
    d_ssim = np.zeros_like(x_all)
    for c in range(3):
        x, y = x_all[:, :, c], y_all[:, :, c]
        mu_x, mu_y, exx, eyy, exy = (blur(v) for v in (x, y, x * x, y * y, x * y))
        var_x, var_y, cov = exx - mu_x * mu_x, eyy - mu_y * mu_y, exy - mu_x * mu_y
        a = 2.0 * mu_x * mu_y + c1
        b = 2.0 * cov + c2
        cc = mu_x * mu_x + mu_y * mu_y + c1
        d = var_x + var_y + c2
        total += (a * b / (cc * d)).sum()
        t_mu = inverse * (2.0 * mu_y * b / (cc * d) - 2.0 * mu_x * a * b / (cc * cc * d))
        t_var = inverse * (-(a * b) / (cc * d * d))
        t_cov = inverse * (2.0 * a / (cc * d))
        terms = {"mu": blur(t_mu), "var": 2.0 * x * blur(t_var), "var_mu": -2.0 * blur(t_var * mu_x),
                 "cov": y * blur(t_cov), "cov_mu": -blur(t_cov * mu_y)}
        d_ssim[:, :, c] = sum(v for k, v in terms.items() if k != omit)
    ssim = total * inverse
    dl = dl - ssim_weight * d_ssim
    return (1.0 - ssim_weight) * l1 + ssim_weight * (1.0 - ssim), dl.ravel()


def errors(g, g64):
    """(max |g - g64| / max |g64|, relative L2) of an evaluation g against g64."""
    e = np.abs(np.asarray(g, dtype=np.float64) - g64)
    return float(e.max() / max(1e-300, np.abs(g64).max())), float(np.linalg.norm(e) / max(1e-300, np.linalg.norm(g64)))
