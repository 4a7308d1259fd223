"""The compact exporters .splat and .spz, against the reference.

The golden fixtures (tests/golden/make_golden.py) hold the reference engine's
own writeSplat / writeSpz output for two seeded clouds that reach every
branch of the quantisers (GaussianCloud.cpp:416-446, 497-560) and for the
gaussians of a 4M cloud whose opacity or rotation bytes sit on the rounding
ties where the build's rewritten expressions decide the byte, and its
readSplat / readSpz result for those files and for hand-laid files that hold
every byte value in every field (:448-493, 562-641).

CPU tests pin the fixtures independently: a numpy restatement of each
quantiser and decoder (round half away from zero, as std::round; Eigen's
packet-order quaternion norm; the optimised build's rewrites: multiplies by
rounded reciprocal constants, 255 / (1 + e) for the opacity byte and
(v + 1) * 127.5 for the rotation bytes) must reproduce them, and the .spz files must be
Python's zlib over the restated planes with the engine's deflate settings.
The library's file checks, which need no device, must reject what the engine
rejects with its messages. GPU tests run the device path through the C ABI:
every .spz byte and every decoded value exact; .splat bytes exact but for
the sigma floats (the correctly rounded exponential, within 1 ulp of the
engine's libm expf) and read log-sigmas within 1 ulp (libm logf).
"""
import gzip
import os
import tempfile
import zlib

import numpy as np
import pytest

from parallax_amd import splat as S

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
f32 = np.float32
C0 = f32(0.28209479177387814)
FIELDS = ("pos", "log_scale", "rot", "opacity", "dc", "rest")


def golden():
    return np.load(os.path.join(HERE, "compact.npz"))


def fixture(name):
    return os.path.join(HERE, name)


def ulp_diff(a, b):
    a = np.asarray(a, f32).view(np.int32).astype(np.int64)
    b = np.asarray(b, f32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


# ---- numpy restatement of the quantisers (float32 arithmetic, no FMA) ----

def to_byte(v):
    """std::max(0, std::min(255, std::round(v))) of float32 values."""
    v = np.asarray(v, np.float64)
    r = np.sign(v) * np.floor(np.abs(v) + 0.5)
    return np.clip(r, 0, 255).astype(np.uint8)


def opacity_byte(logit):
    """quantizeUnit(gaussianOpacity(g)): the optimised build folds
    round(255 * (1 / (1 + e))) into round(255 / (1 + e))."""
    e = np.exp(-np.asarray(logit, np.float64)).astype(f32)
    return to_byte((f32(255) / (f32(1) + e)).astype(f32))


def eigen_normalized(q):
    """Quaternionf::normalized() in (w, x, y, z) order: the squared norm is the
    4-float packet sum (x^2 + z^2) + (y^2 + w^2); zero stays zero."""
    w, x, y, z = (q[:, k].astype(f32) for k in range(4))
    s = (x * x + z * z) + (y * y + w * w)
    n = np.sqrt(s).astype(f32)
    safe = np.where(s > 0, n, f32(1))
    return np.where((s > 0)[:, None], q / safe[:, None], q).astype(f32)


def cloud_of(g, tag):
    n = len(g[f"{tag}_in_opacity"])
    return {k: g[f"{tag}_in_{k}"].reshape(n, -1) for k in FIELDS}


def splat_records(c):
    n = len(c["opacity"])
    rec = np.zeros((n, 32), np.uint8)
    rec[:, :12] = c["pos"].astype(f32).view(np.uint8).reshape(n, 12)
    sigma = np.exp(c["log_scale"].astype(np.float64)).astype(f32)
    rec[:, 12:24] = sigma.view(np.uint8).reshape(n, 12)
    colour = np.clip(f32(0.5) + C0 * c["dc"], f32(0), f32(1)).astype(f32)
    rec[:, 24:27] = to_byte(colour * f32(255))
    rec[:, 27] = opacity_byte(c["opacity"][:, 0])
    rec[:, 28:32] = to_byte(eigen_normalized(c["rot"]) * f32(128) + f32(128))
    return rec


def spz_raw(c, degree, bits, antialiased):
    n = len(c["opacity"])
    rpc = [0, 3, 8, 15][degree]
    head = np.array([0x5053474E, 2, n], np.uint32).tobytes() + bytes([degree, bits, antialiased, 0])
    r = c["pos"].astype(f32) * f32(1 << bits)
    fixed = np.clip(np.sign(r.astype(np.float64)) * np.floor(np.abs(r.astype(np.float64)) + 0.5),
                    -(1 << 23), (1 << 23) - 1).astype(np.int64) & 0xFFFFFF
    pos = np.stack([(fixed >> (8 * j)) & 0xFF for j in range(3)], -1).astype(np.uint8).reshape(n, 9)
    scale = to_byte((c["log_scale"] + f32(10)) * f32(16))
    q = eigen_normalized(c["rot"])
    q = np.where(q[:, :1] < 0, -q, q)
    rot = to_byte((q[:, 1:] + f32(1)) * f32(127.5))  # v * 127.5 + 127.5, factored by the build
    alpha = opacity_byte(c["opacity"][:, 0])
    colour = to_byte((f32(0.5) + C0 * c["dc"]) * f32(255))
    rest = c["rest"][:, : 3 * rpc].reshape(n, 3, rpc).transpose(0, 2, 1)
    sh = to_byte(rest * f32(128) + f32(128))
    return head + b"".join(a.tobytes() for a in (pos, scale, rot, alpha, colour, sh))


def gzip_like_the_engine(raw):
    z = zlib.compressobj(zlib.Z_DEFAULT_COMPRESSION, zlib.DEFLATED, 16 + 15, 8, zlib.Z_DEFAULT_STRATEGY)
    return z.compress(raw) + z.flush()


def logit_of_byte(b):
    alpha = np.clip(np.asarray(b, f32) * (f32(1) / f32(255)), f32(1e-6), f32(1) - f32(1e-6)).astype(f32)
    return np.log((alpha / (f32(1) - alpha)).astype(np.float64)).astype(f32)


def dc_of_byte(b):
    return ((np.asarray(b, f32) * (f32(1) / f32(255)) - f32(0.5)) * (f32(1) / C0)).astype(f32)


def decode_splat(rec):
    n = len(rec)
    out = {"pos": rec[:, :12].copy().view(f32).reshape(n, 3)}
    sigma = rec[:, 12:24].copy().view(f32).reshape(n, 3)
    out["log_scale"] = np.log(np.maximum(sigma, f32(1e-20)).astype(np.float64)).astype(f32)
    out["dc"] = dc_of_byte(rec[:, 24:27])
    out["opacity"] = logit_of_byte(rec[:, 27])
    q = (rec[:, 28:32].astype(f32) - f32(128)) / f32(128)
    w, x, y, z = q.T
    norm = np.sqrt((x * x + z * z) + (y * y + w * w)).astype(f32)
    unit = norm > f32(1e-6)
    out["rot"] = np.where(unit[:, None], q / np.where(unit, norm, f32(1))[:, None], np.array([1, 0, 0, 0], f32))
    return out


def decode_spz(raw):
    n = int(np.frombuffer(raw, np.uint32, 1, 8)[0])
    degree, bits = raw[12], raw[13]
    rpc = [0, 3, 8, 15][degree]
    off = 16
    b = np.frombuffer(raw, np.uint8, 9 * n, off).reshape(n, 3, 3).astype(np.int64)
    off += 9 * n
    v = b[..., 0] | (b[..., 1] << 8) | (b[..., 2] << 16)
    v = np.where(v & 0x800000, v - (1 << 24), v)
    out = {"pos": (v.astype(f32) / f32(1 << bits)).astype(f32)}
    out["log_scale"] = np.frombuffer(raw, np.uint8, 3 * n, off).reshape(n, 3).astype(f32) / f32(16) - f32(10)
    off += 3 * n
    x, y, z = ((np.frombuffer(raw, np.uint8, 3 * n, off).reshape(n, 3).astype(f32) - f32(127.5))
               * (f32(1) / f32(127.5))).T
    off += 3 * n
    w = np.sqrt(np.maximum(f32(0), f32(1) - (x * x + (y * y + z * z)))).astype(f32)
    norm = np.sqrt((x * x + z * z) + (y * y + w * w)).astype(f32)
    unit = norm > f32(1e-6)
    q = np.stack([w, x, y, z], 1)
    out["rot"] = np.where(unit[:, None], q / np.where(unit, norm, f32(1))[:, None], np.array([1, 0, 0, 0], f32))
    out["opacity"] = logit_of_byte(np.frombuffer(raw, np.uint8, n, off))
    off += n
    out["dc"] = dc_of_byte(np.frombuffer(raw, np.uint8, 3 * n, off).reshape(n, 3))
    off += 3 * n
    sh = np.frombuffer(raw, np.uint8, 3 * rpc * n, off).reshape(n, rpc, 3).astype(f32)
    out["rest"] = ((sh - f32(128)) / f32(128)).transpose(0, 2, 1).reshape(n, 3 * rpc)
    assert off + 3 * rpc * n == len(raw)
    return out


# ---- CPU: the restatement reproduces the reference's files and readers ----

@pytest.mark.parametrize("tag", ("sh3", "ties"))
def test_splat_fixture_is_the_quantised_cloud(tag):
    g = golden()
    got = np.fromfile(fixture(f"compact_{tag}.splat"), np.uint8).reshape(-1, 32)
    want = splat_records(cloud_of(g, tag))
    assert np.array_equal(got[:, :12], want[:, :12])
    assert np.array_equal(got[:, 24:], want[:, 24:])
    # sigma: the engine's expf against the correctly rounded exponential
    d = ulp_diff(got[:, 12:24].copy().view(f32), want[:, 12:24].copy().view(f32))
    assert d.max() <= 1 and d.sum() <= max(1, d.size // 100)


@pytest.mark.parametrize("tag", ("sh3", "sh1", "ties"))
def test_spz_fixture_is_the_engine_gzip_of_the_quantised_planes(tag):
    g = golden()
    degree, bits, aa = (int(v) for v in g[f"{tag}_args"])
    data = open(fixture(f"compact_{tag}.spz"), "rb").read()
    raw = spz_raw(cloud_of(g, tag), degree, bits, aa)
    assert gzip.decompress(data) == raw
    assert gzip_like_the_engine(raw) == data


@pytest.mark.parametrize("name, tag, kind", [("laid.splat", "laid_splat", "splat"),
                                             ("compact_sh3.splat", "sh3_splat", "splat"),
                                             ("compact_ties.splat", "ties_splat", "splat"),
                                             ("compact_ties.spz", "ties_spz", "spz"),
                                             ("laid_sh2.spz", "laid_spz", "spz"),
                                             ("compact_sh3.spz", "sh3_spz", "spz"),
                                             ("compact_sh1.spz", "sh1_spz", "spz")])
def test_reference_readers_decode_as_restated(name, tag, kind):
    g = golden()
    if kind == "splat":
        out = decode_splat(np.fromfile(fixture(name), np.uint8).reshape(-1, 32))
    else:
        out = decode_spz(gzip.decompress(open(fixture(name), "rb").read()))
    for k, v in out.items():
        ref = g[f"{tag}_out_{k}"][: v.size]
        if k == "log_scale" and kind == "splat":
            assert ulp_diff(ref, v.ravel()).max() <= 1, k  # libm logf
        else:
            assert np.array_equal(ref, v.ravel()), k


def test_info_reads_the_reference_files():
    assert S.splat_info(fixture("laid.splat")) == 512
    assert S.splat_info(fixture("compact_sh3.splat")) == 600
    assert S.spz_info(fixture("compact_ties.spz")) == (S.splat_info(fixture("compact_ties.splat")), 3, False)
    assert S.spz_info(fixture("compact_sh3.spz")) == (600, 3, True)
    assert S.spz_info(fixture("compact_sh1.spz")) == (200, 1, False)
    assert S.spz_info(fixture("laid_sh2.spz")) == (512, 2, True)


def spz_bytes(count=1, degree=0, bits=12, magic=0x5053474E, version=2, payload=None):
    head = np.array([magic, version, count], np.uint32).tobytes() + bytes([degree, bits, 0, 0])
    body = bytes((19 + 3 * [0, 3, 8, 15][min(degree, 3)]) * count) if payload is None else payload
    return gzip.compress(head + body, mtime=0)


@pytest.mark.parametrize("suffix, body, message", [
    (".splat", b"", ".splat size is not a multiple of the 32-byte record"),
    (".splat", bytes(33), ".splat size is not a multiple of the 32-byte record"),
    (".spz", b"plain bytes, not gzip", ".spz is not a valid gzip stream"),
    (".spz", gzip.compress(bytes(12), mtime=0), ".spz header is truncated"),
    (".spz", spz_bytes()[:-9], ".spz is not a valid gzip stream"),
    (".spz", spz_bytes(magic=0x12345678), "bad .spz magic"),
    (".spz", spz_bytes(version=3), "unsupported .spz version"),
    (".spz", spz_bytes(degree=4, payload=b""), ".spz sh_degree out of range"),
    (".spz", spz_bytes(count=2, payload=bytes(19)), ".spz payload size does not match the declared point count"),
])
def test_info_rejects_what_the_engine_rejects(suffix, body, message):
    with tempfile.NamedTemporaryFile(suffix=suffix, delete=False) as f:
        f.write(body)
        path = f.name
    try:
        with pytest.raises(S.SplatError) as e:
            (S.splat_info if suffix == ".splat" else S.spz_info)(path)
        assert message in str(e.value)
    finally:
        os.unlink(path)


def test_info_reports_a_missing_file():
    for fn in (S.splat_info, S.spz_info):
        with pytest.raises(S.SplatError) as e:
            fn("/nonexistent/cloud.bin")
        assert "cannot read" in str(e.value)


# ---- GPU: the device path through the C ABI ----

def upload(ctx, g, tag):
    from parallax_amd import synthetic as syn
    c = cloud_of(g, tag)
    degree = int(g[f"{tag}_args"][0])
    n = len(c["opacity"])
    cloud = syn.Cloud(c["pos"], c["log_scale"], c["rot"], c["opacity"][:, 0], c["dc"],
                      c["rest"].reshape(n, -1), degree)
    ctx.set_capacity(n, degree)
    ctx.upload_cloud(cloud)
    return n, degree


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ("sh3", "ties"))
def test_write_splat_matches_the_reference(ctx, tag):
    g = golden()
    upload(ctx, g, tag)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out.splat")
        ctx.write_splat(path)
        got = np.fromfile(path, np.uint8).reshape(-1, 32)
    want = np.fromfile(fixture(f"compact_{tag}.splat"), np.uint8).reshape(-1, 32)
    assert got.shape == want.shape
    assert np.array_equal(got[:, :12], want[:, :12])
    assert np.array_equal(got[:, 24:], want[:, 24:])
    d = ulp_diff(got[:, 12:24].copy().view(f32), want[:, 12:24].copy().view(f32))
    print(f"[compact] .splat sigma: {int((d > 0).sum())} of {d.size} floats 1 ulp from the engine's expf", flush=True)
    assert d.max() <= 1


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ("sh3", "sh1", "ties"))
def test_write_spz_is_byte_identical_to_the_reference(ctx, tag):
    g = golden()
    upload(ctx, g, tag)
    _, bits, aa = (int(v) for v in g[f"{tag}_args"])
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out.spz")
        ctx.write_spz(path, bits, bool(aa))
        got = open(path, "rb").read()
    assert got == open(fixture(f"compact_{tag}.spz"), "rb").read()


@pytest.mark.gpu
@pytest.mark.parametrize("name, tag", [("laid.splat", "laid_splat"), ("compact_sh3.splat", "sh3_splat"),
                                       ("compact_ties.splat", "ties_splat")])
def test_read_splat_matches_the_reference(ctx, name, tag):
    g = golden()
    n = S.splat_info(fixture(name))
    ctx.set_capacity(n, 0)
    ctx.read_splat(fixture(name))
    c = ctx.download_cloud()
    assert c.n == n and c.sh_degree == 0
    for k in ("pos", "rot", "opacity", "dc"):
        assert np.array_equal(getattr(c, k).ravel(), g[f"{tag}_out_{k}"]), k
    assert ulp_diff(c.log_scale.ravel(), g[f"{tag}_out_log_scale"]).max() <= 1


@pytest.mark.gpu
@pytest.mark.parametrize("name, tag", [("laid_sh2.spz", "laid_spz"), ("compact_sh3.spz", "sh3_spz"),
                                       ("compact_sh1.spz", "sh1_spz"), ("compact_ties.spz", "ties_spz")])
def test_read_spz_matches_the_reference(ctx, name, tag):
    g = golden()
    n, degree, aa = S.spz_info(fixture(name))
    ctx.set_capacity(n, degree)
    assert ctx.read_spz(fixture(name)) == aa
    c = ctx.download_cloud()
    assert c.n == n and c.sh_degree == degree
    for k in FIELDS:
        got = getattr(c, k).ravel()
        assert np.array_equal(got, g[f"{tag}_out_{k}"][: got.size]), k


@pytest.mark.gpu
def test_spz_requantises_what_it_decoded(ctx):
    """Write, read, write again at 200k gaussians: positions, scales, alphas,
    colours and SH survive the second quantisation byte for byte; rotation
    bytes move by at most one step (w is re-derived from x, y, z)."""
    from parallax_amd import synthetic as syn
    cloud = syn.random_cloud(200_000, 45, 3, (-4, -4, -4), (4, 4, 4))
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    with tempfile.TemporaryDirectory() as d:
        a, b = os.path.join(d, "a.spz"), os.path.join(d, "b.spz")
        ctx.write_spz(a, 12, False)
        ctx.read_spz(a)
        ctx.write_spz(b, 12, False)
        ra, rb = gzip.decompress(open(a, "rb").read()), gzip.decompress(open(b, "rb").read())
    n = cloud.n
    assert len(ra) == len(rb) == 16 + 64 * n
    xa, xb = np.frombuffer(ra, np.uint8), np.frombuffer(rb, np.uint8)
    rot = slice(16 + 12 * n, 16 + 15 * n)
    keep = np.ones(len(ra), bool)
    keep[rot] = False
    assert np.array_equal(xa[keep], xb[keep])
    step = np.abs(xa[rot].astype(int) - xb[rot].astype(int))
    print(f"[compact] re-quantised rotation bytes: {int((step > 0).sum())} of {step.size} moved", flush=True)
    assert step.max() <= 1
