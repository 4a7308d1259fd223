"""Seeding from a point cloud and the 3DGS PLY file, against the reference.

The golden fixtures (tests/golden/make_golden.py) hold the reference engine's
own initializeFromPointCloud output for seeded point sets
(GaussianCloud.cpp:199-256), writeGaussianPly's bytes for a seeded SH2 cloud
(:259-303), and readGaussianPly's cloud for a hand-laid file in another
property order with no normals, no f_dc and quaternions to renormalise
(:305-403).

CPU tests pin the fixtures' semantics independently (the strided
nearest-neighbour spacing, brute force) and check the library's PLY header
parser, which needs no device. GPU tests run the device path through the C
ABI: positions, rotations, opacities, every value read from a PLY file and the
written bytes exact; seeded log sigma to 2 ulp and DC to 2 ulp of its range
(the reference build's -ffast-math turns its divisions by constants into
reciprocal multiplies).
"""
import os
import tempfile

import numpy as np
import pytest

from parallax_amd import splat as S

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("rgb_sh1", "plain_sh0", "big_sh3")


def init_golden():
    return np.load(os.path.join(HERE, "init_points.npz"))


def ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("tag", CASES)
def test_seed_spacing_is_the_sampled_nearest_neighbour_mean(tag):
    """log sigma = log(max(1e-4, mean / 2)) of each sample point's nearest
    other sample point, sample = points 0, s, 2s, ... with s = n / min(n, 512)."""
    g = init_golden()
    xyz = g[f"{tag}_xyz"].astype(np.float64)
    n = len(xyz)
    stride = max(1, n // min(n, 512))
    sample = xyz[::stride]
    d = np.sqrt(((sample[:, None, :] - sample[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    sigma = max(1e-4, d.min(axis=1).mean() * 0.5)
    got = g[f"{tag}_out_log_scale"].reshape(-1, 3)
    assert np.all(got == got[0, 0])
    assert abs(float(got[0, 0]) - np.log(sigma)) <= 2e-6 * abs(np.log(sigma))
    deg, opacity = g[f"{tag}_args"]
    o32 = np.float32(opacity)
    assert ulp_diff(g[f"{tag}_out_opacity"], np.log(o32 / (np.float32(1) - o32))).max() <= 1
    assert np.array_equal(g[f"{tag}_out_pos"].reshape(-1, 3), g[f"{tag}_xyz"])


def test_ply_info_reads_the_reference_file():
    n, deg = S.ply_info(os.path.join(HERE, "cloud_sh2.ply"))
    assert (n, deg) == (500, 2)
    n, deg = S.ply_info(os.path.join(HERE, "laid_sh1.ply"))
    assert (n, deg) == (64, 1)


@pytest.mark.parametrize("body, message", [
    (b"not a ply at all\nend_header\n", "not a PLY file"),
    (b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n", "binary little-endian"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty double x\nend_header\n",
     "must be float"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\nproperty float y\n"
     b"property float z\nproperty float opacity\nend_header\n\x00\x00", "truncated"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\nend_header\n" + b"\x00" * 4,
     "missing core properties"),
])
def test_ply_info_rejects_what_the_engine_rejects(body, message):
    with tempfile.NamedTemporaryFile(suffix=".ply", delete=False) as f:
        f.write(body)
        path = f.name
    try:
        with pytest.raises(S.SplatError) as e:
            S.ply_info(path)
        assert message in str(e.value)
    finally:
        os.unlink(path)


def _ply(lines, body_bytes):
    return ("\n".join(lines) + "\n").encode() + b"\x00" * body_bytes


_CORE = ["property float x", "property float y", "property float z", "property float opacity"]
_LE = ["ply", "format binary_little_endian 1.0"]
_HEADERS = {
    "core only": _ply(_LE + ["element vertex 2"] + _CORE + ["end_header"], 32),
    "sh1 all fields": _ply(_LE + ["element vertex 1"] + _CORE + [f"property float f_dc_{c}" for c in range(3)]
                           + [f"property float f_rest_{k}" for k in range(9)]
                           + [f"property float rot_{k}" for k in range(4)] + ["end_header"], 4 * 20),
    "rest not a multiple of 3": _ply(_LE + ["element vertex 1"] + _CORE
                                     + [f"property float f_rest_{k}" for k in range(10)] + ["end_header"], 56),
    "rest with no degree": _ply(_LE + ["element vertex 1"] + _CORE
                                + [f"property float f_rest_{k}" for k in range(12)] + ["end_header"], 64),
    "rest with a gap": _ply(_LE + ["element vertex 1"] + _CORE
                            + [f"property float f_rest_{k}" for k in (0, 1, 2, 4)] + ["end_header"], 32),
    "later element resets the count": _ply(_LE + ["element vertex 2"] + _CORE
                                           + ["element face 0", "end_header"], 32),
    "element with no count keeps it": _ply(_LE + ["element vertex 2"] + _CORE + ["element extra", "end_header"], 32),
    "count with trailing letters": _ply(_LE + ["element vertex 2abc"] + _CORE + ["end_header"], 16),
    "count that is not a number": _ply(_LE + ["element vertex abc"] + _CORE + ["end_header"], 32),
    "list property": _ply(_LE + ["element vertex 1"] + _CORE + ["property list uchar int idx", "end_header"], 16),
    "float32 and an unnamed property": _ply(_LE + ["element vertex 1", "property float32 x", "property float",
                                                   "property float y", "property float z",
                                                   "property float opacity", "end_header"], 20),
    "duplicate names": _ply(_LE + ["element vertex 1"] + _CORE + ["property float x", "end_header"], 20),
    "carriage returns and tabs": ("ply\r\nformat binary_little_endian 1.0\r\n\telement  vertex\t1\r\n"
                                  + "".join(f" {p}\r\n" for p in _CORE) + "end_header\n").encode() + b"\x00" * 16,
    "end_header inside a line": _ply(_LE + ["element vertex 1"] + _CORE[:3] + ["property float opacityend_header"],
                                     16),
    "format named only in the body": _ply(["ply", "format ascii 1.0", "element vertex 1"] + _CORE
                                          + ["end_header"], 0) + b"binary_little_endian",
    "one byte short": _ply(_LE + ["element vertex 2"] + _CORE + ["end_header"], 31),
    "no magic": _ply(["plx"] + _LE[1:] + ["element vertex 1"] + _CORE + ["end_header"], 16),
    "no end_header": _ply(_LE + ["element vertex 1"] + _CORE, 16),
}


@pytest.mark.parametrize("case", sorted(_HEADERS))
def test_ply_header_rules_match_the_reference_reader(case, tmp_path):
    """splat_ply_info accepts exactly the headers the engine's own
    readGaussianPly accepts (reference build), with the same count and SH
    degree, and rejects the others with the engine's message."""
    import ctypes as C

    import oracle as O
    if not O.ref_available():
        pytest.skip("reference build (oracle/_ref) not present")
    path = tmp_path / "case.ply"
    path.write_bytes(_HEADERS[case])
    n, deg = C.c_int32(), C.c_int32()
    err = C.create_string_buffer(256)
    ref_ok = O.ref().ref_read_ply(str(path).encode(), C.byref(n), C.byref(deg), None, err, 256) == 1
    if ref_ok:
        assert S.ply_info(str(path)) == (n.value, deg.value)
    else:
        with pytest.raises(S.SplatError) as e:
            S.ply_info(str(path))
        assert str(e.value).endswith(": " + err.value.decode())


def test_ply_header_count_that_wraps_is_truncated(tmp_path):
    """A count of -1 reads as 2^64 - 1; the engine's size check
    (GaussianCloud.cpp:372-374) wraps round and its reader then aborts in
    reserve(), so there is no reference result: the library reports a
    truncated body."""
    path = tmp_path / "wrap.ply"
    path.write_bytes(_ply(_LE + ["element vertex -1"] + _CORE + ["end_header"], 32))
    with pytest.raises(S.SplatError) as e:
        S.ply_info(str(path))
    assert "3DGS PLY body is truncated" in str(e.value)


def download(ctx):
    c = ctx.download_cloud()
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("tag", CASES)
def test_init_from_points_matches_the_reference(ctx, tag):
    g = init_golden()
    xyz = g[f"{tag}_xyz"]
    rgb = g[f"{tag}_rgb"] if f"{tag}_rgb" in g else None
    deg, opacity = g[f"{tag}_args"]
    deg = int(deg)
    print(f"[formats] init {tag}: {len(xyz)} points, degree {deg}", flush=True)
    ctx.set_capacity(len(xyz) + 17, deg)
    log_sigma = ctx.init_from_points(xyz, rgb, deg, float(opacity))
    c = download(ctx)
    assert c.n == len(xyz) and c.sh_degree == deg
    assert np.array_equal(c.pos.ravel(), g[f"{tag}_out_pos"])
    assert np.array_equal(c.rot.ravel(), g[f"{tag}_out_rot"])
    assert np.array_equal(c.opacity.ravel(), g[f"{tag}_out_opacity"])
    assert ulp_diff(c.log_scale.ravel(), g[f"{tag}_out_log_scale"]).max() <= 2
    assert ulp_diff(np.float32(log_sigma), g[f"{tag}_out_log_scale"][0]) <= 2
    # (c / 255 - 0.5) / C0 cancels near mid-grey: compare against the value
    # range (2 ulp of 1.77), not per-value ulps
    assert np.abs(c.dc.ravel() - g[f"{tag}_out_dc"]).max() <= 2.4e-7
    assert not np.any(c.rest)


@pytest.mark.gpu
def test_write_ply_is_byte_identical_to_the_reference(ctx):
    g = np.load(os.path.join(HERE, "ply_sh2.npz"))
    from parallax_amd import synthetic as syn
    n = len(g["in_opacity"])
    cloud = syn.Cloud(g["in_pos"].reshape(n, 3), g["in_log_scale"].reshape(n, 3), g["in_rot"].reshape(n, 4),
                      g["in_opacity"], g["in_dc"].reshape(n, 3), g["in_rest"].reshape(n, -1), 2)
    ctx.set_capacity(n, 2)
    ctx.upload_cloud(cloud)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out.ply")
        ctx.write_ply(path)
        got = open(path, "rb").read()
    want = open(os.path.join(HERE, "cloud_sh2.ply"), "rb").read()
    assert got == want


@pytest.mark.gpu
def test_read_ply_matches_the_reference_reader(ctx):
    g = np.load(os.path.join(HERE, "ply_sh2.npz"))
    path = os.path.join(HERE, "laid_sh1.ply")
    n, deg = S.ply_info(path)
    ctx.set_capacity(n, deg)
    ctx.read_ply(path)
    c = download(ctx)
    assert c.n == n and c.sh_degree == int(g["laid_degree"])
    for k in ("pos", "log_scale", "opacity", "dc", "rest"):
        assert np.array_equal(getattr(c, k).ravel(), g[f"laid_out_{k}"][: getattr(c, k).size]), k
    # renormalised by Eigen's norm (the 4-float packet sum (x^2 + z^2) +
    # (y^2 + w^2)) and a division, as the engine's build evaluates it
    assert np.array_equal(c.rot.ravel(), g["laid_out_rot"])


@pytest.mark.gpu
def test_ply_round_trip_keeps_the_cloud(ctx):
    from parallax_amd import synthetic as syn
    cloud = syn.random_cloud(3000, 44, 3, (-2, -2, -2), (2, 2, 2))
    ctx.set_capacity(cloud.n, 3)
    ctx.upload_cloud(cloud)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "rt.ply")
        ctx.write_ply(path)
        ctx.read_ply(path)
    c = download(ctx)
    for k in ("pos", "log_scale", "opacity", "dc", "rest"):
        assert np.array_equal(getattr(c, k), getattr(cloud, k)), k
    assert ulp_diff(c.rot, cloud.rot).max() <= 2
