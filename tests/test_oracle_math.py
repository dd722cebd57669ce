"""Pins the oracle's arithmetic building blocks against published known
answers, closed forms and statistical invariants (CPU only).

Reference algorithms: deps/embree/random_sampler.h (MurmurHash3 mix and
finalizer, Numerical Recipes LCG), render/sampler.h:49-91 (concentric disk),
render/sampler.cc + sampler.h:112-128 (cosine hemisphere), render/camera.h:
128-209 (pinhole camera), render/tile.cc:317-391 (blocking tiles).
"""
import ctypes as C
import math

import numpy as np
import pytest

import orc


def murmur3_x86_32(data: bytes, seed: int) -> int:
    """Full MurmurHash3_x86_32 of a key whose length is a multiple of 4,
    composed from the oracle's block mix and finalizer."""
    lib = orc.lib()
    assert len(data) % 4 == 0
    h = seed & 0xFFFFFFFF
    for i in range(0, len(data), 4):
        h = lib.orc_murmur_mix(h, int.from_bytes(data[i:i + 4], "little"))
    return lib.orc_murmur_fmix(h ^ len(data))


# Published MurmurHash3_x86_32 test vectors (SMHasher verification set).
MURMUR_VECTORS = [
    (b"", 0, 0x00000000),
    (b"", 1, 0x514E28B7),
    (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\x00\x00\x00\x00", 0, 0x2362F9DE),
    (b"\xff\xff\xff\xff", 0, 0x76293B50),
    (b"\x21\x43\x65\x87", 0, 0xF55B516B),
    (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE),
    (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"abcd", 0x9747B28C, 0xF0478627),
    (b"test", 0, 0xBA6BD213),
]


@pytest.mark.parametrize("data,seed,want", MURMUR_VECTORS)
def test_murmur3_published_vectors(data, seed, want):
    assert murmur3_x86_32(data, seed) == want


def nr_lcg(s):
    return (s * 1664525 + 1013904223) & 0xFFFFFFFF


def test_lcg_numerical_recipes_sequence():
    # Numerical Recipes ranqd1 from state 0 (published sequence)
    s, seq = 0, []
    for _ in range(6):
        s = nr_lcg(s)
        seq.append(s)
    assert seq == [1013904223, 1196435762, 3519870697, 2868466484, 1649599747, 2670642822]


@pytest.mark.parametrize("rid", [0, 1, 7, 12345, 2**31 - 1, -5])
def test_rng_stream(rid):
    """seed = fmix(mix(0, id)) (no length term); each draw advances the LCG
    and returns (state >> 1) * 2^-31 in float32."""
    lib = orc.lib()
    seed = lib.orc_rng_seed(rid)
    assert seed == lib.orc_murmur_fmix(lib.orc_murmur_mix(0, rid & 0xFFFFFFFF))
    out = np.zeros(8, np.float32)
    lib.orc_rng_floats(rid, 8, orc.p(out))
    s = seed
    for k in range(8):
        s = nr_lcg(s)
        want = np.float32(np.float32(s >> 1) * np.float32(2.0 ** -31))
        assert out[k] == want, k
        assert 0.0 <= out[k] < 1.0 or out[k] == 1.0


def test_rng_uniformity():
    lib = orc.lib()
    vals = []
    for rid in range(2000):
        out = np.zeros(4, np.float32)
        lib.orc_rng_floats(rid, 4, orc.p(out))
        vals.append(out)
    v = np.concatenate(vals)
    hist, _ = np.histogram(v, bins=20, range=(0, 1))
    exp = len(v) / 20
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 50  # 19 dof: p ~ 1e-4


def test_sincos_matches_float64_libm():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-10, 10, 200000), np.linspace(-7, 7, 10001),
                        [0.0, math.pi / 4, math.pi / 2, math.pi, 2 * math.pi, -math.pi]]).astype(np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    orc.lib().orc_sincos(len(x), orc.p(x), orc.p(s), orc.p(c))
    ws = np.sin(x.astype(np.float64)).astype(np.float32)
    wc = np.cos(x.astype(np.float64)).astype(np.float32)

    def ulps(a, b):
        ai = a.view(np.int32).astype(np.int64)
        bi = b.view(np.int32).astype(np.int64)
        ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
        bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
        return np.abs(ai - bi)
    assert ulps(s, ws).max() <= 1 and ulps(c, wc).max() <= 1
    assert (s == ws).mean() > 0.999 and (c == wc).mean() > 0.999


def concentric(u1, u2):
    u1 = np.ascontiguousarray(u1, np.float32)
    u2 = np.ascontiguousarray(u2, np.float32)
    dx = np.zeros_like(u1)
    dy = np.zeros_like(u1)
    orc.lib().orc_concentric(len(u1), orc.p(u1), orc.p(u2), orc.p(dx), orc.p(dy))
    return dx, dy


def test_concentric_disk_closed_forms():
    h = math.sqrt(0.5)
    cases = [((0.5, 0.5), (0.0, 0.0)), ((1.0, 0.5), (1.0, 0.0)), ((0.5, 1.0), (0.0, 1.0)),
             ((0.0, 0.5), (-1.0, 0.0)), ((0.5, 0.0), (0.0, -1.0)), ((1.0, 1.0), (h, h)),
             ((0.0, 0.0), (-h, -h)), ((0.75, 0.5), (0.5, 0.0))]
    u1 = [c[0][0] for c in cases]
    u2 = [c[0][1] for c in cases]
    dx, dy = concentric(u1, u2)
    for i, (_, (ex, ey)) in enumerate(cases):
        assert abs(dx[i] - ex) < 1e-6 and abs(dy[i] - ey) < 1e-6, cases[i]


def test_concentric_disk_maps_squares_to_circles():
    """Radius = max(|2u1-1|, |2u2-1|) (square rings map to circles) and the
    angle advances linearly along each ring; the image is area-uniform."""
    rng = np.random.default_rng(11)
    u1 = rng.uniform(0, 1, 200000).astype(np.float32)
    u2 = rng.uniform(0, 1, 200000).astype(np.float32)
    dx, dy = concentric(u1, u2)
    sx = 2.0 * u1.astype(np.float64) - 1
    sy = 2.0 * u2.astype(np.float64) - 1
    r = np.maximum(np.abs(sx), np.abs(sy))
    assert np.allclose(np.hypot(dx, dy), r, atol=2e-6)
    # on the ring of radius r, the boundary position (0..8r, counter-clockwise
    # from the +x axis) fixes the angle: theta = (pi/4) * position / r
    ang = np.mod(np.arctan2(dy.astype(np.float64), dx.astype(np.float64)), 2 * np.pi)
    pos = np.where((sx >= -sy) & (sx > sy), sy / r,
          np.where(sx >= -sy, 2 - sx / r, np.where(sx <= sy, 4 - sy / r, 6 + sx / r)))
    want = np.mod(pos * np.pi / 4, 2 * np.pi)
    d = np.abs(ang - want)
    d = np.minimum(d, 2 * np.pi - d)
    keep = r > 1e-3
    assert d[keep].max() < 1e-4
    # uniform area: counts in 8 sectors x 4 equal-area rings
    ring = np.minimum((r ** 2 * 4).astype(int), 3)
    sect = np.minimum((ang / (np.pi / 4)).astype(int), 7)
    hist = np.bincount(ring * 8 + sect, minlength=32)
    exp = len(u1) / 32
    assert ((hist - exp) ** 2 / exp).sum() < 70  # 31 dof


def test_cosine_hemisphere_invariants_and_distribution():
    rng = np.random.default_rng(7)
    n = 100000
    u1 = rng.uniform(0, 1, n).astype(np.float32)
    u2 = rng.uniform(0, 1, n).astype(np.float32)
    N = rng.normal(size=(n, 3))
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    N = N.astype(np.float32)
    wi = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    orc.lib().orc_cosine_sample(n, orc.p(u1), orc.p(u2), orc.p(N), orc.p(wi), orc.p(pdf))
    assert np.allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-5)
    cos = (wi.astype(np.float64) * N).sum(1)
    assert cos.min() >= -1e-6
    # pdf is cos/pi of the local z (the sample before rotation)
    assert np.allclose(pdf, np.maximum(cos, 0) / np.pi, atol=2e-6)
    # cosine-weighted: cos^2 is uniform on [0, 1]
    hist, _ = np.histogram(cos ** 2, bins=20, range=(0, 1))
    exp = n / 20
    assert ((hist - exp) ** 2 / exp).sum() < 50
    # the same (u1, u2) with a fixed normal reproduces the disk sample in the tangent plane
    Nz = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    orc.lib().orc_cosine_sample(n, orc.p(u1), orc.p(u2), orc.p(Nz), orc.p(wi), orc.p(pdf))
    dx, dy = concentric(u1, u2)
    assert np.allclose(np.hypot(wi[:, 0], wi[:, 1]), np.hypot(dx, dy), atol=1e-5)


def camera_dirs(pos, lookat, up, fov, w, h, xs, ys):
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    out = np.zeros((len(xs), 3), np.float32)
    a = [np.asarray(v, np.float32) for v in (pos, lookat, up)]
    orc.lib().orc_camera_dirs(orc.p(a[0]), orc.p(a[1]), orc.p(a[2]), C.c_float(fov), w, h, len(xs),
                              orc.p(xs), orc.p(ys), orc.p(out))
    return out


@pytest.mark.parametrize("w,h,fov", [(64, 64, 90.0), (1920, 1080, 60.0), (100, 300, 45.0)])
def test_camera_closed_form(w, h, fov):
    """Looking down -z from the origin with up +y: pixel (x, y) maps to
    normalize((2x/W - 1) * a * t, (2y/H - 1) * t, -1), t = tan(fov/2)."""
    t = math.tan(math.radians(fov) / 2)
    a = w / h
    xs = np.array([0, w / 2, w, 0, w, 13.5, w - 1], np.float32)
    ys = np.array([0, h / 2, h, h, 0, 7.25, h - 1], np.float32)
    d = camera_dirs([0, 0, 0], [0, 0, -1], [0, 1, 0], fov, w, h, xs, ys)
    want = np.stack([(2 * xs / w - 1) * a * t, (2 * ys / h - 1) * t, -np.ones_like(xs)], 1).astype(np.float64)
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.abs(d - want).max() < 2e-6
    # an arbitrary frame: the image centre looks at `lookat`
    pos = np.array([2.0, 3.0, 9.0])
    look = np.array([-1.0, 0.5, 2.0])
    d = camera_dirs(pos, look, [0, 1, 0], fov, w, h, [w / 2], [h / 2])
    v = (look - pos) / np.linalg.norm(look - pos)
    assert np.abs(d[0] - v).max() < 2e-6


@pytest.mark.parametrize("w,h,spp,tiles,rows,last", [
    (512, 512, 1, 1, 512, 512),       # AO, 64 domains
    (1024, 1024, 8, 8, 128, 128),     # path tracing, 64 domains
    (1920, 1080, 16, 33, 33, 24),     # path tracing, 256 domains
    (3840, 2160, 16, 128, 17, 1),     # path tracing, 1024 domains
])
def test_blocking_tiles(w, h, spp, tiles, rows, last):
    """render/tile.cc:317-391 at 1,048,576 samples per tile."""
    n = orc.num_tiles(image_w=w, image_h=h, pixel_samples=spp, max_samples_per_rank=1 << 20)
    assert n == tiles
    est = -(-(w * h * spp) // (1 << 20))
    assert h // est == rows and h - (tiles - 1) * rows == last
