"""Pins the oracle's scene formats, triangle test, BVH, domain list,
interpolation and normals against the reference's fixtures, closed forms
and brute force (CPU only).

Reference: io/scene_loader.cc, io/ply_loader.cc (formats; the fixtures are
examples/wavelet/wavelet_example.spray:1-22 with the vertex/face/bound lines
written by apps/ply_header_reader.cc:100-103, and examples/wavelets64),
render/aabb.h:139-169 (slab test), render/rays.h:57-93 +
render/wbvh_embree.cc:126-148 (domain list), render/trimesh_buffer.cc:227-360
(normals, colour interpolation with integer truncation).
"""
import math
import os
import re

import numpy as np
import pytest

import orc
import scenes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ply_read(path):
    import ctypes as C
    nv, nf, col = C.c_longlong(), C.c_longlong(), C.c_int()
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    rc = orc.lib().orc_ply_read(str(path).encode(), C.byref(nv), C.byref(nf), C.byref(col), orc.p(lo), orc.p(hi))
    if rc != 0:
        raise orc.OracleError(orc.lib().orc_last_error().decode())
    return nv.value, nf.value, bool(col.value), lo, hi


def spray_declared(path):
    """The domain blocks of a .spray file as written (file, vertex, face, bound, translate)."""
    doms, cur = [], None
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "domain":
            cur = {}
            doms.append(cur)
        elif cur is not None and t[0] in ("file", "vertex", "face", "bound", "translate"):
            cur[t[0]] = t[1:]
    return doms


def test_wavelet_example_headers_match_declared():
    """ASCII PLY with colour: counts and the 16-digit bounds the reference's
    header reader wrote into the scene file are reproduced exactly."""
    for d in spray_declared(os.path.join(GOLD, "wavelet_example.spray")):
        nv, nf, col, lo, hi = ply_read(os.path.join(GOLD, d["file"][0]))
        assert (nv, nf) == (int(d["vertex"][0]), int(d["face"][0]))
        assert col
        want = np.array([float(x) for x in d["bound"]], np.float64).astype(np.float32)
        assert np.array_equal(lo, want[:3]) and np.array_equal(hi, want[3:])
    assert [ply_read(os.path.join(GOLD, f))[:2] for f in ("wavelet1.ply", "wavelet0.ply")] == \
        [(958, 1748), (909, 1684)]


def test_wavelet_binary_ply_and_lattice_scene():
    nv, nf, col, lo, hi = ply_read(os.path.join(GOLD, "wavelet.ply"))
    assert (nv, nf, col) == (2840, 5480, True)
    assert np.allclose(lo, [-10, -10, -10], atol=1e-6) and np.allclose(hi, [10, 9.324713, 10], atol=1e-6)
    path = os.path.join(GOLD, "wavelets64.spray")
    s = orc.Scene(path, GOLD, parse_only=True)
    assert s.num_domains == 64 and s.num_lights == 1
    decl = spray_declared(path)
    for i, d in enumerate(decl):
        nv_, nf_, lo_, hi_, _ = s.domain(i)
        assert (nv_, nf_) == (2840, 5480)
        b = np.array([float(x) for x in d["bound"]], np.float32)
        tr = np.array([float(x) for x in d["translate"]], np.float32)
        assert np.array_equal(lo_, b[:3] + tr) and np.array_equal(hi_, b[3:] + tr)
    wlo, whi = s.bound()
    assert np.allclose(wlo, [-10, -10, -10]) and np.allclose(whi, [70, 9.324713 + 3 * 19.324713, 70], atol=1e-4)
    # translate steps of the lattice: 20 / 19.324713 / 20
    ts = np.array([[float(x) for x in d["translate"]] for d in decl])
    assert set(np.round(np.unique(ts[:, 0]), 4)) == {0, 20, 40, 60}
    assert np.allclose(np.unique(ts[:, 1]), [0, 19.324713, 38.649426, 57.974139], atol=1e-4)


def test_binary_and_ascii_ply_agree(tmp_path):
    rng = np.random.default_rng(2)
    v = rng.uniform(-3, 3, (50, 3)).astype(np.float32)
    f = rng.integers(0, 50, (80, 3))
    c = rng.integers(0, 256, (50, 3))
    scenes.write_ply(str(tmp_path / "a.ply"), v, f, c)
    scenes.write_ply(str(tmp_path / "b.ply"), v, f, c, binary=True)
    a = ply_read(tmp_path / "a.ply")
    b = ply_read(tmp_path / "b.ply")
    assert a[:3] == b[:3] == (50, 80, True)
    assert np.array_equal(a[3], v.min(0)) and np.array_equal(b[4], v.max(0))


@pytest.mark.parametrize("text,msg", [
    ("domain\nfile x.ply\nbogus 1\n", "unknown tag"),
    ("light point 1 2 3\ndomain\n", "light point"),
    ("light spot 0 0 0 1 1 1\ndomain\n", "unknown light"),
    ("domain\nmtl velvet 1 1 1\n", "unknown material"),
    ("domain\nrotate w 10\n", "invalid axis"),
    ("# nothing\n", "no domain"),
    ("domain\nfile a.ply\nvertex 3\nface 1\n", "invalid bound"),
])
def test_malformed_scene_files_are_errors(tmp_path, text, msg):
    p = tmp_path / "bad.spray"
    p.write_text(text)
    with pytest.raises(orc.OracleError, match=msg):
        orc.Scene(str(p), str(tmp_path), parse_only=True)


def test_malformed_ply_is_error(tmp_path):
    (tmp_path / "x.ply").write_text("plx\nformat ascii 1.0\n")
    with pytest.raises(orc.OracleError):
        ply_read(tmp_path / "x.ply")
    with pytest.raises(orc.OracleError):
        ply_read(tmp_path / "missing.ply")


def tri_hits(org, dirs, tris, tnear=0.0, tfar=math.inf):
    import ctypes as C
    org = np.ascontiguousarray(org, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    tris = np.ascontiguousarray(tris, np.float32)
    n = len(org)
    t = np.zeros(n, np.float32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    hit = np.zeros(n, np.int32)
    orc.lib().orc_tri_intersect(n, orc.p(org), orc.p(dirs), orc.p(tris), C.c_float(tnear), C.c_float(tfar),
                                orc.p(t), orc.p(u), orc.p(v), orc.p(hit))
    return hit.astype(bool), t, u, v


def test_triangle_closed_form_unit_square():
    """Rays straight down onto the unit square z=0 split along its diagonal:
    t = height, barycentrics equal the in-plane coordinates."""
    rng = np.random.default_rng(4)
    n = 4000
    xy = rng.uniform(0.01, 0.99, (n, 2))
    h = rng.uniform(0.5, 5.0, n)
    org = np.stack([xy[:, 0], xy[:, 1], h], 1)
    dirs = np.tile([0.0, 0.0, -1.0], (n, 1))
    lower = np.tile([0, 0, 0, 1, 0, 0, 1, 1, 0], (n, 1))  # (0,0) (1,0) (1,1)
    upper = np.tile([0, 0, 0, 1, 1, 0, 0, 1, 0], (n, 1))  # (0,0) (1,1) (0,1)
    h1, t1, u1, v1 = tri_hits(org, dirs, lower)
    h2, t2, u2, v2 = tri_hits(org, dirs, upper)
    x, y = org[:, 0].astype(np.float32), org[:, 1].astype(np.float32)
    below = x > y
    assert np.array_equal(h1, below) and np.array_equal(h2, ~below)
    assert np.allclose(t1[h1], org[h1, 2].astype(np.float32), rtol=1e-6)
    # lower triangle: p = v0 + u (v1 - v0) + v (v2 - v0) -> u = x - y, v = y
    assert np.allclose(u1[h1], (x - y)[h1], atol=2e-6) and np.allclose(v1[h1], y[h1], atol=2e-6)
    # upper: u = x, v = y - x
    assert np.allclose(u2[h2], x[h2], atol=2e-6) and np.allclose(v2[h2], (y - x)[h2], atol=2e-6)
    # tnear / tfar bounds
    hh, _, _, _ = tri_hits(org, dirs, lower, tfar=0.4)
    assert not hh.any()


def test_triangle_vs_float64_random():
    """Random rays and triangles against a float64 Moller-Trumbore: hit/miss
    agrees away from edges, t agrees to float rounding."""
    rng = np.random.default_rng(8)
    n = 20000
    tris = rng.uniform(-1, 1, (n, 9))
    org = rng.uniform(-3, 3, (n, 3))
    tgt = tris[:, 0:3] * 0.4 + tris[:, 3:6] * 0.3 + tris[:, 6:9] * 0.3 + rng.normal(0, 0.5, (n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tris32, org32, d32 = tris.astype(np.float32), org.astype(np.float32), d.astype(np.float32)
    hit, t, u, v = tri_hits(org32, d32, tris32, tnear=1e-3)
    a, b, c = (tris32[:, 0:3].astype(np.float64), tris32[:, 3:6].astype(np.float64),
               tris32[:, 6:9].astype(np.float64))
    o, dd = org32.astype(np.float64), d32.astype(np.float64)
    e1, e2 = b - a, c - a
    pv = np.cross(dd, e2)
    det = (e1 * pv).sum(1)
    tv = o - a
    uu = (tv * pv).sum(1) / det
    qv = np.cross(tv, e1)
    vv = (dd * qv).sum(1) / det
    tt = (e2 * qv).sum(1) / det
    ref = (uu >= 0) & (vv >= 0) & (uu + vv <= 1) & (tt > 1e-3)
    margin = np.minimum(np.minimum(np.abs(uu), np.abs(vv)), np.abs(1 - uu - vv))
    clear = (margin > 1e-4) & (np.abs(det) > 1e-6)
    assert np.array_equal(hit[clear], ref[clear])
    both = hit & ref & clear
    assert both.sum() > 1000
    assert np.allclose(t[both], tt[both], rtol=1e-5, atol=1e-6)
    assert np.allclose(u[both], uu[both], atol=1e-4) and np.allclose(v[both], vv[both], atol=1e-4)


def test_watertight_closed_box(tmp_path):
    """Rays from inside a closed box, including rays aimed exactly at its
    vertices and edge midpoints, always hit it (no leaks through shared edges)."""
    box = scenes.cube((0, 0, 0), (1, 1, 1))
    path = scenes.write_scene(str(tmp_path), "box", [box], ["light point 0 5 0 1 1 1"])
    s = orc.Scene(path, str(tmp_path))
    rng = np.random.default_rng(9)
    corners = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float64)
    mids = np.array([[0.5, 0, 0], [0.5, 1, 1], [0, 0.5, 1], [1, 0.5, 0], [0, 1, 0.5], [1, 0, 0.5],
                     [0.5, 0.5, 0], [0.5, 0.5, 1], [0, 0.5, 0.5], [1, 0.5, 0.5]], np.float64)
    targets = np.concatenate([corners, mids, rng.uniform(0, 1, (2000, 3))])
    org = np.tile([0.5, 0.5, 0.5], (len(targets), 1)) + rng.uniform(-0.2, 0.2, (len(targets), 3)) * \
        (np.arange(len(targets)) >= 18)[:, None]
    d = targets - org
    faces = np.abs(d) > 0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t, u, v, prim = s.closest(0, org.astype(np.float32), d.astype(np.float32), tnear=1e-4)
    assert (prim >= 0).all()
    assert faces.any()
    # brute force agrees on the reported triangle and t
    tb, _, _, pb = s.closest(0, org.astype(np.float32), d.astype(np.float32), tnear=1e-4, brute=True)
    assert np.array_equal(prim, pb) and np.array_equal(t, tb)


@pytest.fixture(scope="module")
def wavelet_scene():
    return orc.Scene(os.path.join(GOLD, "wavelet_example.spray"), GOLD, nthreads=4)


def random_rays_at(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    org = c + rng.normal(size=(n, 3)) * ext
    tgt = rng.uniform(lo, hi, (n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


def test_bvh_closest_equals_brute_force(wavelet_scene):
    s = wavelet_scene
    for dom in range(s.num_domains):
        _, _, lo, hi, _ = s.domain(dom)
        org, d = random_rays_at(lo, hi, 3000, dom)
        t, u, v, p = s.closest(dom, org, d)
        tb, ub, vb, pb = s.closest(dom, org, d, brute=True)
        assert (p >= 0).mean() > 0.3
        assert np.array_equal(p, pb) and np.array_equal(t, tb) and np.array_equal(u, ub)


def test_any_hit_equals_closest_before_tmax(wavelet_scene):
    s = wavelet_scene
    _, _, lo, hi, _ = s.domain(0)
    org, d = random_rays_at(lo, hi, 4000, 21)
    t, _, _, p = s.closest(0, org, d)
    rng = np.random.default_rng(3)
    tmax = rng.uniform(0, 40, len(org)).astype(np.float32)
    occ = s.any_hit(0, org, d, tmax)
    want = (p >= 0) & (t < tmax)
    assert np.array_equal(occ, want)


def test_domain_list_vs_float64_slabs():
    s = orc.Scene(os.path.join(GOLD, "wavelets64.spray"), GOLD, parse_only=True)
    nd = s.num_domains
    boxes = np.array([np.concatenate(s.domain(i)[2:4]) for i in range(nd)], np.float64)
    rng = np.random.default_rng(6)
    n = 3000
    wlo, whi = s.bound()
    org = rng.uniform(wlo - 20, whi + 20, (n, 3))
    inside = rng.uniform(wlo, whi, (n // 3, 3))
    org[: n // 3] = inside  # origins inside domains: negative entry distance
    tgt = rng.uniform(wlo, whi, (n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org32, d32 = org.astype(np.float32), d.astype(np.float32)
    cnt, ids, ts = s.domain_list(org32, d32, maxk=64)
    o, dd = org32.astype(np.float64), d32.astype(np.float64)
    checked = 0
    for i in range(n):
        with np.errstate(divide="ignore", invalid="ignore"):
            t0 = (boxes[:, 0:3] - o[i]) / dd[i]
            t1 = (boxes[:, 3:6] - o[i]) / dd[i]
        tmin = np.minimum(t0, t1).max(1)
        tmax = np.maximum(t0, t1).min(1)
        hit = (tmin <= tmax) & (tmax > 1e-3)
        # skip near-ties and grazing cases (float rounding decides those)
        gap = np.abs(tmax - tmin)[hit]
        srt = np.sort(tmin[hit])
        if (gap < 1e-3).any() or (np.diff(srt) < 1e-3).any() or (np.abs(tmax - 1e-3) < 1e-3).any():
            continue
        want = np.flatnonzero(hit)[np.argsort(tmin[hit], kind="stable")]
        assert cnt[i] == len(want)
        assert list(ids[i, :cnt[i]]) == list(want)
        assert np.allclose(ts[i, :cnt[i]], tmin[want], rtol=1e-5, atol=1e-4)
        if cnt[i] and i < n // 3:
            assert ts[i, 0] <= 0.0  # the origin's own domain first
        checked += 1
    assert checked > 0.9 * n


def test_colour_interpolation_truncates(tmp_path):
    """Vertex colours are interpolated in float and truncated to integers
    (render/trimesh_buffer.cc:345-350 + reflection.h:241-262)."""
    tri = dict(verts=[[0, 0, 0], [1, 0, 0], [0, 1, 0]], faces=[[0, 1, 2]],
               rgb=[[0, 10, 179], [255, 11, 179], [0, 13, 179]])
    path = scenes.write_scene(str(tmp_path), "tri", [tri], ["light point 0 5 0 1 1 1"])
    s = orc.Scene(path, str(tmp_path))
    u = np.array([0.0, 0.5, 0.25, 1.0, 0.0, 0.3], np.float32)
    v = np.array([0.0, 0.0, 0.25, 0.0, 1.0, 0.3], np.float32)
    col, _ = s.interpolate(0, np.zeros(6, np.int32), u, v)
    r, g, b = (col >> 16) & 255, (col >> 8) & 255, col & 255
    # red: 255 u -> 0, 127 (127.5 truncated), 63 (63.75), 255, 0, 76 (76.5)
    assert list(r) == [0, 127, 63, 255, 0, 76]
    # green: 10 w + 11 u + 13 v
    assert list(g) == [10, 10, 11, 11, 13, 11]
    # blue: constant 179 gives 179 or 178 (float sum below the integer)
    assert set(b) <= {178, 179} and b[0] == 179


def test_tetrahedron_vertex_normals(tmp_path):
    """Area-weighted vertex normals of a regular tetrahedron point from the
    centroid through each vertex."""
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float32)
    f = [[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]]
    # orientation: outward
    tet = dict(verts=v, faces=f, rgb=[[255, 255, 255]] * 4)
    path = scenes.write_scene(str(tmp_path), "tet", [tet], ["light point 0 5 0 1 1 1"])
    s = orc.Scene(path, str(tmp_path))
    vv, ff, rgb, n = s.mesh(0)
    for i in range(4):
        c = n[i] / np.linalg.norm(n[i])
        want = v[i] / np.linalg.norm(v[i])
        assert np.allclose(c, want, atol=1e-6), i
    # interpolation at a vertex returns that vertex's (unnormalised) normal
    _, ns = s.interpolate(0, np.array([0], np.int32), np.array([0.0], np.float32), np.array([0.0], np.float32))
    assert np.allclose(ns[0], n[f[0][0]])


def test_axis_parallel_rays_on_box_planes(tmp_path):
    """Rays with a zero direction component whose origin lies on a grid plane
    of the mesh (0 * inf = NaN in a slab test): the BVH still finds exactly
    the brute-force hits, and the domain list still holds every domain the
    ray passes through (deviation Q13: NaN-safe slabs)."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "row", 6, 1, 1, 10, rgb=0xB3B3B3, extra_lines="light point 3 3 0.5 1 1 1")
    s = orc.Scene(os.path.join(d, "row.spray"), d)
    rng = np.random.default_rng(12)
    n = 3000
    org = np.zeros((n, 3), np.float32)
    dirs = np.zeros((n, 3), np.float32)
    for i in range(n):
        k = i % 3                      # the axis the ray is parallel to
        o = rng.uniform(0.0, 1.0, 3)
        o[0] = rng.uniform(0.0, 6.0)
        o[k] = np.round(o[k] * 10) / 10  # on a cell plane of the 10^3 grid
        dv = rng.normal(size=3)
        dv[k] = 0.0
        org[i] = o
        dirs[i] = dv / np.linalg.norm(dv)
    assert (dirs == 0).sum() >= n
    mism = 0
    for dom in range(6):
        t, _, _, p = s.closest(dom, org, dirs)
        tb, _, _, pb = s.closest(dom, org, dirs, brute=True)
        assert np.array_equal(p, pb) and np.array_equal(t, tb), dom
        mism += int((p >= 0).sum())
        occ = s.any_hit(dom, org, dirs, np.inf)
        assert np.array_equal(occ, pb >= 0), dom
    assert mism > 100
    # a ray along the row inside the plane z = 0.5 crosses all six domains
    cnt, ids, ts = s.domain_list(np.array([[-1.0, 0.5, 0.5]], np.float32),
                                 np.array([[1.0, 0.0, 0.0]], np.float32))
    assert cnt[0] == 6 and list(ids[0, :6]) == [0, 1, 2, 3, 4, 5]
    # along the face y = 1 shared by two rows: only the upper row (half-open
    # domain membership keeps exact ties between domains out)
    sa.scenegen_write(d, "two", 3, 2, 1, 10, rgb=0xB3B3B3, extra_lines="light point 1 3 0.5 1 1 1")
    s2 = orc.Scene(os.path.join(d, "two.spray"), d, parse_only=True)
    assert s2.domain(3)[2][1] == 1.0 and s2.domain(0)[3][1] == 1.0  # the rows meet at y = 1
    cnt, ids, ts = s2.domain_list(np.array([[-1.0, 1.0, 0.5]], np.float32),
                                  np.array([[1.0, 0.0, 0.0]], np.float32))
    assert cnt[0] == 3 and list(ids[0, :3]) == [3, 4, 5]
