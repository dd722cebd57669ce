"""GPU parity: the HIP path through the C ABI against the oracle on the same
inputs (closest hit, any hit, domain lists, normals, whole frames)."""
import math
import os

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu

C1_CAM = dict(pos=[2.0, 2.0, 2.0 + 2.0 * math.sqrt(3.0)], lookat=[2.0, 2.0, 2.0], up=[0.0, 1.0, 0.0])
RECT_LIGHT = "light rect 0.5 4.5 0.5 3 0 0 0 0 3 4 4 4"


@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    import spray_amd as sa
    d = str(tmp_path_factory.mktemp("c1"))
    sa.scenegen_write(d, "c1", 4, 4, 4, 64, rgb=0xFFFFFF)
    sa.scenegen_write(d, "c2", 4, 4, 4, 64, rgb=0xB3B3B3, extra_lines=RECT_LIGHT)
    return d


@pytest.fixture(scope="module")
def ctx_c1(c1):
    import spray_amd as sa
    cfg = sa.default_config(image_w=512, image_h=512, pixel_samples=1, ao_mode=1, ao_samples=8,
                            bounces=1, ao_max_dist=0.5, has_camera=1, camera_pos=C1_CAM["pos"],
                            camera_lookat=C1_CAM["lookat"], camera_up=C1_CAM["up"], nthreads=8)
    ctx = sa.Context(cfg)
    ctx.load_scene(os.path.join(c1, "c1.spray"), c1)
    yield ctx
    ctx.close()


@pytest.fixture(scope="module")
def ref_c1(c1):
    return orc.Scene(os.path.join(c1, "c1.spray"), c1)


def eye_rays(w, h, spp=1):
    ys, xs = np.mgrid[0:h, 0:w]
    xs = xs.ravel().astype(np.float32)
    ys = ys.ravel().astype(np.float32)
    dirs = np.zeros((w * h, 3), np.float32)
    orc.lib().orc_camera_dirs(orc.p(np.array(C1_CAM["pos"], np.float32)),
                              orc.p(np.array(C1_CAM["lookat"], np.float32)),
                              orc.p(np.array(C1_CAM["up"], np.float32)), 90.0, w, h, w * h,
                              orc.p(xs), orc.p(ys), orc.p(dirs))
    org = np.tile(np.array(C1_CAM["pos"], np.float32), (w * h, 1))
    return org, dirs


def test_closest_hit_c1_domain(ctx_c1, ref_c1):
    """All 262,144 C1 eye rays against single domains: primID bit-exact on
    >= 99.99% of rays, t within 1e-5 relative."""
    org, dirs = eye_rays(512, 512)
    for dom in (21, 42, 0):
        t, u, v, p = ctx_c1.trace_rays(org, dirs, dom)
        rt, ru, rv, rp = ref_c1.closest(dom, org, dirs)
        agree = (p == rp)
        print(f"domain {dom}: hits={int((rp >= 0).sum())} prim agreement {agree.mean():.6f}")
        assert agree.mean() >= 0.9999
        both = (p >= 0) & (rp >= 0) & agree
        assert np.all(np.abs(t[both] - rt[both]) <= 1e-5 * np.abs(rt[both]))
        # identical arithmetic: expected bit-exact
        assert np.array_equal(t[both], rt[both]) and np.array_equal(u[both], ru[both])


def test_closest_hit_random_rays(ctx_c1, ref_c1):
    rng = np.random.default_rng(7)
    n = 200_000
    org = rng.uniform(-0.5, 4.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dirs = d.astype(np.float32)
    for dom in (5, 63):
        t, u, v, p = ctx_c1.trace_rays(org, dirs, dom)
        rt, ru, rv, rp = ref_c1.closest(dom, org, dirs)
        assert (p == rp).mean() >= 0.9999
        both = (p >= 0) & (p == rp)
        assert np.array_equal(t[both], rt[both])


def test_any_hit(ctx_c1, ref_c1):
    rng = np.random.default_rng(3)
    n = 200_000
    org = rng.uniform(0.0, 4.0, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tfar = rng.uniform(0.01, 2.0, n).astype(np.float32)
    for dom in (21, 22):
        occ = ctx_c1.trace_rays(org, dirs, dom, mode=1, tfar=tfar)
        rocc = ref_c1.any_hit(dom, org, dirs, tfar)
        assert np.array_equal(occ, rocc)


def test_ploc_builder(c1, ctx_c1, ref_c1):
    """The PLOC builder gives the same hits as the default LBVH tree and the
    oracle: hits must not depend on the tree."""
    import spray_amd as sa
    rng = np.random.default_rng(11)
    n = 100_000
    org = rng.uniform(-0.5, 4.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tfar = rng.uniform(0.01, 2.0, n).astype(np.float32)
    with sa.Context(ctx_c1.cfg) as ctx:
        ctx.set_bvh_builder(1)
        ctx.load_scene(os.path.join(c1, "c1.spray"), c1)
        for dom in (5, 42):
            t, u, v, p = ctx.trace_rays(org, dirs, dom)
            t0, u0, v0, p0 = ctx_c1.trace_rays(org, dirs, dom)
            rt, ru, rv, rp = ref_c1.closest(dom, org, dirs)
            assert (p == rp).mean() >= 0.9999
            assert np.array_equal(p, p0)
            both = (p >= 0) & (p == rp)
            assert np.array_equal(t[both], rt[both])
            occ = ctx.trace_rays(org, dirs, dom, mode=1, tfar=tfar)
            assert np.array_equal(occ, ref_c1.any_hit(dom, org, dirs, tfar))


def test_domain_list(ctx_c1, ref_c1):
    rng = np.random.default_rng(11)
    n = 100_000
    org = rng.uniform(-1.0, 5.0, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    cnt, ids, ts = ctx_c1.domain_list(org, dirs, maxk=16)
    rc, rids, rts = ref_c1.domain_list(org, dirs, maxk=16)
    assert np.array_equal(cnt, rc)
    for i in range(n):
        k = min(cnt[i], 16)
        assert np.array_equal(ids[i, :k], rids[i, :k])
        assert np.array_equal(ts[i, :k], rts[i, :k])


def test_normals_bit_exact(ctx_c1, ref_c1):
    for dom in (0, 21, 63):
        n = ctx_c1.normals(dom)
        rn = ref_c1.mesh(dom)[3]
        assert np.array_equal(n, rn)


def _img_close(a, b):
    err = np.abs(a[..., :3] - b[..., :3])
    return float(np.sqrt((err ** 2).mean())), float(err.max())


@pytest.mark.parametrize("schedule", [0, 1])
def test_ao_frame_c1(ctx_c1, ref_c1, schedule):
    """C1 AO frame (512x512, 1 spp, 8 AO rays, max dist 0.5) vs the oracle."""
    ctx_c1.set_schedule(schedule)
    img, st = ctx_c1.trace_frame()
    rimg, rst = ref_c1.render(image_w=512, image_h=512, pixel_samples=1, ao_mode=1, ao_samples=8,
                              bounces=1, ao_max_dist=0.5, nthreads=8, cam_pos=C1_CAM["pos"],
                              cam_lookat=C1_CAM["lookat"], cam_up=C1_CAM["up"])
    rmse, mx = _img_close(img, rimg)
    print(f"schedule {schedule}: rmse {rmse:.3g} max {mx:.3g} useful {st['useful_rays']} "
          f"oracle {rst['useful_rays']} trace_ms {st['trace_ms']:.2f}")
    assert st["useful_rays"] == rst["useful_rays"]
    assert st["final_hits"] == rst["final_hits"]
    assert rmse <= 1e-4 and mx <= 1e-3


PT_SMALL = dict(image_w=128, image_h=128, pixel_samples=4, ao_mode=0, ao_samples=1, bounces=5,
                rr_start=3, max_samples_per_rank=16384)  # 4 blocking tiles of 32 rows


@pytest.fixture(scope="module")
def ref_pt_small(c1):
    ref = orc.Scene(os.path.join(c1, "c2.spray"), c1)
    return ref.render(nthreads=8, cam_pos=C1_CAM["pos"], cam_lookat=C1_CAM["lookat"],
                      cam_up=C1_CAM["up"], **PT_SMALL)


@pytest.mark.parametrize("schedule,tiles", [(0, 0), (1, 0), (1, 1), (1, 3)])
def test_pt_frame_small(c1, ref_pt_small, schedule, tiles):
    """Path tracing with the rect light, depth 5, RR from bounce 3, four
    blocking tiles and two speculation windows, vs the oracle; the batched
    form traces the tiles all together (0 = automatic), one by one, or in
    groups of 3 + 1, with the same result."""
    import spray_amd as sa
    cfg = sa.default_config(has_camera=1, camera_pos=C1_CAM["pos"], camera_lookat=C1_CAM["lookat"],
                            camera_up=C1_CAM["up"], nthreads=8, schedule=schedule, **PT_SMALL)
    with sa.Context(cfg) as ctx:
        ctx.load_scene(os.path.join(c1, "c2.spray"), c1)
        ctx.set_tiles_in_flight(tiles)
        img, st = ctx.trace_frame()
    rimg, rst = ref_pt_small
    rmse, mx = _img_close(img, rimg)
    print(f"PT schedule {schedule} tiles {tiles}: rmse {rmse:.3g} max {mx:.3g} "
          f"useful {st['useful_rays']} oracle {rst['useful_rays']}")
    assert st["tiles"] == 4
    assert st["useful_rays"] == rst["useful_rays"]
    assert st["final_hits"] == rst["final_hits"]
    assert rmse <= 1e-4 and mx <= 1e-3


def test_long_domain_lists(tmp_path):
    """Rays along a row of 24 domains cross more than 16 of them (the
    scheduler's score weighs the 16 nearest); the image centre's rays run
    along faces shared by domains.  Frames still match the oracle."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "row", 24, 1, 1, 10, rgb=0xB3B3B3, extra_lines="light point 12 3 0.5 1 1 1")
    cam = dict(pos=[-3.0, 0.5, 0.5], lookat=[24.0, 0.5, 0.5], up=[0.0, 1.0, 0.0])
    kw = dict(image_w=64, image_h=64, pixel_samples=2, ao_mode=0, ao_samples=1, bounces=3, fov=20.0,
              max_samples_per_rank=4096)
    cfg = sa.default_config(has_camera=1, camera_pos=cam["pos"], camera_lookat=cam["lookat"],
                            camera_up=cam["up"], nthreads=8, schedule=1, **kw)
    ref = orc.Scene(os.path.join(d, "row.spray"), d)
    rimg, rst = ref.render(nthreads=8, cam_pos=cam["pos"], cam_lookat=cam["lookat"], cam_up=cam["up"], **kw)
    with sa.Context(cfg) as ctx:
        ctx.load_scene(os.path.join(d, "row.spray"), d)
        n, ids, ts = ctx.domain_list(np.tile(np.array(cam["pos"], np.float32), (1, 1)),
                                     np.array([[1.0, 0.0, 0.0]], np.float32), 32)
        assert n[0] >= 20
        img, st = ctx.trace_frame()
    rmse, mx = _img_close(img, rimg)
    print(f"long lists: rmse {rmse:.3g} max {mx:.3g} useful {st['useful_rays']} oracle {rst['useful_rays']}")
    assert st["useful_rays"] == rst["useful_rays"]
    assert rmse <= 1e-4 and mx <= 1e-3


def test_odd_domain_lattice(tmp_path):
    """27 domains (3x3x3), boxes that fill their lattice cells: the domain
    grid lists each domain once and rays graze cell faces, edges and corners
    of neighbouring domains.  Path-traced frames (batched and plain) match
    the oracle."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "g27", 3, 3, 3, 24, rgb=0xB3B3B3,
                      extra_lines="light rect 0.5 3.4 0.5 2 0 0 0 0 2 4 4 4")
    cam = dict(pos=[1.5, 1.5, 1.5 + 3.0], lookat=[1.5, 1.5, 1.5], up=[0.0, 1.0, 0.0])
    kw = dict(image_w=64, image_h=64, pixel_samples=2, ao_mode=0, ao_samples=1, bounces=4,
              rr_start=3, max_samples_per_rank=4096)
    ref = orc.Scene(os.path.join(d, "g27.spray"), d)
    rimg, rst = ref.render(nthreads=8, cam_pos=cam["pos"], cam_lookat=cam["lookat"], cam_up=cam["up"], **kw)
    for schedule in (1, 0):
        cfg = sa.default_config(has_camera=1, camera_pos=cam["pos"], camera_lookat=cam["lookat"],
                                camera_up=cam["up"], nthreads=8, schedule=schedule, **kw)
        with sa.Context(cfg) as ctx:
            ctx.load_scene(os.path.join(d, "g27.spray"), d)
            img, st = ctx.trace_frame()
        rmse, mx = _img_close(img, rimg)
        print(f"27 domains, schedule {schedule}: rmse {rmse:.3g} max {mx:.3g} useful {st['useful_rays']}")
        assert st["useful_rays"] == rst["useful_rays"]
        assert st["final_hits"] == rst["final_hits"]
        assert rmse <= 1e-4 and mx <= 1e-3


def test_axis_parallel_rays(ctx_c1, ref_c1):
    """Rays with zero direction components and origins on grid planes (0 * inf
    in slab tests): closest hits, any hits and domain lists equal the oracle's."""
    rng = np.random.default_rng(21)
    n = 20000
    org = rng.uniform(0.0, 4.0, (n, 3))
    k = np.arange(n) % 3
    org[np.arange(n), k] = np.round(org[np.arange(n), k] * 64) / 64  # on a cell plane
    dirs = rng.normal(size=(n, 3))
    dirs[np.arange(n), k] = 0.0
    rows = np.arange(n // 4)
    dirs[rows, (k[rows] + 1) % 3] = 0.0  # some rays along an axis
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    org, dirs = org.astype(np.float32), dirs.astype(np.float32)
    cnt, ids, ts = ctx_c1.domain_list(org, dirs, 16)
    rcnt, rids, rts = ref_c1.domain_list(org, dirs, 16)
    assert np.array_equal(cnt, rcnt)
    m = np.arange(16)[None, :] < np.minimum(cnt, 16)[:, None]
    assert np.array_equal(ids[m], rids[m]) and np.array_equal(ts[m], rts[m])
    for dom in (0, 21, 42):
        t, u, v, p = ctx_c1.trace_rays(org, dirs, dom)
        rt, ru, rv, rp = ref_c1.closest(dom, org, dirs)
        assert (p >= 0).sum() > 100
        assert np.array_equal(p, rp) and np.array_equal(t, rt), dom
        occ = ctx_c1.trace_rays(org, dirs, dom, mode=1)
        assert np.array_equal(occ.astype(bool), rp >= 0), dom


def test_far_origins_at_vertices_and_edges(ctx_c1, ref_c1):
    """Rays from origins 2 to 4000 units away (so |org / dir| is large and
    the fused box test's pad carries the bound), aimed exactly at mesh
    vertices, at edge midpoints and at points one float step off them, and
    rays skimming the domain's box faces: closest hits and any hits equal the
    oracle's, bit for bit."""
    rng = np.random.default_rng(77)
    for dom in (0, 21, 63):
        v, f, _, _ = ref_c1.mesh(dom)
        v = v.reshape(-1, 3).astype(np.float64)
        f = f.reshape(-1, 3)
        nv, nf = 4000, 4000
        tv = v[rng.integers(0, len(v), nv)]
        fe = f[rng.integers(0, len(f), nf)]
        k = rng.integers(0, 3, nf)
        te = 0.5 * (v[fe[np.arange(nf), k]] + v[fe[np.arange(nf), (k + 1) % 3]])
        tgt = np.concatenate([tv, te]).astype(np.float32)
        off = np.nextafter(tgt, np.float32(np.inf))  # one float step off
        tgt = np.concatenate([tgt, off])
        n = len(tgt)
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        dist = np.array([2.0, 60.0, 4000.0])[rng.integers(0, 3, n)]
        org = (tgt.astype(np.float64) + u * dist[:, None]).astype(np.float32)
        d = tgt.astype(np.float64) - org.astype(np.float64)
        dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        # skimming rays: along a face of the domain box, just inside it
        _, _, lo, hi, _ = ref_c1.domain(dom)
        lo, hi = lo.astype(np.float64), hi.astype(np.float64)
        m = 4000
        so = rng.uniform(lo, hi, (m, 3))
        ax = rng.integers(0, 3, m)
        so[np.arange(m), ax] = np.where(rng.random(m) < 0.5, lo[ax], hi[ax])
        sd = rng.normal(size=(m, 3))
        sd[np.arange(m), ax] = rng.normal(size=m) * 1e-4
        sd /= np.linalg.norm(sd, axis=1, keepdims=True)
        so = so - sd * 50.0
        org = np.concatenate([org, so.astype(np.float32)])
        dirs = np.concatenate([dirs, sd.astype(np.float32)])
        t, uu, vv, p = ctx_c1.trace_rays(org, dirs, dom)
        rt, ru, rv, rp = ref_c1.closest(dom, org, dirs)
        # where the two traversals differ, and on a sample, brute force over
        # every triangle of the domain decides (it depends on no box test)
        check = np.flatnonzero((p != rp) | (t != rt))
        check = np.union1d(check, rng.choice(len(p), 300, replace=False))
        bt, bu, bv, bp = ref_c1.closest(dom, org[check], dirs[check], brute=True)
        print(f"domain {dom}: {(p >= 0).sum()} hits of {len(p)}, {int(((p != rp) | (t != rt)).sum())} "
              f"differ from the oracle's traversal")
        assert (p >= 0).sum() > len(p) // 4
        assert np.array_equal(p[check], bp) and np.array_equal(t[check], bt), dom
        assert np.array_equal(uu[check], bu) and np.array_equal(vv[check], bv), dom
        assert (p == rp).mean() >= 0.9999
        occ = ctx_c1.trace_rays(org, dirs, dom, mode=1)
        assert np.array_equal(occ[check], bp >= 0), dom
        assert (occ == (rp >= 0)).mean() >= 0.9999


@pytest.mark.parametrize("nx,schedule", [(8, 0), (8, 1), (5, 1)],
                         ids=["512-plain", "512-batched", "320-batched"])
def test_domain_grid_frame_matches_oracle(tmp_path, nx, schedule):
    """512 domains (8x8x8), the largest scene among the parity tests, and 320
    (5x8x8, mask rows of 5 words padded to 6): the domain-list pass walks the
    domain grid and hands the scatter a packed list of hit-domain ids for rays
    that hit at most 16 domains and a full mask for the others (the form for
    more than 256 domains); the frame matches the oracle's brute-force domain
    lists."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "g", nx, 8, 8, 8, rgb=0xB3B3B3,
                      extra_lines=f"light rect 1 1 8.5 0 6 0 {nx - 2} 0 0 4 4 4")
    cx = nx / 2.0
    cam = dict(camera_pos=[cx, 4.0, 18.0], camera_lookat=[cx, 4.0, 4.0], camera_up=[0.0, 1.0, 0.0])
    pt = dict(image_w=96, image_h=96, pixel_samples=2, ao_mode=0, ao_samples=1, bounces=5, rr_start=3,
              max_samples_per_rank=4608)
    ref = orc.Scene(os.path.join(d, "g.spray"), d, nthreads=8)
    rimg, rst = ref.render(nthreads=8, cam_pos=cam["camera_pos"], cam_lookat=cam["camera_lookat"],
                           cam_up=cam["camera_up"], **pt)
    with sa.Context(sa.default_config(has_camera=1, nthreads=8, schedule=schedule, **cam, **pt)) as ctx:
        ctx.load_scene(os.path.join(d, "g.spray"), d)
        assert ctx.num_domains == nx * 64
        img, st = ctx.trace_frame()
    err = np.abs(img[..., :3] - rimg[..., :3])
    assert st["useful_rays"] == rst["useful_rays"]
    assert st["final_hits"] == rst["final_hits"]
    assert float(np.sqrt((err ** 2).mean())) <= 1e-4 and float(err.max()) <= 1e-3
    assert float(rimg[..., :3].mean()) > 0.01  # a lit image
