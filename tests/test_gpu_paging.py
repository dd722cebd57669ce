"""GPU tests of the paged domain pool (a domain cache smaller than the scene,
render/lru_cache.cc semantics): frames match the oracle, residency and
hit/miss counts follow the restated LRU cache, queries on non-resident
domains still answer exactly."""
import math
import os
import sys

import numpy as np
import pytest

import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from lru_ref import LruRef  # noqa: E402

pytestmark = pytest.mark.gpu

CAM = dict(pos=[2.0, 2.0, 2.0 + 2.0 * math.sqrt(3.0)], lookat=[2.0, 2.0, 2.0], up=[0.0, 1.0, 0.0])
RECT_LIGHT = "light rect 0.5 4.5 0.5 3 0 0 0 0 3 4 4 4"
PT = dict(image_w=128, image_h=128, pixel_samples=4, ao_mode=0, ao_samples=1, bounces=5,
          rr_start=3, max_samples_per_rank=16384)


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    import spray_amd as sa
    d = str(tmp_path_factory.mktemp("paged"))
    sa.scenegen_write(d, "c2", 4, 4, 4, 64, rgb=0xB3B3B3, extra_lines=RECT_LIGHT)
    return d


@pytest.fixture(scope="module")
def ref(scene):
    return orc.Scene(os.path.join(scene, "c2.spray"), scene)


def _ctx(scene, cache, schedule=0):
    import spray_amd as sa
    cfg = sa.default_config(has_camera=1, camera_pos=CAM["pos"], camera_lookat=CAM["lookat"],
                            camera_up=CAM["up"], nthreads=8, schedule=schedule, **PT)
    ctx = sa.Context(cfg)
    ctx.set_cache_size(cache)
    ctx.load_scene(os.path.join(scene, "c2.spray"), scene)
    return ctx


def test_paged_frame_matches_oracle(scene, ref):
    """Path tracing with 8 of 64 domains in HBM (plain schedule, domains paged
    in per pass) gives the oracle's image and counts."""
    rimg, rst = ref.render(nthreads=8, cam_pos=CAM["pos"], cam_lookat=CAM["lookat"], cam_up=CAM["up"], **PT)
    with _ctx(scene, 8) as ctx:
        img, st = ctx.trace_frame()
        cap, lru, hits, misses, nbytes = ctx.cache_state()
    err = np.abs(img[..., :3] - rimg[..., :3])
    rmse, mx = float(np.sqrt((err ** 2).mean())), float(err.max())
    print(f"paged: rmse {rmse:.3g} max {mx:.3g} hits {st['cache_hits']} misses {st['cache_misses']} "
          f"bytes {st['cache_bytes']}")
    assert st["useful_rays"] == rst["useful_rays"]
    assert st["final_hits"] == rst["final_hits"]
    assert rmse <= 1e-4 and mx <= 1e-3
    assert cap == 8 and len(lru) == 8
    assert st["cache_misses"] > 0 and st["cache_bytes"] > 0
    assert (hits, misses, nbytes) == (st["cache_hits"], st["cache_misses"], st["cache_bytes"])


def test_paged_lru_sequence_and_hits(scene, ref):
    """Each spray_trace_rays call is one domain access: residency order and
    hit/miss counts follow the reference's LRU cache, and hits on domains
    that were paged out and back in match the oracle."""
    rng = np.random.default_rng(5)
    n = 20000
    org = rng.uniform(-0.5, 4.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    with _ctx(scene, 4) as ctx:
        cap, lru0, _, _, _ = ctx.cache_state()
        assert cap == 4 and lru0 == [60, 61, 62, 63]  # the last domains built stay resident
        model = LruRef(64, 4)
        for dom in lru0:
            model.load(dom)
        model.hits = model.misses = 0
        for dom in (0, 61, 5, 0, 62, 7, 8, 0, 63, 5, 21):
            t, u, v, p = ctx.trace_rays(org, dirs, dom)
            model.load(dom)
            cap, lru, hits, misses, _ = ctx.cache_state()
            assert lru == model.lru_order(), (dom, lru, model.lru_order())
            assert (hits, misses) == (model.hits, model.misses)
            if dom in (0, 21):
                rt, ru, rv, rp = ref.closest(dom, org, dirs)
                assert (p == rp).mean() >= 0.9999
                both = (p >= 0) & (p == rp)
                assert np.array_equal(t[both], rt[both])


def test_paged_normals(scene, ref):
    with _ctx(scene, 4) as ctx:
        for dom in (0, 33, 63):
            assert np.array_equal(ctx.normals(dom), ref.mesh(dom)[3])


@pytest.mark.parametrize("cache", [8, 3])
def test_paged_batched_frame_matches_oracle(scene, ref, cache):
    """The batched schedule over a partial cache traces chunks of consecutive
    scheduled domains (half the cache each, the next chunk loading meanwhile,
    alternate passes in reverse) and gives the oracle's image and counts."""
    rimg, rst = ref.render(nthreads=8, cam_pos=CAM["pos"], cam_lookat=CAM["lookat"], cam_up=CAM["up"], **PT)
    with _ctx(scene, cache, schedule=1) as ctx:
        img, st = ctx.trace_frame()
    err = np.abs(img[..., :3] - rimg[..., :3])
    rmse, mx = float(np.sqrt((err ** 2).mean())), float(err.max())
    assert st["useful_rays"] == rst["useful_rays"]
    assert st["final_hits"] == rst["final_hits"]
    assert rmse <= 1e-4 and mx <= 1e-3
    assert st["cache_misses"] > 0


def test_paged_learned_needs_whole_scene(scene):
    import spray_amd as sa
    with _ctx(scene, 8, schedule=2) as ctx:
        with pytest.raises(sa.SprayError):
            ctx.trace_frame()
