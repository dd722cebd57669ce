"""The batched form's autotuner and the parameters it chooses between: every
(queue chunk, waves per block) pair renders the oracle's image; the tuner
picks one of them on its own tuning frame, records it in its cache under the
device/frame/scene key and reuses it without tuning again."""
import math
import os

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu

CAM = dict(pos=[1.5, 1.5, 1.5 + 3.0], lookat=[1.5, 1.5, 1.5], up=[0.0, 1.0, 0.0])
KW = dict(image_w=64, image_h=64, pixel_samples=2, ao_mode=0, ao_samples=1, bounces=5, rr_start=3,
          max_samples_per_rank=4096)


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    import spray_amd as sa
    d = str(tmp_path_factory.mktemp("tune"))
    sa.scenegen_write(d, "g", 3, 3, 3, 16, rgb=0xB3B3B3, extra_lines="light rect 0.5 3.4 0.5 2 0 0 0 0 2 4 4 4")
    ref = orc.Scene(os.path.join(d, "g.spray"), d, nthreads=8)
    rimg, rst = ref.render(nthreads=8, cam_pos=CAM["pos"], cam_lookat=CAM["lookat"], cam_up=CAM["up"], **KW)
    return d, rimg, rst


def _ctx(d):
    import spray_amd as sa
    cfg = sa.default_config(has_camera=1, camera_pos=CAM["pos"], camera_lookat=CAM["lookat"],
                            camera_up=CAM["up"], nthreads=8, schedule=1, **KW)
    ctx = sa.Context(cfg)
    ctx.load_scene(os.path.join(d, "g.spray"), d)
    return ctx


def test_every_pair_matches_oracle(scene):
    d, rimg, rst = scene
    with _ctx(d) as ctx:
        for chunk in (8, 16, 32, 64):
            for tile in (1, 2, 4):
                ctx.set_batch_params(chunk, tile)
                img, st = ctx.trace_frame()
                err = np.abs(img[..., :3] - rimg[..., :3])
                assert st["useful_rays"] == rst["useful_rays"], (chunk, tile)
                assert st["final_hits"] == rst["final_hits"], (chunk, tile)
                assert float(np.sqrt((err ** 2).mean())) <= 1e-4 and float(err.max()) <= 1e-3, (chunk, tile)


def test_autotune_cache(scene, tmp_path, monkeypatch):
    d, rimg, rst = scene
    cache = tmp_path / "cache" / "autotune.txt"
    monkeypatch.setenv("SPRAY_AUTOTUNE_CACHE", str(cache))
    with _ctx(d) as ctx:
        chunk, tile = ctx.autotune()
        assert chunk in (8, 16, 32, 64) and tile in (1, 2, 4)
        lines = cache.read_text().splitlines()
        assert len(lines) == 1
        key, c2, t2, ms = lines[0].split()
        assert (int(c2), int(t2)) == (chunk, tile) and float(ms) > 0
        assert "spp2" in key and "64x64" in key  # the key names the frame it tunes for
        img, st = ctx.trace_frame()  # the frame itself is rendered at full samples
        assert st["useful_rays"] == rst["useful_rays"]
    with _ctx(d) as ctx:  # a new context finds the key and does not tune again
        assert ctx.autotune() == (chunk, tile)
        assert len(cache.read_text().splitlines()) == 1
        ctx.set_camera([1.5, 1.4, 4.5], CAM["lookat"], CAM["up"], 90.0)
        assert ctx.autotune() == (chunk, tile)  # the camera is not part of the key
    # another scene is another key
    import spray_amd as sa
    d2 = str(tmp_path / "other")
    os.makedirs(d2)
    sa.scenegen_write(d2, "g", 3, 3, 3, 12, rgb=0xB3B3B3, extra_lines="light rect 0.5 3.4 0.5 2 0 0 0 0 2 4 4 4")
    with _ctx(d2) as ctx:
        ctx.autotune()
        assert len(cache.read_text().splitlines()) == 2
