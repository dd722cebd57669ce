"""The C ABI's failure modes on the product path: every reference abort
(glog CHECK / LOG(FATAL)) that the boundary can meet becomes a status code
and a message, and the capacity path (pools grown and the frame rendered
again) gives the oracle's image."""
import math
import os

import numpy as np
import pytest

import orc
import scenes

pytestmark = pytest.mark.gpu

SPRAY_ERR_INVALID, SPRAY_ERR_IO, SPRAY_ERR_FORMAT, SPRAY_ERR_STATE = 1, 2, 3, 7
CAM = dict(pos=[1.0, 1.0, 1.0 + 2.0 * math.sqrt(3.0)], lookat=[1.0, 1.0, 1.0], up=[0.0, 1.0, 0.0])
PT = dict(image_w=64, image_h=64, pixel_samples=4, ao_mode=0, ao_samples=1, bounces=5, rr_start=3,
          max_samples_per_rank=4096)


def _ctx(**kw):
    import spray_amd as sa
    cfg = dict(has_camera=1, camera_pos=CAM["pos"], camera_lookat=CAM["lookat"], camera_up=CAM["up"],
               nthreads=4)
    cfg.update(PT)
    cfg.update(kw)
    return sa.Context(sa.default_config(**cfg))


def _code(fn, *a, **kw):
    import spray_amd as sa
    with pytest.raises(sa.SprayError) as e:
        fn(*a, **kw)
    return e.value.code, str(e.value)


def test_missing_ply_is_io_error(tmp_path):
    """io/ply_loader.cc:182 (LOG(FATAL) on a file that does not open)."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "s", 2, 1, 1, 8)
    os.remove(os.path.join(d, "s_d0001.ply"))
    with _ctx() as ctx:
        code, msg = _code(ctx.load_scene, os.path.join(d, "s.spray"), d)
        assert code == SPRAY_ERR_IO and "s_d0001.ply" in msg, msg
        code, _ = _code(ctx.trace_frame)  # nothing loaded: call-order error, no abort
        assert code == SPRAY_ERR_STATE


def test_missing_scene_file_and_bad_formats(tmp_path):
    """io/scene_loader.cc:67 (the .spray must open); a scene without domains, a
    truncated binary PLY and an unknown light are format errors."""
    with _ctx() as ctx:
        code, _ = _code(ctx.load_scene, str(tmp_path / "none.spray"), str(tmp_path))
        assert code == SPRAY_ERR_IO
        empty = tmp_path / "empty.spray"
        empty.write_text("light point 0 5 0 1 1 1\n# no domains\n")
        code, msg = _code(ctx.load_scene, str(empty), str(tmp_path))
        assert code == SPRAY_ERR_FORMAT, msg
        bad = tmp_path / "bad.spray"
        bad.write_text("light spot 0 5 0 1 1 1\n")
        code, msg = _code(ctx.load_scene, str(bad), str(tmp_path))
        assert code == SPRAY_ERR_FORMAT and "spot" in msg, msg
        q = scenes.quad(0.0, -1, 1, -1, 1)
        scenes.write_ply(str(tmp_path / "t.ply"), q["verts"], q["faces"], q["rgb"], binary=True)
        raw = (tmp_path / "t.ply").read_bytes()
        (tmp_path / "t.ply").write_bytes(raw[:-5])
        (tmp_path / "t.spray").write_text("light point 0 5 0 1 1 1\ndomain\nfile t.ply\n"
                                          "bound -1 0 -1 1 0 1\nvertex 4\nface 2\nmtl diffuse 1 1 1\n")
        code, msg = _code(ctx.load_scene, str(tmp_path / "t.spray"), str(tmp_path))
        assert code == SPRAY_ERR_FORMAT, msg


def test_bad_rays_and_arguments(tmp_path):
    """Rays with a NaN or all-zero direction (undefined for Embree) and bad
    domain ids, modes and cache sizes (lru_cache.cc:44 CHECK_NE(cache_size,
    0)) are invalid arguments; good rays in the same context still work."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "s", 2, 2, 2, 8)
    with _ctx() as ctx:
        ctx.load_scene(os.path.join(d, "s.spray"), d)
        org = np.full((4, 3), 1.0, np.float32)
        dirs = np.array([[0, 0, -1], [0, 1, 0], [0, 0, 0], [1, 0, 0]], np.float32)
        code, msg = _code(ctx.trace_rays, org, dirs, 0)
        assert code == SPRAY_ERR_INVALID and "ray 2" in msg, msg
        dirs[2] = [np.nan, 0.0, 1.0]
        code, msg = _code(ctx.trace_rays, org, dirs, 0, mode=1)
        assert code == SPRAY_ERR_INVALID and "ray 2" in msg, msg
        code, msg = _code(ctx.domain_list, org, dirs)
        assert code == SPRAY_ERR_INVALID and "ray 2" in msg, msg
        dirs[2] = [0.0, 0.0, 1.0]
        t, u, v, p = ctx.trace_rays(org, dirs, 0)
        assert p.shape == (4,)
        code, _ = _code(ctx.trace_rays, org, dirs, 8)
        assert code == SPRAY_ERR_INVALID
        code, _ = _code(ctx.trace_rays, org, dirs, 0, mode=5)
        assert code == SPRAY_ERR_INVALID
        code, _ = _code(ctx.set_cache_size, 0)
        assert code == SPRAY_ERR_INVALID
        code, _ = _code(ctx.set_schedule, 3)
        assert code == SPRAY_ERR_INVALID


def test_pool_regrow_matches_oracle(tmp_path):
    """Ray pools and queues started at 1/64 of the estimate overflow; the frame
    is rendered again with grown pools (spray_stats.regrows > 0) and equals the
    oracle's, for the plain and the batched schedule."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "g", 2, 2, 2, 16, rgb=0xB3B3B3, extra_lines="light rect 0.5 2.4 0.5 1 0 0 0 0 1 4 4 4")
    # 128x128x4 samples in 4 tiles: the pools' fixed slack of 4096 rays is
    # small against them
    kw = dict(PT, image_w=128, image_h=128, max_samples_per_rank=16384)
    ref = orc.Scene(os.path.join(d, "g.spray"), d, nthreads=4)
    rimg, rst = ref.render(nthreads=4, cam_pos=CAM["pos"], cam_lookat=CAM["lookat"], cam_up=CAM["up"], **kw)
    for schedule in (0, 1):
        with _ctx(schedule=schedule, **kw) as ctx:
            ctx.load_scene(os.path.join(d, "g.spray"), d)
            ctx.set_pool_headroom(1.0 / 64.0)
            img, st = ctx.trace_frame()
            img2, st2 = ctx.trace_frame()  # the grown pools stay
        err = np.abs(img[..., :3] - rimg[..., :3])
        print(f"schedule {schedule}: regrows {st['regrows']} then {st2['regrows']}, rmse "
              f"{float(np.sqrt((err ** 2).mean())):.3g}")
        assert st["regrows"] > 0 and st2["regrows"] == 0
        assert st["useful_rays"] == rst["useful_rays"] and st["final_hits"] == rst["final_hits"]
        assert float(np.sqrt((err ** 2).mean())) <= 1e-4 and float(err.max()) <= 1e-3
        assert np.abs(img2 - img).max() <= 1e-5


def test_path_tracing_without_lights(tmp_path):
    """The reference's PT shader loops over no lights (ooc_shader_pt.h:114):
    black image, bounces still traced; the same as the oracle."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "g", 2, 2, 2, 12, rgb=0xB3B3B3)
    ref = orc.Scene(os.path.join(d, "g.spray"), d, nthreads=4)
    rimg, rst = ref.render(nthreads=4, cam_pos=CAM["pos"], cam_lookat=CAM["lookat"], cam_up=CAM["up"], **PT)
    with _ctx(schedule=1) as ctx:
        ctx.load_scene(os.path.join(d, "g.spray"), d)
        img, st = ctx.trace_frame()
    assert float(np.abs(img[..., :3]).max()) == 0.0 and float(np.abs(rimg[..., :3]).max()) == 0.0
    assert st["useful_rays"] == rst["useful_rays"] > st["eye_rays"]
