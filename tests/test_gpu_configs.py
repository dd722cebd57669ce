"""GPU tests on the benchmark configurations' own scenes.

C3: 256 domains on an 8x8x4 lattice with log-uniform per-domain triangle
budgets (10k..2M, seeded field frequencies).  The full 120M-triangle scene
at 1920x1080, 16 spp, depth 6 is checked through size-independent
properties: plain, batched and learned schedules give the same useful rays,
final hits and image.  The same lattice and budget law at a 3M-triangle
total is rendered against the oracle at a reduced image size.
C4: the C2 scene (64 domains, 64^3 cells each) with the learned schedule
against the oracle at a reduced image size.
C5: the 16x8x8, 160^3-cell in-situ scene generated and remeshed on the GPU
at phase 0.05: sampled domains bit-exact against the CPU generator, and a
full-size frame (3840x2160, 16 spp, depth 8) batched vs learned with the
predictor trained on the frame's own hit log."""
import math
import os

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu

RECT3 = "light rect 1 8.5 0.5 6 0 0 0 0 3 4 4 4"
CAM3 = dict(pos=[4.0, 4.0, 8.0], lookat=[4.0, 4.0, 2.0], up=[0.0, 1.0, 0.0])
C3GEN = dict(per_domain_random=True, seed=0xC3, rgb=0xB3B3B3)


def _close(a, b):
    err = np.abs(a[..., :3].astype(np.float64) - b[..., :3])
    return float(np.sqrt((err ** 2).mean())), float(err.max())


def _cfg(cam, **kw):
    import spray_amd as sa
    return sa.default_config(has_camera=1, camera_pos=cam["pos"], camera_lookat=cam["lookat"],
                             camera_up=cam["up"], nthreads=16, **kw)


def _train(ctx, log_schedule, samples=4_000_000, epochs=2):
    ctx.set_schedule(log_schedule)
    ctx.predictor_log(True, samples)
    ctx.trace_frame(want_image=False)
    ctx.predictor_log(False)
    assert ctx.predictor_num_samples() > 10_000
    ctx.predictor_train(epochs, 4096, 2e-3)


def test_c3_reduced_vs_oracle(tmp_path):
    """C3's lattice and budget law at 3M triangles, 128x72, 4 spp, depth 6:
    plain, batched and learned frames equal the oracle's."""
    import spray_amd as sa
    d = str(tmp_path)
    tris = sa.scenegen_write(d, "c3r", 8, 8, 4, 0, tri_total=3e6, extra_lines=RECT3, **C3GEN)
    assert 2.5e6 < tris < 3.5e6
    kw = dict(image_w=128, image_h=72, pixel_samples=4, ao_mode=0, ao_samples=1, bounces=6, rr_start=3,
              max_samples_per_rank=128 * 18 * 4)
    ref = orc.Scene(os.path.join(d, "c3r.spray"), d, nthreads=16)
    rimg, rst = ref.render(nthreads=16, cam_pos=CAM3["pos"], cam_lookat=CAM3["lookat"], cam_up=CAM3["up"], **kw)
    assert rst["tiles"] == 4 and float(rimg[..., :3].mean()) > 1e-3
    with sa.Context(_cfg(CAM3, schedule=0, **kw)) as ctx:
        ctx.load_scene(os.path.join(d, "c3r.spray"), d)
        assert ctx.num_domains == 256
        frames = {}
        for name, s in (("plain", 0), ("batched", 1)):
            ctx.set_schedule(s)
            frames[name] = ctx.trace_frame()
        _train(ctx, 0)
        ctx.set_schedule(2)
        frames["learned"] = ctx.trace_frame()
    for name, (img, st) in frames.items():
        rmse, mx = _close(img, rimg)
        print(f"C3 reduced {name}: rmse {rmse:.3g} max {mx:.3g} useful {st['useful_rays']} "
              f"oracle {rst['useful_rays']} radiance visits {st['radiance_visits']}")
        assert st["useful_rays"] == rst["useful_rays"], name
        assert st["final_hits"] == rst["final_hits"], name
        assert rmse <= 1e-4 and mx <= 1e-3, name


def test_c3_full_scene_schedules_agree():
    """The full C3 workload (about 120M triangles, 1920x1080, 16 spp, depth 6):
    plain, batched and learned frames agree (useful rays, final hits, image);
    a few domains of the GPU-generated scene equal the CPU generator's."""
    import spray_amd as sa
    kw = dict(image_w=1920, image_h=1080, pixel_samples=16, ao_mode=0, ao_samples=1, bounces=6, rr_start=3)
    with sa.Context(_cfg(CAM3, schedule=0, **kw)) as ctx:
        ctx.scenegen_load("c3", 8, 8, 4, 0, tri_total=120e6, extra_lines=RECT3, **C3GEN)
        tris = sum(ctx.mesh_size(i)[1] for i in range(ctx.num_domains))
        assert 1.0e8 < tris < 1.5e8
        for dom in (0, 97, 255):
            v, f, c = ctx.domain_mesh(dom)
            cv, cf, cc = sa.scenegen_domain(dom, 8, 8, 4, 0, tri_total=120e6, **C3GEN)
            assert np.array_equal(v, cv) and np.array_equal(f, cf) and np.array_equal(c, cc), dom
        frames = {}
        ctx.set_schedule(0)
        frames["plain"] = ctx.trace_frame()
        ctx.set_schedule(1)
        frames["batched"] = ctx.trace_frame()
        _train(ctx, 1, samples=16_000_000, epochs=1)
        ctx.set_schedule(2)
        frames["learned"] = ctx.trace_frame()
    pimg, pst = frames["plain"]
    assert pst["tiles"] == 33 and float(pimg[..., :3].mean()) > 1e-3
    for name in ("batched", "learned"):
        img, st = frames[name]
        rmse, mx = _close(img, pimg)
        print(f"C3 {name} vs plain: rmse {rmse:.3g} max {mx:.3g} useful {st['useful_rays']} "
              f"radiance visits {st['radiance_visits']} (plain {pst['radiance_visits']}) "
              f"trace {st['trace_ms']:.1f} ms (plain {pst['trace_ms']:.1f} ms)")
        assert st["useful_rays"] == pst["useful_rays"] and st["final_hits"] == pst["final_hits"]
        assert rmse <= 1e-4 and mx <= 1e-3


def test_c4_learned_vs_oracle(tmp_path):
    """C4: the C2 scene (64 domains, 64^3 cells) and the predictor trained on
    plain-frame hit logs; learned frame at 256x256, 4 spp, depth 5 equals the
    oracle's frame, and the predictor saves visits over the batched form."""
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "c2", 4, 4, 4, 64, rgb=0xB3B3B3, extra_lines="light rect 0.5 4.5 0.5 3 0 0 0 0 3 4 4 4")
    cam = dict(pos=[2.0, 2.0, 2.0 + 2.0 * math.sqrt(3.0)], lookat=[2.0, 2.0, 2.0], up=[0.0, 1.0, 0.0])
    kw = dict(image_w=256, image_h=256, pixel_samples=4, ao_mode=0, ao_samples=1, bounces=5, rr_start=3,
              max_samples_per_rank=65536)
    ref = orc.Scene(os.path.join(d, "c2.spray"), d, nthreads=16)
    rimg, rst = ref.render(nthreads=16, cam_pos=cam["pos"], cam_lookat=cam["lookat"], cam_up=cam["up"], **kw)
    with sa.Context(_cfg(cam, schedule=1, **kw)) as ctx:
        ctx.load_scene(os.path.join(d, "c2.spray"), d)
        _, bst = ctx.trace_frame()
        _train(ctx, 0)
        ctx.set_schedule(2)
        img, st = ctx.trace_frame()
    rmse, mx = _close(img, rimg)
    print(f"C4 learned vs oracle: rmse {rmse:.3g} max {mx:.3g} useful {st['useful_rays']} "
          f"radiance visits {st['radiance_visits']} (batched {bst['radiance_visits']})")
    assert st["useful_rays"] == rst["useful_rays"] and st["final_hits"] == rst["final_hits"]
    assert rmse <= 1e-4 and mx <= 1e-3
    assert st["radiance_visits"] < bst["radiance_visits"]


def test_c5_remesh_and_frame():
    """C5: 1024 domains of 160^3 cells generated on the GPU, remeshed at phase
    0.05; sampled domains equal the CPU generator's meshes bit for bit; the
    full-size frame (3840x2160, 16 spp, depth 8) is the same for the batched
    and the learned schedule."""
    import spray_amd as sa
    cam = dict(pos=[8.0, 4.0, 4.0 + 0.5 * math.sqrt(16.0 ** 2 + 8.0 ** 2 + 8.0 ** 2)],
               lookat=[8.0, 4.0, 4.0], up=[0.0, 1.0, 0.0])
    kw = dict(image_w=3840, image_h=2160, pixel_samples=16, ao_mode=0, ao_samples=1, bounces=8, rr_start=3)
    gen = dict(rgb=0xB3B3B3)
    with sa.Context(_cfg(cam, schedule=1, **kw)) as ctx:
        ctx.scenegen_load("c5", 16, 8, 8, 160, extra_lines="light rect 2 1 8.5 0 6 0 12 0 0 4 4 4", **gen)
        assert ctx.num_domains == 1024
        ctx.remesh(0.05)
        for dom in (0, 333, 700, 1023):
            v, f, c = ctx.domain_mesh(dom)
            cv, cf, cc = sa.scenegen_domain(dom, 16, 8, 8, 160, phase=0.05, **gen)
            assert len(f) > 300_000
            assert np.array_equal(v, cv) and np.array_equal(f, cf) and np.array_equal(c, cc), dom
        bimg, bst = ctx.trace_frame()
        _train(ctx, 1, samples=4_000_000, epochs=1)
        ctx.set_schedule(2)
        limg, lst = ctx.trace_frame()
    rmse, mx = _close(limg, bimg)
    print(f"C5 learned vs batched: rmse {rmse:.3g} max {mx:.3g} useful {lst['useful_rays']} "
          f"trace {lst['trace_ms']:.1f} / {bst['trace_ms']:.1f} ms")
    assert lst["useful_rays"] == bst["useful_rays"] and lst["final_hits"] == bst["final_hits"]
    assert rmse <= 1e-4 and mx <= 1e-3
    assert float(bimg[..., :3].mean()) > 1e-3
