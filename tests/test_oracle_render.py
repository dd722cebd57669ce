"""Pins the oracle's tracer (eye rays, AO and path-tracing shaders, shadow
far bound, rect light, Russian roulette, speculative scheduling) against
closed forms and against its own canonical (non-speculative) form (CPU only).

Reference: ooc/ooc_tracer.inl (frame loop), ooc/ooc_shader_ao.h:81-197,
ooc/ooc_shader_pt.h:93-227, render/light.h:41-89 (point and hemisphere
lights), render/reflection.h:202-214 (Blinn-Phong; the direct term carries no
1/pi).  The rect light, the AO/shadow far bound and Russian roulette are
additions; their closed forms are below.
"""
import math
import os

import numpy as np
import pytest

import orc
import scenes

DOWN_UP = [0.0, 0.0, -1.0]  # up vector for a camera looking straight down


def pinhole_dirs(pos, lookat, up, fov, w, h, xs, ys):
    """float64 pinhole camera: normalize(-w + (2x/W-1) hw u + (2y/H-1) hh v)."""
    pos, lookat, up = (np.asarray(a, np.float64) for a in (pos, lookat, up))
    W = (pos - lookat) / np.linalg.norm(pos - lookat)
    U = np.cross(up, W)
    U /= np.linalg.norm(U)
    V = np.cross(W, U)
    hh = math.tan(math.radians(fov) / 2)
    hw = hh * w / h
    d = -W[None] + ((2 * xs / w - 1) * hw)[:, None] * U[None] + ((2 * ys / h - 1) * hh)[:, None] * V[None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def floor_hits(cam, fov, w, h):
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    d = pinhole_dirs(cam, [cam[0], 0, cam[2]], DOWN_UP, fov, w, h, xs.ravel(), ys.ravel())
    t = -cam[1] / d[:, 1]
    return (np.asarray(cam)[None] + t[:, None] * d).reshape(h, w, 3)


def near_either(img, want_hi, want_lo, rtol):
    """True where img matches the expectation with kd = c/255 or (c-1)/255
    (the colour interpolation truncates a float sum that can fall just below c)."""
    return (np.abs(img - want_hi) <= rtol * np.abs(want_hi) + 1e-7) | \
           (np.abs(img - want_lo) <= rtol * np.abs(want_lo) + 1e-7)


def render(tmp_path, domains, lights, **kw):
    path = scenes.write_scene(str(tmp_path), "s", domains, lights)
    s = orc.Scene(path, str(tmp_path), nthreads=4)
    kw.setdefault("nthreads", 4)
    return s.render(**kw)


def test_ao_furnace_open_plane(tmp_path):
    """An unoccluded diffuse plane under AO returns exactly its colour: each
    AO sample carries kd (cos/pi)/(N pdf) with pdf = cos/pi."""
    c = (200, 100, 0)
    floor = scenes.quad(0.0, -50, 50, -50, 50, c=c)
    img, st = render(tmp_path, [floor], ["light point 0 5 0 1 1 1"], image_w=48, image_h=40,
                     cam_pos=[0.2, 3.0, 0.1], cam_lookat=[0.2, 0.0, 0.1], cam_up=DOWN_UP, fov=60.0,
                     pixel_samples=1, ao_mode=1, ao_samples=8, bounces=1)
    for k in range(3):
        ok = near_either(img[..., k], c[k] / 255.0, max(c[k] - 1, 0) / 255.0, 2e-5)
        assert ok.all(), k
    assert st["eye_rays"] == 48 * 40
    assert st["useful_rays"] == 48 * 40 * (1 + 8)


@pytest.mark.parametrize("max_dist,fraction", [(0.5, 1.0), (2.0, 0.25), (math.inf, 0.0)])
def test_ao_max_distance_under_ceiling(tmp_path, max_dist, fraction):
    """Floor at y=0 under a ceiling at y=1: an AO ray at polar angle theta
    reaches the ceiling after 1/cos(theta).  With far bound D it escapes when
    cos(theta) < 1/D; for cosine-weighted directions P(cos < x) = x^2, so the
    mean AO is kd * min(1, 1/D)^2 (D = 0.5: 1, D = 2: 1/4, D = inf: 0)."""
    floor = scenes.quad(0.0, -50, 50, -50, 50, c=(255, 255, 255))
    ceil = scenes.quad(1.0, -50, 50, -50, 50, c=(255, 255, 255), up=False)
    img, st = render(tmp_path, [floor, ceil], ["light point 0 5 0 1 1 1"], image_w=64, image_h=64,
                     cam_pos=[0.0, 0.9, 0.0], cam_lookat=[0.0, 0.0, 0.0], cam_up=DOWN_UP, fov=60.0,
                     pixel_samples=1, ao_mode=1, ao_samples=16, bounces=1, ao_max_dist=max_dist)
    got = img[..., 0]
    if fraction == 1.0:
        assert near_either(got, 1.0, 254.0 / 255.0, 2e-5).all()
    elif fraction == 0.0:
        # only rays within ~1 degree of the horizon pass the 50-unit ceiling's edge
        assert got.mean() < 1e-3
    else:
        assert abs(got.mean() - fraction) < 0.01  # 65536 samples: sigma ~ 0.0017


def test_pt_point_light_closed_form(tmp_path):
    """One bounce, point light over a plane: pixel = I kd cos(theta) (no 1/r^2,
    no 1/pi: light.h:41-60 and reflection.h:202-214)."""
    c = (179, 230, 20)
    I = np.array([0.8, 0.5, 1.5])
    L = np.array([0.3, 2.0, -0.2])
    cam, fov, w, h = [0.0, 3.0, 0.0], 60.0, 40, 32
    floor = scenes.quad(0.0, -50, 50, -50, 50, c=c)
    img, st = render(tmp_path, [floor], ["light point " + " ".join(repr(float(x)) for x in (*L, *I))], image_w=w, image_h=h,
                     cam_pos=cam, cam_lookat=[cam[0], 0, cam[2]], cam_up=DOWN_UP, fov=fov,
                     pixel_samples=1, ao_mode=0, ao_samples=1, bounces=1)
    p = floor_hits(cam, fov, w, h)
    toL = L[None, None] - p
    cos = toL[..., 1] / np.linalg.norm(toL, axis=-1)
    for k in range(3):
        hi = I[k] * c[k] / 255.0 * cos
        lo = I[k] * (c[k] - 1) / 255.0 * cos
        assert near_either(img[..., k], hi, lo, 2e-5).all(), k
    assert st["shades"] == w * h and st["useful_rays"] == 2 * w * h


def rect_irradiance(px, pz, hgt, x0, x1, z0, z1):
    """integral of cos(theta_p) cos(theta_l) / r^2 over the rectangle
    [x0,x1] x [z0,z1] at height hgt, parallel to the receiving plane
    (= pi x the point-to-rectangle form factor; corner formula + signed
    decomposition at the foot point)."""
    def g(a, b):
        X, Y = a / hgt, b / hgt
        return 0.5 * (X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X)) +
                      Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y)))

    def S(x, z):
        return math.copysign(1, x - px) * math.copysign(1, z - pz) * g(abs(x - px), abs(z - pz))
    return S(x1, z1) - S(x0, z1) - S(x1, z0) + S(x0, z0)


def test_rect_light_form_factor_closed_form(tmp_path):
    """A rect light of radiance Le facing a plane: E[pixel] = kd Le pi F with
    F the point-to-rectangle form factor."""
    c = (255, 255, 255)
    Le = 2.0
    hgt = 1.0
    cam, fov, w, h = [0.0, 0.8, 0.0], 70.0, 48, 48
    floor = scenes.quad(0.0, -50, 50, -50, 50, c=c)
    light = f"light rect -0.5 {hgt} -0.25 1 0 0 0 0 0.5 {Le} {Le} {Le}"
    img, st = render(tmp_path, [floor], [light], image_w=w, image_h=h, cam_pos=cam,
                     cam_lookat=[0, 0, 0], cam_up=DOWN_UP, fov=fov, pixel_samples=1, ao_mode=0,
                     ao_samples=32, bounces=1)
    p = floor_hits(cam, fov, w, h)
    want = np.array([[Le * rect_irradiance(q[0], q[2], hgt, -0.5, 0.5, -0.25, 0.25) for q in row] for row in p])
    got = img[..., 0].astype(np.float64)
    assert abs(got.mean() / want.mean() - 1) < 0.01
    rel = (got - want) / want
    assert np.sqrt((rel ** 2).mean()) < 0.15  # per-pixel Monte Carlo noise, 32 light samples
    # the estimator is unbiased per pixel: the error averages out over rows
    assert np.abs(rel.mean(1)).max() < 0.06


def sandwich(tmp_path, light):
    floor = scenes.quad(0.0, -20, 20, -20, 20, c=(200, 200, 200))
    ceil = scenes.quad(1.0, -20, 20, -20, 20, c=(200, 200, 200), up=False)
    path = scenes.write_scene(str(tmp_path), "sw", [floor, ceil], [light])
    return orc.Scene(path, str(tmp_path), nthreads=4)


def test_point_light_shadow_rays_are_unbounded(tmp_path):
    """Shadow rays toward a point light have tfar = inf (render/rays.h:398-399):
    a ceiling above the light shadows the whole floor below it."""
    s = sandwich(tmp_path, "light point 0.1 0.5 0.2 0.3 0.3 0.3")
    kw = dict(image_w=32, image_h=32, cam_pos=[0.0, 0.9, 0.0], cam_lookat=[0, 0, 0], cam_up=DOWN_UP,
              fov=80.0, pixel_samples=1, ao_mode=0, ao_samples=1, bounces=1, nthreads=4)
    img, st = s.render(**kw)
    assert st["shades"] == 32 * 32 and img[..., :3].max() == 0.0


def test_russian_roulette_keeps_the_mean(tmp_path):
    s = sandwich(tmp_path, "light rect -0.5 0.5 -0.5 1 0 0 0 0 1 1 1 1")
    kw = dict(image_w=32, image_h=32, cam_pos=[0.0, 0.9, 0.0], cam_lookat=[0, 0, 0], cam_up=DOWN_UP,
              fov=80.0, pixel_samples=16, ao_mode=0, ao_samples=1, bounces=6, nthreads=4)
    a, sa = s.render(rr_start=0, **kw)
    b, sb = s.render(rr_start=1, **kw)
    ma, mb = a[..., 0].mean(), b[..., 0].mean()
    assert abs(mb / ma - 1) < 0.02
    assert sb["shades"] < sa["shades"]  # roulette terminated paths
    # without roulette the multi-bounce part is strictly positive
    c, _ = s.render(rr_start=0, **dict(kw, bounces=1))
    assert ma > 1.1 * c[..., 0].mean()


@pytest.fixture(scope="module")
def mc_scene(tmp_path_factory):
    import spray_amd as sa
    d = str(tmp_path_factory.mktemp("mc"))
    sa.scenegen_write(d, "mc", 2, 2, 2, 12, rgb=0xB3B3B3, extra_lines="light rect 0.5 2.5 0.5 1 0 0 0 0 1 4 4 4")
    return orc.Scene(os.path.join(d, "mc.spray"), d, nthreads=4)


CAM2 = dict(cam_pos=[1.0, 1.0, 1.0 + 2.0 * math.sqrt(3.0)], cam_lookat=[1.0, 1.0, 1.0], cam_up=[0, 1, 0])


@pytest.mark.parametrize("mode", ["ao", "pt"])
def test_speculative_equals_canonical(mc_scene, mode):
    """The speculative schedule (per-domain queues, VBuf history, pending
    window restarts) renders exactly what a ray-by-ray tracer renders."""
    if mode == "ao":
        kw = dict(pixel_samples=1, ao_mode=1, ao_samples=4, bounces=1, ao_max_dist=0.5)
    else:
        kw = dict(pixel_samples=2, ao_mode=0, ao_samples=1, bounces=6, rr_start=3)
    kw.update(CAM2, image_w=48, image_h=48, nthreads=4, max_samples_per_rank=2048)
    a, sa = mc_scene.render(canonical=0, **kw)
    b, sb = mc_scene.render(canonical=1, **kw)
    assert sa["tiles"] > 1
    assert sa["useful_rays"] == sb["useful_rays"]
    assert np.abs(a - b).max() <= 1e-6 * max(1.0, np.abs(b).max())
    assert a[..., :3].max() > 0
    # speculation visits more (ray, domain) pairs than the useful rays
    assert sa["radiance_visits"] >= sa["final_hits"]


def test_speculative_order_independent_of_threads(mc_scene):
    kw = dict(CAM2, image_w=32, image_h=32, pixel_samples=2, ao_mode=0, ao_samples=1, bounces=5,
              rr_start=3, max_samples_per_rank=1024)
    a, sa = mc_scene.render(nthreads=1, **kw)
    b, sb = mc_scene.render(nthreads=4, **kw)
    assert sa["useful_rays"] == sb["useful_rays"]
    assert np.abs(a - b).max() <= 1e-6
