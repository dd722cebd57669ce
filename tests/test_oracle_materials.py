"""Pins the oracle's path-tracing branches that the diffuse scenes never run
against float64 closed forms (CPU only): the Blinn-Phong specular term with
ks > 0 (reflection.h:202-214), the hemisphere light (light.h:67-89), and the
mirror, glass and transmission BSDFs with total internal reflection
(ooc_shader_pt.h:177-214, reflection.h:126-173,264-364).  The scenes and
their closed forms are in material_scenes.py."""
import math

import numpy as np
import pytest

import material_scenes as ms
import orc
import scenes

DETERMINISTIC = [k for k, v in ms.all_scenes().items() if not v.get("stochastic")]


def oracle_render(tmp_path, sc, nthreads=4):
    path = scenes.write_scene(str(tmp_path), "m", sc["doms"], sc["lights"])
    ref = orc.Scene(path, str(tmp_path), nthreads=nthreads)
    cam = sc["cam"]
    return ref.render(nthreads=nthreads, cam_pos=cam, cam_lookat=[cam[0], 0.0, cam[2]],
                      cam_up=ms.DOWN_UP, **sc["kw"])


def matches_closed_form(img, sc, rtol=1e-4):
    """Per pixel and channel: the closed form with kd = c/255 or (c-1)/255
    (the colour interpolation truncates a float sum that can land just below
    c); exact zeros where the closed form is zero."""
    c = np.asarray(sc["c"], np.float64)
    hi, valid = sc["expect"](c / 255.0)
    lo, _ = sc["expect"](np.maximum(c - 1, 0) / 255.0)
    got = img[..., :3].astype(np.float64)
    ok = (np.abs(got - hi) <= rtol * np.abs(hi) + 1e-7) | (np.abs(got - lo) <= rtol * np.abs(lo) + 1e-7)
    return ok | ~valid[..., None], hi, valid


@pytest.mark.parametrize("name", DETERMINISTIC)
def test_closed_form(tmp_path, name):
    sc = ms.all_scenes()[name]
    img, st = oracle_render(tmp_path, sc)
    ok, want, valid = matches_closed_form(img, sc)
    bad = ~ok.all(axis=-1)
    print(f"{name}: {int(valid.sum())} pixels compared, {int(bad.sum())} off, "
          f"lit {float((want.max(-1) > 0).mean()):.2f}, mean {float(img[..., :3].mean()):.4g}")
    assert valid.mean() > 0.8
    assert (want.max(-1) > 0)[valid].mean() > 0.1  # the branch under test contributes
    assert not bad.any()
    if "useful" in sc:
        assert st["useful_rays"] == sc["useful"]


def test_glass_normal_incidence(tmp_path):
    """At normal incidence a 1.0 -> 1.5 interface transmits 1 - 0.04: the centre
    pixel's ray (through the image centre, straight down) reaches the floor
    with weight 0.96 (R0 = ((1 - 1.5) / (1 + 1.5))^2)."""
    sc = ms.dielectric_transmit_scene("glass", (1.0, 1.5), 40.0, True)
    img, _ = oracle_render(tmp_path, sc)
    w, h = sc["kw"]["image_w"], sc["kw"]["image_h"]
    p = np.array([0.0, 0.0, 0.0])
    L = np.array([6.0, 0.5, 0.0])
    cos_f = L[1] / np.linalg.norm(L - p)
    kd = np.asarray(sc["c"], np.float64) / 255.0
    I = np.array([1.2, 0.9, 0.7])
    ratio = img[h // 2, w // 2, :3] / (kd * I * cos_f)
    assert np.allclose(ratio, 0.96, rtol=2e-4) or np.allclose(
        img[h // 2, w // 2, :3] / ((np.asarray(sc["c"]) - 1) / 255.0 * I * cos_f), 0.96, rtol=2e-4)


def test_critical_angle(tmp_path):
    """Glass seen from its dense side (1.5 -> 1.0): transmitted light reaches the
    floor only inside the cone sin(theta) < 1/1.5 (41.8 degrees); outside it the
    ray is totally reflected and escapes, so those pixels are black."""
    sc = ms.dielectric_transmit_scene("glass", (1.5, 1.0), 120.0, True)
    img, _ = oracle_render(tmp_path, sc)
    w, h = sc["kw"]["image_w"], sc["kw"]["image_h"]
    d = ms.down_rays(sc["cam"], sc["kw"]["fov"], w, h)
    theta = np.degrees(np.arccos(-d[:, 1])).reshape(h, w)
    crit = math.degrees(math.asin(1.0 / 1.5))
    lum = img[..., :3].sum(-1)
    assert (lum[theta > crit + 0.5] == 0).all()
    inner = theta < crit - 0.5
    _, valid = sc["expect"](np.asarray(sc["c"]) / 255.0)
    want, _ = sc["expect"](np.asarray(sc["c"]) / 255.0)
    lit_inner = inner & valid & (want.max(-1) > 0)
    assert lit_inner.sum() > 50 and (lum[lit_inner] > 0).all()


def test_total_internal_reflection_weight(tmp_path):
    """Beyond the critical angle all light is reflected (fr = 1): the reflected
    pixel is 1/|cos_i| times the ceiling term, for glass and transmission
    alike; inside the cone transmission reflects nothing and glass reflects
    the Fresnel fraction."""
    g = ms.dielectric_reflect_scene("glass", (1.5, 1.0), 120.0, True)
    t = ms.dielectric_reflect_scene("transmission", (1.5, 1.0), 120.0, True)
    (tmp_path / "g").mkdir()
    (tmp_path / "t").mkdir()
    gimg, _ = oracle_render(tmp_path / "g", g)
    timg, _ = oracle_render(tmp_path / "t", t)
    w, h = g["kw"]["image_w"], g["kw"]["image_h"]
    d = ms.down_rays(g["cam"], g["kw"]["fov"], w, h)
    theta = np.degrees(np.arccos(-d[:, 1])).reshape(h, w)
    crit = math.degrees(math.asin(1.0 / 1.5))
    outer = theta > crit + 0.5
    inner = theta < crit - 0.5
    assert outer.sum() > 100 and inner.sum() > 100
    # beyond the critical angle both materials reflect everything
    assert np.allclose(gimg[outer][:, :3], timg[outer][:, :3], rtol=1e-6, atol=0)
    # inside it transmission reflects nothing, glass a Fresnel fraction < 1
    assert (timg[inner][:, :3] == 0).all()
    assert (gimg[inner][:, :3].sum(-1) > 0).mean() > 0.9


def test_hemisphere_light_half_ceiling(tmp_path):
    """Stochastic closed form: a half-plane ceiling at height 1 over x > 0
    shadows a floor point at x0 with probability (1 + x0/sqrt(1 + x0^2))/2
    under cosine-distributed light directions.  The mean over the image and
    over each half (x0 < 0, x0 > 0) matches within 5 sigma."""
    sc = ms.hemisphere_half_ceiling_scene()
    img, st = oracle_render(tmp_path, sc)
    kd = np.asarray(sc["c"], np.float64) / 255.0
    want, x0 = sc["expect"](kd)
    got = img[..., 0].astype(np.float64)
    S = sc["kw"]["ao_samples"]
    unit = kd[0] * math.pi  # value of one unoccluded pixel
    for mask in (np.ones_like(x0, bool), x0 < 0, x0 > 0):
        p = want[..., 0][mask].mean() / unit
        sigma = math.sqrt(p * (1 - p) / (mask.sum() * S))
        print(f"mean lit fraction {got[mask].mean() / unit:.4f} expected {p:.4f} (sigma {sigma:.4f})")
        assert abs(got[mask].mean() / unit - p) < 5 * sigma + 1e-4
    # the lit side is brighter: a sign error in the frame would swap the halves
    assert got[x0 < -0.2].mean() > got[x0 > 0.2].mean() * 1.5


def test_fresnel_helper_known_values():
    """The test's own Fresnel: R0 = 0.04 for 1.0 -> 1.5; zero p-reflectance at
    Brewster's angle atan(1.5); 1 beyond the critical angle from the dense
    side; energy-consistent limits at grazing incidence."""
    assert math.isclose(float(ms.fresnel(np.array(1.0), 1.0, 1.5)), 0.04, rel_tol=1e-12)
    tb = math.atan(1.5)
    ti, tt = tb, math.asin(math.sin(tb) / 1.5)
    assert abs(math.tan(ti - tt) / math.tan(ti + tt)) < 1e-12  # r_p vanishes
    rs = -math.sin(ti - tt) / math.sin(ti + tt)
    assert math.isclose(float(ms.fresnel(np.array(math.cos(tb)), 1.0, 1.5)), rs * rs / 2, rel_tol=1e-9)
    assert float(ms.fresnel(np.array(math.cos(math.radians(45))), 1.5, 1.0)) == 1.0
    assert float(ms.fresnel(np.array(1e-6), 1.0, 1.5)) > 0.999
