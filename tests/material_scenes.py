"""Scenes with closed-form images for the path-tracing shader's branches that
the diffuse benchmark scenes never reach: the Blinn-Phong specular term
(ks > 0), the hemisphere ("light diffuse") area light, and the delta BSDFs
(mirror, glass, transmission) including total internal reflection.

Reference semantics (ooc/ooc_shader_pt.h:93-227, render/reflection.h:
126-173,202-364, render/light.h:41-89):
  * direct light, point light: Lin * I * (kd cos + ks (n.h)^s) / pdf, pdf = 1
    (no 1/pi, no 1/r^2); hemisphere light: cosine-sampled directions, S
    samples, Lin * L * kd cos / (pdf S) with pdf = cos/pi, so an unoccluded
    point receives kd L pi;
  * delta bounce: cos_i = dot(wo, Ns) with the *unnormalised* interpolated
    vertex normal Ns, clamped to [-1, 1]; reflection weight fr / |cos_i|,
    transmission weight (1 - fr) / |cos_i|; fr = Fresnel (glass), 1 (mirror),
    0 or 1 (transmission: refraction or total internal reflection);
  * delta surfaces get no direct light.

Every expectation here is computed in float64 from textbook optics, not from
the shader's formulas: Snell's law in angle form, the Fresnel equations in
their angle form (r_s = -sin(ti - tt)/sin(ti + tt), r_p = tan(ti - tt)/
tan(ti + tt)) with the normal-incidence limit ((n1 - n2)/(n1 + n2))^2, the
critical angle asin(n2/n1), mirror reflection by flipping the normal
component, and the projected-solid-angle integral of a half-plane occluder.

Sheets whose vertex normals must have unit length ("unit" sheets) are built
from unit right triangles with unshared vertices: each vertex normal is its
one face's cross product, of length 1, so Ns is a unit vector and cos_i is
the geometric cosine.  A "large" sheet is one quad of two big triangles: its
vertex normals are thousands long, cos_i clamps to +-1, and the reference
then uses normal incidence for every pixel (fr = R0, weight fr, refraction
direction normalize(eta d + (eta - 1) n)).
"""
import math

import numpy as np

DOWN_UP = [0.0, 0.0, -1.0]  # camera up vector when looking straight down


# ---------------------------------------------------------------- geometry
def quad(y, x0, x1, z0, z1, c=(255, 255, 255), up=True, mtl=None):
    v = [[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]]
    f = [[0, 2, 1], [0, 3, 2]] if up else [[0, 1, 2], [0, 2, 3]]
    d = dict(verts=v, faces=f, rgb=[c] * 4)
    if mtl:
        d["mtl"] = mtl
    return d


def unit_sheet(y, x0, x1, z0, z1, up=True, mtl=None, c=(255, 255, 255)):
    """Plane y = const over integer extents, two unit right triangles per unit
    square, no shared vertices: every vertex normal is (0, +-1, 0) exactly."""
    verts, faces = [], []
    for i in range(x0, x1):
        for k in range(z0, z1):
            a, b, cc, d = [i, y, k], [i + 1, y, k], [i + 1, y, k + 1], [i, y, k + 1]
            for tri in ((a, cc, b), (a, d, cc)) if up else ((a, b, cc), (a, cc, d)):
                n = len(verts)
                verts += [list(p) for p in tri]
                faces.append([n, n + 1, n + 2])
    out = dict(verts=verts, faces=faces, rgb=[c] * len(verts))
    if mtl:
        out["mtl"] = mtl
    return out


def pinhole_dirs(pos, lookat, up, fov, w, h, xs, ys):
    """float64 pinhole camera (render/camera.h:128-209): normalize(-W + (2x/W
    - 1) hw U + (2y/H - 1) hh V); 1 spp rays go through pixel corners."""
    pos, lookat, up = (np.asarray(a, np.float64) for a in (pos, lookat, up))
    W = (pos - lookat) / np.linalg.norm(pos - lookat)
    U = np.cross(up, W)
    U /= np.linalg.norm(U)
    V = np.cross(W, U)
    hh = math.tan(math.radians(fov) / 2)
    hw = hh * w / h
    d = -W[None] + ((2 * xs / w - 1) * hw)[:, None] * U[None] + ((2 * ys / h - 1) * hh)[:, None] * V[None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def down_rays(cam, fov, w, h):
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    return pinhole_dirs(cam, [cam[0], 0.0, cam[2]], DOWN_UP, fov, w, h, xs.ravel(), ys.ravel())


def hit_plane(p, d, y):
    """Points where rays p + s d meet the plane at height y (s > 0 assumed)."""
    s = (y - p[..., 1]) / d[..., 1]
    return p + s[..., None] * d


def snell(cos_i, n1, n2):
    """(sin_t, cos_t, tir) for incidence cosine cos_i from index n1 into n2."""
    sin_i = np.sqrt(np.maximum(0.0, 1.0 - cos_i ** 2))
    sin_t = n1 / n2 * sin_i
    tir = sin_t >= 1.0
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - np.minimum(sin_t, 1.0) ** 2))
    return sin_t, cos_t, tir


def fresnel(cos_i, n1, n2):
    """Unpolarised dielectric reflectance in the angle form of the Fresnel
    equations; 1 under total internal reflection, the normal-incidence limit
    ((n1 - n2)/(n1 + n2))^2 where the angles vanish."""
    ti = np.arccos(np.clip(cos_i, -1.0, 1.0))
    sin_t, cos_t, tir = snell(cos_i, n1, n2)
    tt = np.arcsin(np.minimum(sin_t, 1.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        rs = -np.sin(ti - tt) / np.sin(ti + tt)
        rp = np.tan(ti - tt) / np.tan(ti + tt)
        f = 0.5 * (rs ** 2 + rp ** 2)
    r0 = ((n1 - n2) / (n1 + n2)) ** 2
    f = np.where(ti < 1e-6, r0, f)
    return np.where(tir, 1.0, f)


def refracted_dir(d, sin_t, cos_t, down=True):
    """Transmitted direction below (down=True) a horizontal interface: the
    horizontal part of d keeps its heading with length sin_t."""
    hz = d[..., [0, 2]]
    hn = np.linalg.norm(hz, axis=-1, keepdims=True)
    hu = np.where(hn > 0, hz / np.where(hn > 0, hn, 1.0), 0.0)
    out = np.zeros_like(d)
    out[..., 0] = sin_t * hu[..., 0]
    out[..., 2] = sin_t * hu[..., 1]
    out[..., 1] = -cos_t if down else cos_t
    return out


def segment_hits_rect(p, q, y, x0, x1, z0, z1):
    """Whether the unbounded shadow ray from p through q (the reference's point
    lights have no far bound) meets the rectangle at height y beyond p, and
    how far that crossing lies from the rectangle's border (pixels within a
    hair of it are left out of exact comparisons)."""
    d = q - p
    with np.errstate(divide="ignore", invalid="ignore"):
        lam = (y - p[..., 1]) / d[..., 1]
    x = p[..., 0] + lam * d[..., 0]
    z = p[..., 2] + lam * d[..., 2]
    hit = (lam > 1e-6) & (x >= x0) & (x <= x1) & (z >= z0) & (z <= z1)
    margin = np.minimum(np.minimum(np.abs(x - x0), np.abs(x - x1)),
                        np.minimum(np.abs(z - z0), np.abs(z - z1)))
    margin = np.where(lam > 1e-6, margin, np.inf)
    return hit, np.nan_to_num(margin, nan=np.inf)


def point_light_term(p, n, L, I, kd):
    """Diffuse direct light of a point light at p with normal n: I kd cos."""
    to = L[None] - p
    cos = (to @ n) / np.linalg.norm(to, axis=-1)
    return np.clip(cos, 0.0, 1.0)[:, None] * (I[None] * kd[None])


# ---------------------------------------------------------------- scenes
# Each scene: domains, light lines, camera (looking straight down), render
# settings, the vertex colour c and expect(kd) -> (image, valid): the
# float64 closed-form image and the pixels it speaks for exactly (pixels
# whose shadow ray grazes an occluder's border, or whose incidence lies
# within a hair of the critical angle, are left out).
def _light(L, I):
    return "light point " + " ".join(repr(float(x)) for x in (*L, *I))


def _all(h, w):
    return np.ones((h, w), bool)


def specular_scene():
    """Blinn-Phong with ks > 0 under a point light: I (kd cos + ks (n.h)^s)."""
    c = (179, 230, 20)
    I = np.array([0.8, 0.5, 1.5])
    L = np.array([0.3, 2.0, -0.2])
    ks, shin = (0.3, 0.5, 0.2), 8.0
    cam, fov, w, h = [0.0, 3.0, 0.0], 60.0, 40, 32
    doms = [quad(0.0, -50, 50, -50, 50, c=c)]
    kw = dict(image_w=w, image_h=h, pixel_samples=1, ao_mode=0, ao_samples=1, bounces=1, fov=fov,
              ks=ks, shininess=shin)

    def expect(kd):
        d = down_rays(cam, fov, w, h)
        p = hit_plane(np.asarray(cam, np.float64)[None], d, 0.0)
        wi = L[None] - p
        wi /= np.linalg.norm(wi, axis=-1, keepdims=True)
        hv = wi - d
        hv /= np.linalg.norm(hv, axis=-1, keepdims=True)
        ndh = np.clip(hv[:, 1], 0.0, 1.0)
        img = I[None] * (kd[None] * wi[:, 1:2] + np.asarray(ks)[None] * (ndh ** shin)[:, None])
        return img.reshape(h, w, 3), _all(h, w)
    return dict(doms=doms, lights=[_light(L, I)], cam=cam, kw=kw, c=c, expect=expect, useful=2 * w * h)


def hemisphere_open_scene():
    """'light diffuse L' over an open plane: every sample carries kd L pi."""
    c = (200, 150, 90)
    Lr = np.array([0.9, 0.6, 0.3])
    cam, fov, w, h = [0.2, 2.0, -0.1], 50.0, 36, 30
    doms = [quad(0.0, -50, 50, -50, 50, c=c)]
    S = 4
    kw = dict(image_w=w, image_h=h, pixel_samples=1, ao_mode=0, ao_samples=S, bounces=1, fov=fov)

    def expect(kd):
        img = np.broadcast_to(kd[None, None] * Lr[None, None] * math.pi, (h, w, 3)).copy()
        return img, _all(h, w)
    return dict(doms=doms, lights=["light diffuse " + " ".join(repr(float(x)) for x in Lr)], cam=cam,
                kw=kw, c=c, expect=expect, useful=(1 + S) * w * h)


def hemisphere_half_ceiling_scene():
    """'light diffuse' over a floor at y = 0 with a ceiling at y = 1 over x > 0.
    A cosine-distributed direction reaches y = 1 at horizontal offset X =
    tan(theta) cos(phi), whose density is 1 / (2 (1 + X^2)^(3/2)); the floor
    point at x0 is shadowed with probability (1 + x0 / sqrt(1 + x0^2)) / 2.
    The hemisphere light's shadow rays are unbounded (light.h:67-89).  The
    image is stochastic: expect() gives each pixel's mean."""
    c = (255, 255, 255)
    Lr = np.array([1.0, 1.0, 1.0])
    cam, fov, w, h = [0.0, 0.9, 0.0], 60.0, 64, 64
    doms = [quad(0.0, -60, 60, -60, 60, c=c), quad(1.0, 0, 60, -60, 60, c=c, up=False)]
    S = 16
    kw = dict(image_w=w, image_h=h, pixel_samples=1, ao_mode=0, ao_samples=S, bounces=1, fov=fov)

    def expect(kd):
        d = down_rays(cam, fov, w, h)
        p = hit_plane(np.asarray(cam, np.float64)[None], d, 0.0)
        x0 = p[:, 0]
        lit = 1.0 - 0.5 * (1.0 + x0 / np.sqrt(1.0 + x0 ** 2))
        return (lit[:, None] * kd[None] * Lr[None] * math.pi).reshape(h, w, 3), x0.reshape(h, w)
    return dict(doms=doms, lights=["light diffuse 1 1 1"], cam=cam, kw=kw, c=c, expect=expect,
                stochastic=True)


def mirror_scene(unit=True):
    """Eye rays hit a mirror at y = 0, reflect to a ceiling at y = 3 (facing
    down) lit by a point light beside it; pixel = (1/|cos_i|) kd I cos_c with
    cos_i the geometric cosine for unit normals, 1 for the clamped large
    sheet.  The unbounded shadow rays leave sideways, missing the mirror."""
    c = (220, 180, 140)
    I = np.array([2.0, 1.5, 1.0])
    L = np.array([6.0, 2.0, 0.0])
    cam, fov, w, h = [0.0, 2.0, 0.0], 30.0, 32, 28
    mirror = (unit_sheet(0.0, -2, 2, -2, 2, mtl="mirror 0.5 0.5 0.5") if unit
              else quad(0.0, -2, 2, -2, 2, mtl="mirror 0.5 0.5 0.5"))
    doms = [mirror, quad(3.0, -10, 10, -10, 10, c=c, up=False)]
    kw = dict(image_w=w, image_h=h, pixel_samples=1, ao_mode=0, ao_samples=1, bounces=2, fov=fov)

    def expect(kd):
        d = down_rays(cam, fov, w, h)
        p1 = hit_plane(np.asarray(cam, np.float64)[None], d, 0.0)
        r = d * np.array([1.0, -1.0, 1.0])
        p2 = hit_plane(p1, r, 3.0)
        weight = 1.0 / np.abs(d[:, 1]) if unit else np.ones(len(d))
        img = weight[:, None] * point_light_term(p2, np.array([0.0, -1.0, 0.0]), L, I, kd)
        occ, margin = segment_hits_rect(p2, L[None], 0.0, -2, 2, -2, 2)
        img[occ] = 0.0
        return img.reshape(h, w, 3), (margin > 1e-3).reshape(h, w)
    return dict(doms=doms, lights=[_light(L, I)], cam=cam, kw=kw, c=c, expect=expect, useful=3 * w * h)


def dielectric_transmit_scene(mtl="glass", eta=(1.0, 1.5), fov=40.0, unit=True):
    """Eye rays from y = 3 meet a dielectric sheet at y = 1 (normal +y, so they
    enter: n1 = eta[0], n2 = eta[1]); the transmitted ray refracts to a
    floor at y = 0 lit by a low point light.  pixel = T / |cos_i| * kd I
    cos_f, T = 1 - Fresnel (glass) or 1 (transmission); 0 beyond the critical
    angle (the reflected ray escapes upward) and where the floor point's
    unbounded shadow ray crosses the sheet.  Large sheet: cos_i = 1."""
    n1, n2 = eta
    c = (200, 210, 160)
    I = np.array([1.2, 0.9, 0.7])
    L = np.array([6.0, 0.5, 0.0])
    cam, w, h = [0.0, 3.0, 0.0], 40, 36
    ext = 8
    mt = f"{mtl} {n1!r} {n2!r}"
    sheet = unit_sheet(1.0, -ext, ext, -ext, ext, mtl=mt) if unit else quad(1.0, -ext, ext, -ext, ext, mtl=mt)
    fext = 60  # larger triangles would cost the float32 hit point its precision
    doms = [sheet, quad(0.0, -fext, fext, -fext, fext, c=c)]
    kw = dict(image_w=w, image_h=h, pixel_samples=1, ao_mode=0, ao_samples=1, bounces=2, fov=fov)

    def expect(kd):
        d = down_rays(cam, fov, w, h)
        p1 = hit_plane(np.asarray(cam, np.float64)[None], d, 1.0)
        if unit:
            cos_i = -d[:, 1]
            sin_t, cos_t, tir = snell(cos_i, n1, n2)
            wt = refracted_dir(d, sin_t, cos_t)
            weight = 1.0 / cos_i
        else:  # clamped: normal incidence for every pixel
            cos_i = np.ones(len(d))
            sin_t = np.zeros(len(d))
            eta_r = n1 / n2
            wt = eta_r * d + (eta_r - 1.0) * np.array([0.0, 1.0, 0.0])[None]
            wt /= np.linalg.norm(wt, axis=-1, keepdims=True)
            tir = np.zeros(len(d), bool)
            weight = np.ones(len(d))
        T = (1.0 - fresnel(cos_i, n1, n2)) if mtl == "glass" else np.ones(len(d))
        with np.errstate(divide="ignore", invalid="ignore"):
            p2 = hit_plane(p1, wt, 0.0)
        img = (weight * T)[:, None] * point_light_term(p2, np.array([0.0, 1.0, 0.0]), L, I, kd)
        occ, margin = segment_hits_rect(p2, L[None], 1.0, -ext, ext, -ext, ext)
        reach = np.nan_to_num(np.abs(p2[:, [0, 2]]).max(axis=1), nan=np.inf)
        img[occ | tir | (reach > fext)] = 0.0
        img = np.nan_to_num(img)
        valid = (margin > 1e-3) & (np.abs(sin_t - 1.0) > 1e-2) & (np.abs(reach - fext) > 1e-2)
        return img.reshape(h, w, 3), valid.reshape(h, w)
    return dict(doms=doms, lights=[_light(L, I)], cam=cam, kw=kw, c=c, expect=expect)


def dielectric_reflect_scene(mtl="glass", eta=(1.0, 1.5), fov=120.0, unit=True):
    """Eye rays from y = 3 meet a dielectric sheet at y = 1; the reflected ray
    rises to a ceiling at y = 7 (facing down) lit by a point light beside it
    and the transmitted ray escapes downward.  pixel = R / |cos_i| kd I cos_c,
    R = Fresnel (glass; 1 beyond the critical angle) or, for transmission, 1
    under total internal reflection and 0 otherwise."""
    n1, n2 = eta
    c = (240, 200, 120)
    I = np.array([1.0, 1.3, 0.8])
    L = np.array([40.0, 6.0, 0.0])
    cam, w, h = [0.0, 3.0, 0.0], 48, 40
    ext = 8
    mt = f"{mtl} {n1!r} {n2!r}"
    sheet = unit_sheet(1.0, -ext, ext, -ext, ext, mtl=mt) if unit else quad(1.0, -ext, ext, -ext, ext, mtl=mt)
    cext = 120
    doms = [sheet, quad(7.0, -cext, cext, -cext, cext, c=c, up=False)]
    kw = dict(image_w=w, image_h=h, pixel_samples=1, ao_mode=0, ao_samples=1, bounces=2, fov=fov)

    def expect(kd):
        d = down_rays(cam, fov, w, h)
        p1 = hit_plane(np.asarray(cam, np.float64)[None], d, 1.0)
        inside = (np.abs(p1[:, 0]) <= ext) & (np.abs(p1[:, 2]) <= ext)
        cos_i = -d[:, 1] if unit else np.ones(len(d))
        sin_t, _, tir = snell(cos_i, n1, n2)
        R = fresnel(cos_i, n1, n2) if mtl == "glass" else tir.astype(np.float64)
        r = d * np.array([1.0, -1.0, 1.0])
        p2 = hit_plane(p1, r, 7.0)
        img = (R / cos_i)[:, None] * point_light_term(p2, np.array([0.0, -1.0, 0.0]), L, I, kd)
        occ, margin = segment_hits_rect(p2, L[None], 1.0, -ext, ext, -ext, ext)
        out = (np.abs(p2[:, 0]) > cext) | (np.abs(p2[:, 2]) > cext)
        img[occ | out | ~inside] = 0.0
        valid = (margin > 1e-3) & (np.abs(sin_t - 1.0) > 1e-2) & inside
        return img.reshape(h, w, 3), valid.reshape(h, w)
    return dict(doms=doms, lights=[_light(L, I)], cam=cam, kw=kw, c=c, expect=expect)


def all_scenes():
    """name -> scene, for the GPU-vs-oracle frame tests."""
    return {
        "specular": specular_scene(),
        "hemisphere_open": hemisphere_open_scene(),
        "hemisphere_half_ceiling": hemisphere_half_ceiling_scene(),
        "mirror_unit": mirror_scene(True),
        "mirror_large": mirror_scene(False),
        "glass_transmit_unit": dielectric_transmit_scene("glass", (1.0, 1.5), 40.0, True),
        "glass_transmit_large": dielectric_transmit_scene("glass", (1.0, 1.5), 40.0, False),
        "glass_transmit_tir": dielectric_transmit_scene("glass", (1.5, 1.0), 120.0, True),
        "transmission_transmit": dielectric_transmit_scene("transmission", (1.0, 1.5), 40.0, True),
        "glass_reflect_unit": dielectric_reflect_scene("glass", (1.0, 1.5), 120.0, True),
        "glass_reflect_large": dielectric_reflect_scene("glass", (1.0, 1.5), 120.0, False),
        "glass_reflect_tir": dielectric_reflect_scene("glass", (1.5, 1.0), 120.0, True),
        "transmission_reflect_tir": dielectric_reflect_scene("transmission", (1.5, 1.0), 120.0, True),
    }
