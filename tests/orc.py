"""ctypes wrapper of the oracle (oracle/liborc.so): test infrastructure only."""
import ctypes as C
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.environ.get("ORC_LIB") or os.path.join(ROOT, "oracle", "liborc.so")


class TraceConfig(C.Structure):
    _fields_ = [
        ("image_w", C.c_int), ("image_h", C.c_int),
        ("cam_pos", C.c_float * 3), ("cam_lookat", C.c_float * 3), ("cam_up", C.c_float * 3),
        ("fov", C.c_float),
        ("pixel_samples", C.c_int), ("ao_samples", C.c_int), ("ao_mode", C.c_int),
        ("bounces", C.c_int),
        ("ks", C.c_float * 3), ("shininess", C.c_float),
        ("max_samples_per_rank", C.c_longlong),
        ("nthreads", C.c_int),
        ("ao_max_dist", C.c_float),
        ("rr_start", C.c_int),
        ("ref_pending_bug", C.c_int),
        ("canonical", C.c_int),
        ("tile_begin", C.c_int), ("tile_end", C.c_int),
    ]


class TraceStats(C.Structure):
    _fields_ = [(n, C.c_ulonglong) for n in (
        "useful_rays", "eye_rays", "radiance_visits", "shadow_visits", "shades", "final_hits",
        "passes", "windows", "tiles")] + [("seconds", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        l = C.CDLL(LIB)
        l.orc_last_error.restype = C.c_char_p
        l.orc_scene_new.restype = C.c_void_p
        for n in ("orc_scene_free", "orc_scene_num_domains", "orc_scene_num_lights"):
            getattr(l, n).argtypes = [C.c_void_p]
        l.orc_scene_load.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]
        l.orc_scene_parse.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        l.orc_scene_bound.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_scene_domain.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_mesh_sizes.restype = C.c_longlong
        l.orc_mesh_sizes.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        l.orc_mesh_copy.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_closest_hit.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_void_p, C.c_void_p, C.c_float,
                                      C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_any_hit.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_void_p, C.c_void_p, C.c_float,
                                  C.c_void_p, C.c_void_p]
        l.orc_domain_list.argtypes = [C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
        l.orc_interpolate.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
        l.orc_tri_intersect.argtypes = [C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_rng_seed.restype = C.c_uint32
        l.orc_rng_seed.argtypes = [C.c_int]
        l.orc_murmur_mix.restype = C.c_uint32
        l.orc_murmur_mix.argtypes = [C.c_uint32, C.c_uint32]
        l.orc_murmur_fmix.restype = C.c_uint32
        l.orc_murmur_fmix.argtypes = [C.c_uint32]
        l.orc_rng_floats.argtypes = [C.c_int, C.c_int, C.c_void_p]
        l.orc_sincos.argtypes = [C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_concentric.argtypes = [C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_cosine_sample.argtypes = [C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_camera_dirs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.c_int,
                                      C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_num_tiles.argtypes = [C.POINTER(TraceConfig)]
        l.orc_render.argtypes = [C.c_void_p, C.POINTER(TraceConfig), C.c_void_p, C.POINTER(TraceStats)]
        l.orc_render_schedule.argtypes = [C.c_void_p, C.POINTER(TraceConfig), C.c_void_p, C.POINTER(TraceStats),
                                          C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_ply_read.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        assert l.orc_sizeof_config() == C.sizeof(TraceConfig), "TraceConfig layout mismatch"
        assert l.orc_sizeof_stats() == C.sizeof(TraceStats), "TraceStats layout mismatch"
        _lib = l
    return _lib


def p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    pass


class Scene:
    def __init__(self, spray_file=None, ply_path="", nthreads=8, parse_only=False):
        self.l = lib()
        self.h = self.l.orc_scene_new()
        if spray_file is not None:
            fn = self.l.orc_scene_parse if parse_only else self.l.orc_scene_load
            args = (self.h, str(spray_file).encode(), str(ply_path).encode())
            rc = fn(*args) if parse_only else fn(*args, nthreads)
            if rc != 0:
                raise OracleError(self.l.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.l.orc_scene_free(self.h)
            self.h = None

    @property
    def num_domains(self):
        return self.l.orc_scene_num_domains(self.h)

    @property
    def num_lights(self):
        return self.l.orc_scene_num_lights(self.h)

    def bound(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        self.l.orc_scene_bound(self.h, p(lo), p(hi))
        return lo, hi

    def domain(self, i):
        nv, nf = C.c_longlong(), C.c_longlong()
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        mtl = C.c_int()
        assert self.l.orc_scene_domain(self.h, i, C.byref(nv), C.byref(nf), p(lo), p(hi), C.byref(mtl)) == 0
        return nv.value, nf.value, lo, hi, mtl.value

    def mesh(self, i):
        nf = C.c_longlong()
        nv = self.l.orc_mesh_sizes(self.h, i, C.byref(nf))
        v = np.zeros((nv, 3), np.float32)
        f = np.zeros((nf.value, 3), np.uint32)
        rgb = np.zeros(nv, np.uint32)
        n = np.zeros((nv, 3), np.float32)
        self.l.orc_mesh_copy(self.h, i, p(v), p(f), p(rgb), p(n))
        return v, f, rgb, n

    def closest(self, dom, org, dirs, tnear=0.001, tfar=math.inf, brute=False):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(org)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        pr = np.zeros(n, np.int32)
        assert self.l.orc_closest_hit(self.h, dom, n, p(org), p(dirs), tnear, tfar, int(brute),
                                      p(t), p(u), p(v), p(pr)) == 0
        return t, u, v, pr

    def any_hit(self, dom, org, dirs, tfar, tnear=0.001):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(org)
        tf = np.broadcast_to(np.asarray(tfar, np.float32), (n,)).copy()
        out = np.zeros(n, np.int32)
        assert self.l.orc_any_hit(self.h, dom, n, p(org), p(dirs), tnear, p(tf), p(out)) == 0
        return out.astype(bool)

    def domain_list(self, org, dirs, maxk=32):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(org)
        cnt = np.zeros(n, np.int32)
        ids = np.zeros((n, maxk), np.int32)
        ts = np.zeros((n, maxk), np.float32)
        self.l.orc_domain_list(self.h, n, p(org), p(dirs), maxk, p(cnt), p(ids), p(ts))
        return cnt, ids, ts

    def interpolate(self, dom, prim, u, v):
        prim = np.ascontiguousarray(prim, np.int32)
        u = np.ascontiguousarray(u, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        n = len(prim)
        col = np.zeros(n, np.uint32)
        ns = np.zeros((n, 3), np.float32)
        self.l.orc_interpolate(self.h, dom, n, p(prim), p(u), p(v), p(col), p(ns))
        return col, ns

    def render(self, **kw):
        cfg = make_config(**kw)
        img = np.zeros((cfg.image_h, cfg.image_w, 4), np.float32)
        st = TraceStats()
        if self.l.orc_render(self.h, C.byref(cfg), p(img), C.byref(st)) != 0:
            raise OracleError(self.l.orc_last_error().decode())
        return img, {f: getattr(st, f) for f, _ in TraceStats._fields_}


def _render_schedule(self, max_records=256, **kw):
    """render() plus the DomainStats schedule of the first max_records passes:
    (img, stats, meta[n, 4] = (tile, window, pass, entries), scores[n, nd],
    order[n, nd])."""
    cfg = make_config(**kw)
    img = np.zeros((cfg.image_h, cfg.image_w, 4), np.float32)
    st = TraceStats()
    nd = self.num_domains
    nrec = C.c_int(0)
    meta = np.zeros((max_records, 4), np.int32)
    scores = np.zeros((max_records, nd), np.int64)
    order = np.zeros((max_records, nd), np.int32)
    if self.l.orc_render_schedule(self.h, C.byref(cfg), p(img), C.byref(st), max_records, C.byref(nrec),
                                  p(meta), p(scores), p(order)) != 0:
        raise OracleError(self.l.orc_last_error().decode())
    n = nrec.value
    return img, {f: getattr(st, f) for f, _ in TraceStats._fields_}, meta[:n], scores[:n], order[:n]


Scene.render_schedule = _render_schedule


def make_config(**kw):
    c = TraceConfig()
    c.image_w, c.image_h = 400, 400
    c.cam_lookat[2] = -1.0
    c.cam_up[1] = 1.0
    c.fov = 90.0
    c.pixel_samples = 4
    c.ao_samples = 8
    c.bounces = 1
    c.shininess = 1.0
    c.max_samples_per_rank = 1024 * 1024
    c.nthreads = 1
    c.ao_max_dist = math.inf
    c.tile_begin = 0
    c.tile_end = 1 << 30
    for k, v in kw.items():
        if k in ("cam_pos", "cam_lookat", "cam_up", "ks"):
            for i in range(3):
                getattr(c, k)[i] = float(v[i])
        else:
            setattr(c, k, v)
    return c


def num_tiles(**kw):
    return lib().orc_num_tiles(C.byref(make_config(**kw)))
