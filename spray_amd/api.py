"""Python view of the C ABI: configuration, scene generation and a tracer
context.  Mirrors the reference's driver flow (render/spray_renderer.inl:
Config -> Scene::init -> Camera::init -> Tracer::trace) one rank at a time.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L


class SprayError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"spray error {code}: {msg}")
        self.code = code


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def default_config(**kw):
    """spray_config with the reference defaults (render/config.cc:33-73),
    overridden by keyword arguments (lists for the 3-vectors)."""
    lib = L.load()
    cfg = L.spray_config()
    lib.spray_config_defaults(C.byref(cfg))
    for k, v in kw.items():
        if k in ("camera_pos", "camera_lookat", "camera_up", "ks"):
            for i in range(3):
                getattr(cfg, k)[i] = float(v[i])
        else:
            setattr(cfg, k, v)
    return cfg


def _gen_params(nx, ny, nz, res, a=6.0, b=17.0, phase=0.0, rgb=0xFFFFFF, per_domain_random=False,
                seed=0, tri_total=0.0):
    return L.spray_scenegen_params(nx=nx, ny=ny, nz=nz, res=res, a=a, b=b, phase=phase,
                                   per_domain_random=int(per_domain_random), seed=seed,
                                   tri_total=tri_total, rgb=rgb)


def scenegen_write(directory, name, nx, ny, nz, res, a=6.0, b=17.0, phase=0.0, rgb=0xFFFFFF,
                   per_domain_random=False, seed=0, tri_total=0.0, extra_lines=""):
    """Writes a marching-cubes scene (name.spray + PLYs) and returns the triangle count."""
    lib = L.load()
    p = _gen_params(nx, ny, nz, res, a, b, phase, rgb, per_domain_random, seed, tri_total)
    tris = C.c_longlong(0)
    rc = lib.spray_scenegen_write(C.byref(p), str(directory).encode(), name.encode(),
                                  extra_lines.encode(), C.byref(tris))
    if rc != L.SPRAY_OK:
        raise SprayError(rc, lib.spray_last_error(None).decode())
    return tris.value


def scenegen_domain(domain, nx, ny, nz, res, a=6.0, b=17.0, phase=0.0, rgb=0xFFFFFF,
                    per_domain_random=False, seed=0, tri_total=0.0):
    """One domain's mesh from the CPU generator: (verts[n, 3], faces[m, 3], rgb[n])."""
    lib = L.load()
    p = _gen_params(nx, ny, nz, res, a, b, phase, rgb, per_domain_random, seed, tri_total)
    nv, nf = C.c_longlong(0), C.c_longlong(0)
    rc = lib.spray_scenegen_domain(C.byref(p), domain, None, None, None, C.byref(nv), C.byref(nf))
    if rc != L.SPRAY_OK:
        raise SprayError(rc, "spray_scenegen_domain")
    v = np.zeros((nv.value, 3), np.float32)
    f = np.zeros((nf.value, 3), np.uint32)
    c = np.zeros(nv.value, np.uint32)
    rc = lib.spray_scenegen_domain(C.byref(p), domain, _fp(v), f.ctypes.data_as(C.POINTER(C.c_uint)),
                                   c.ctypes.data_as(C.POINTER(C.c_uint)), C.byref(nv), C.byref(nf))
    if rc != L.SPRAY_OK:
        raise SprayError(rc, "spray_scenegen_domain")
    return v, f, c


def write_ply_info(infile, outfile, info):
    """The ply_header_reader's output file (apps/ply_header_reader.cc:90-104):
    file, vertex and face lines and the bound line with 16 significant
    digits (std::setprecision(digits10 + 1) on the float bounds)."""
    nv, nf, lo, hi = info
    b = " ".join("%.16g" % float(x) for x in (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]))
    with open(outfile, "w") as fp:
        fp.write(f"file {infile}\nvertex {nv}\nface {nf}\nbound {b}\n")


def write_ppm(path, rgba):
    """Writes an (h, w, 4) float image as the reference's P3 / 1023 PPM."""
    import numpy as np
    lib = L.load()
    a = np.ascontiguousarray(rgba, np.float32)
    h, w = a.shape[:2]
    rc = lib.spray_write_ppm(_fp(a), w, h, str(path).encode())
    if rc != L.SPRAY_OK:
        raise SprayError(rc, lib.spray_last_error(None).decode())


class Context:
    """One rank's tracer (spray_ctx)."""

    def __init__(self, cfg):
        self.lib = L.load()
        self.cfg = cfg
        h = C.c_void_p()
        rc = self.lib.spray_create(C.byref(cfg), C.byref(h))
        if rc != L.SPRAY_OK:
            raise SprayError(rc, self.lib.spray_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != L.SPRAY_OK:
            raise SprayError(rc, self.lib.spray_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.spray_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_scene(self, spray_file, ply_path=""):
        self._check(self.lib.spray_load_scene(self.h, str(spray_file).encode(), str(ply_path).encode()))

    def scenegen_load(self, name, nx, ny, nz, res, a=6.0, b=17.0, phase=0.0, rgb=0xFFFFFF,
                      per_domain_random=False, seed=0, tri_total=0.0, extra_lines=""):
        """Generates the marching-cubes scene on the GPU straight into the domain pool."""
        p = _gen_params(nx, ny, nz, res, a, b, phase, rgb, per_domain_random, seed, tri_total)
        self._check(self.lib.spray_scenegen_load(self.h, C.byref(p), name.encode(), extra_lines.encode()))

    def remesh(self, phase):
        self._check(self.lib.spray_remesh(self.h, float(phase)))

    def domain_mesh(self, i):
        """(vertices (nv, 3) float32, faces (nf, 3) uint32, colours (nv,) uint32) of domain i."""
        nv, nf = self.mesh_size(i)
        v = np.zeros((nv, 3), np.float32)
        f = np.zeros((nf, 3), np.uint32)
        rgb = np.zeros(nv, np.uint32)
        up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint))
        self._check(self.lib.spray_domain_mesh(self.h, i, _fp(v), up(f), up(rgb), None, None))
        return v, f, rgb

    def write_scene(self, directory, name):
        self._check(self.lib.spray_write_scene(self.h, str(directory).encode(), name.encode()))

    def ply_info(self, path):
        """(vertex count, face count, lo[3], hi[3]) of a PLY file."""
        nv, nf = C.c_longlong(), C.c_longlong()
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        self._check(self.lib.spray_ply_info(self.h, str(path).encode(), C.byref(nv), C.byref(nf),
                                            _fp(lo), _fp(hi)))
        return nv.value, nf.value, lo, hi

    @property
    def num_domains(self):
        return self.lib.spray_num_domains(self.h)

    def domain_info(self, i):
        nv, nf = C.c_longlong(), C.c_longlong()
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        self._check(self.lib.spray_domain_info(self.h, i, C.byref(nv), C.byref(nf), _fp(lo), _fp(hi)))
        return nv.value, nf.value, lo, hi

    def mesh_size(self, i):
        """(vertices, faces) of domain i as loaded."""
        nv, nf = C.c_longlong(), C.c_longlong()
        self._check(self.lib.spray_domain_mesh(self.h, i, None, None, None, C.byref(nv), C.byref(nf)))
        return nv.value, nf.value

    def normals(self, i):
        nv = self.mesh_size(i)[0]
        out = np.zeros((nv, 3), np.float32)
        self._check(self.lib.spray_domain_normals(self.h, i, _fp(out)))
        return out

    def set_camera(self, pos, lookat, up, fov):
        a = [np.asarray(x, np.float32) for x in (pos, lookat, up)]
        self._check(self.lib.spray_set_camera(self.h, _fp(a[0]), _fp(a[1]), _fp(a[2]), float(fov)))

    def set_config(self, cfg):
        self._check(self.lib.spray_set_config(self.h, C.byref(cfg)))
        self.cfg = cfg

    def set_schedule(self, s):
        self._check(self.lib.spray_set_schedule(self.h, s))

    def set_batch_params(self, chunk, tile):
        self._check(self.lib.spray_set_batch_params(self.h, chunk, tile))

    def set_tiles_in_flight(self, tiles):
        self._check(self.lib.spray_set_tiles_in_flight(self.h, int(tiles)))

    def set_cache_size(self, domains):
        """Domains kept in HBM from the next load_scene (< 0: all, the default)."""
        self._check(self.lib.spray_set_cache_size(self.h, int(domains)))

    def cache_state(self):
        """(capacity, resident domains least to most recently used, hits, misses, bytes)."""
        cap = C.c_int(0)
        self._check(self.lib.spray_cache_state(self.h, C.byref(cap), None, None))
        doms = (C.c_int * cap.value)()
        cnt = (C.c_ulonglong * 3)()
        self._check(self.lib.spray_cache_state(self.h, C.byref(cap), doms, cnt))
        return cap.value, [d for d in doms if d >= 0], int(cnt[0]), int(cnt[1]), int(cnt[2])

    def set_bvh_builder(self, builder):
        """0 = Karras LBVH (default), 1 = PLOC; applies to the next load_scene."""
        self._check(self.lib.spray_set_bvh_builder(self.h, int(builder)))

    def autotune(self):
        ch, tl = C.c_int(), C.c_int()
        self._check(self.lib.spray_autotune(self.h, C.byref(ch), C.byref(tl)))
        return ch.value, tl.value

    def trace_frame(self, want_image=True):
        st = L.spray_stats()
        img = np.zeros((self.cfg.image_h, self.cfg.image_w, 4), np.float32) if want_image else None
        self._check(self.lib.spray_trace_frame(self.h, _fp(img) if want_image else None, C.byref(st)))
        return img, {f: getattr(st, f) for f, _ in L.spray_stats._fields_}

    def trace_rays(self, org, dirs, domain, mode=L.QUERY_CLOSEST, tnear=0.001, tfar=math.inf):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = org.shape[0]
        cols = [np.ascontiguousarray(org[:, k]) for k in range(3)] + \
               [np.ascontiguousarray(dirs[:, k]) for k in range(3)]
        tn = np.full(n, tnear, np.float32) if np.isscalar(tnear) else np.ascontiguousarray(tnear, np.float32)
        tf = np.full(n, tfar, np.float32) if np.isscalar(tfar) else np.ascontiguousarray(tfar, np.float32)
        rays = L.spray_rays_soa(*[_fp(c) for c in cols], _fp(tn), _fp(tf))
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        p = np.zeros(n, np.int32)
        hits = L.spray_hits_soa(_fp(t), _fp(u), _fp(v), _ip(p))
        self._check(self.lib.spray_trace_rays(self.h, C.byref(rays), n, C.byref(hits), mode, domain))
        if mode == L.QUERY_ANY:
            return p.astype(bool)
        return t, u, v, p

    def domain_list(self, org, dirs, maxk=32):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = org.shape[0]
        cnt = np.zeros(n, np.int32)
        ids = np.zeros((n, maxk), np.int32)
        ts = np.zeros((n, maxk), np.float32)
        self._check(self.lib.spray_domain_list(self.h, _fp(org), _fp(dirs), n, maxk, _ip(cnt), _ip(ids), _fp(ts)))
        return cnt, ids, ts

    def profile_kernels(self, enable):
        self._check(self.lib.spray_profile_kernels(self.h, int(enable)))

    def kernel_time(self, reset=True):
        n, ms = C.c_longlong(), C.c_double()
        self._check(self.lib.spray_kernel_time(self.h, C.byref(n), C.byref(ms), int(reset)))
        return n.value, ms.value

    def set_pool_headroom(self, factor):
        """Factor on the initial ray-pool estimate (smaller: overflow and regrow)."""
        self._check(self.lib.spray_set_pool_headroom(self.h, float(factor)))

    def schedule_log(self, max_records):
        """Records the plain schedule's first max_records passes from the next
        frame on (DomainStats scores and visit order per pass)."""
        self._check(self.lib.spray_schedule_log(self.h, int(max_records)))
        self._slog_max = int(max_records)

    def schedule_read(self):
        """(meta[n, 4] = (tile, window, pass, entries), scores[n, nd], order[n, nd])."""
        nd, m = self.num_domains, max(1, getattr(self, "_slog_max", 0))
        n = C.c_int(0)
        meta = np.zeros((m, 4), np.int32)
        scores = np.zeros((m, nd), np.uint64)
        order = np.zeros((m, nd), np.int32)
        self._check(self.lib.spray_schedule_read(self.h, C.byref(n), _ip(meta),
                                                 scores.ctypes.data_as(C.POINTER(C.c_ulonglong)), _ip(order)))
        k = n.value
        return meta[:k], scores[:k].astype(np.int64), order[:k]

    # learned schedule
    def predictor_log(self, enable, max_samples=0):
        self._check(self.lib.spray_predictor_log(self.h, int(enable), int(max_samples)))

    def predictor_num_samples(self):
        return self.lib.spray_predictor_num_samples(self.h)

    def predictor_train(self, epochs, batch, lr):
        loss = C.c_float()
        self._check(self.lib.spray_predictor_train(self.h, epochs, batch, lr, C.byref(loss)))
        return loss.value

    def predictor_eval(self):
        loss, acc = C.c_float(), C.c_float()
        self._check(self.lib.spray_predictor_eval(self.h, C.byref(loss), C.byref(acc)))
        return loss.value, acc.value

    def mlp_num_params(self):
        return self.lib.spray_mlp_num_params(self.h)

    def mlp_grad(self, samples, params):
        samples = np.ascontiguousarray(samples, np.float32)
        params = np.ascontiguousarray(params, np.float32)
        g = np.zeros(self.mlp_num_params(), np.float32)
        loss = C.c_float()
        self._check(self.lib.spray_mlp_grad(self.h, _fp(samples), len(samples), _fp(params),
                                            C.byref(loss), _fp(g)))
        return loss.value, g

    def mlp_logits(self, samples, params=None):
        samples = np.ascontiguousarray(samples, np.float32)
        z = np.zeros(len(samples), np.float32)
        pp = _fp(np.ascontiguousarray(params, np.float32)) if params is not None else None
        self._check(self.lib.spray_mlp_logits(self.h, _fp(samples), len(samples), pp, _fp(z)))
        return z

    def predictor_save(self, path):
        self._check(self.lib.spray_predictor_save(self.h, str(path).encode()))

    def predictor_load(self, path):
        self._check(self.lib.spray_predictor_load(self.h, str(path).encode()))
