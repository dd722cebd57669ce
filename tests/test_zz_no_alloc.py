"""The no-allocation entry point spray_trace_rays_device (callers that cannot
allocate: graph capture, serving loops) against spray_trace_rays on the same
rays: closest and any hits bit-equal, on two calls in a row, the second with
fewer rays (its static scratch keeps the last call's contents)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(rng, n):
    org = rng.uniform(-0.5, 2.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tnear = np.full(n, 0.001, np.float32)
    tfar = rng.uniform(0.05, 3.0, n).astype(np.float32)
    tfar[: n // 2] = np.inf
    return org, dirs, tnear, tfar


class _Hip:
    """Device arrays from the HIP runtime the library itself links
    (libamdhip64.so.7, already loaded with it), not from a second runtime."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _device_query(ctx, org, dirs, tnear, tfar, mode, dom):
    hip = _Hip()
    n = len(org)
    try:
        d_org, d_dir, d_tn, d_tf = (hip.put(x) for x in (org, dirs, tnear, tfar))
        # outputs start as garbage so that nothing unwritten can pass
        junk_f = np.full(n, -7.0, np.float32)
        junk_i = np.full(n, -7, np.int32)
        d_t, d_u, d_v = (hip.put(junk_f) for _ in range(3))
        d_p = hip.put(junk_i)
        fp = lambda x: C.cast(x, C.POINTER(C.c_float))
        ip = lambda x: C.cast(x, C.POINTER(C.c_int))
        closest = mode == 0
        rc = ctx.lib.spray_trace_rays_device(ctx.h, fp(d_org), fp(d_dir), fp(d_tn), fp(d_tf), n,
                                             fp(d_t) if closest else None, fp(d_u) if closest else None,
                                             fp(d_v) if closest else None, ip(d_p), mode, dom)
        assert rc == 0, ctx.lib.spray_last_error(ctx.h)
        # the entry point has waited for its stream (it reads the error word)
        assert hip.rt.hipDeviceSynchronize() == 0
        return (hip.get(d_t, junk_f), hip.get(d_u, junk_f), hip.get(d_v, junk_f),
                hip.get(d_p, junk_i))
    finally:
        hip.free()


def test_no_alloc_matches_trace_rays(tmp_path):
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "s", 2, 2, 2, 24, rgb=0xFFFFFF)
    cam = dict(camera_pos=[1.0, 1.0, 1.0 + 2.0 * math.sqrt(3.0)], camera_lookat=[1.0, 1.0, 1.0],
               camera_up=[0.0, 1.0, 0.0])
    cfg = sa.default_config(image_w=64, image_h=64, has_camera=1, **cam)
    rng = np.random.default_rng(17)
    with sa.Context(cfg) as ctx:
        ctx.load_scene(os.path.join(d, "s.spray"), d)
        # more rays than the scratch's 65,536 lanes (grid-stride), then fewer
        for n, dom in ((200_003, 3), (1_537, 5)):
            org, dirs, tnear, tfar = _rays(rng, n)
            for mode in (0, 1):
                t, u, v, p = _device_query(ctx, org, dirs, tnear, tfar, mode, dom)
                if mode == 0:
                    rt, ru, rv, rp = ctx.trace_rays(org, dirs, dom, tnear=tnear, tfar=tfar)
                    print(f"n {n} domain {dom}: hits {int((rp >= 0).sum())}")
                    assert (rp >= 0).sum() > 100
                    assert np.array_equal(p, rp)
                    assert np.array_equal(t, rt) and np.array_equal(u, ru) and np.array_equal(v, rv)
                else:
                    rocc = ctx.trace_rays(org, dirs, dom, mode=1, tnear=tnear, tfar=tfar)
                    assert rocc.sum() > 100
                    assert np.array_equal(p.astype(bool), rocc.astype(bool))
                    assert set(np.unique(p)) <= {0, 1}


def test_no_alloc_rejects_bad_input(tmp_path):
    import spray_amd as sa
    d = str(tmp_path)
    sa.scenegen_write(d, "s", 1, 1, 1, 8, rgb=0xFFFFFF)
    with sa.Context(sa.default_config(image_w=8, image_h=8)) as ctx:
        ctx.load_scene(os.path.join(d, "s.spray"), d)
        lib = ctx.lib
        assert lib.spray_trace_rays_device(ctx.h, None, None, None, None, 0, None, None, None, None, 0, 0) == 0
        assert lib.spray_trace_rays_device(ctx.h, None, None, None, None, 4, None, None, None, None, 0, 0) == 1
        assert lib.spray_trace_rays_device(ctx.h, None, None, None, None, 4, None, None, None, None, 0, 9) == 1
        assert lib.spray_trace_rays_device(ctx.h, None, None, None, None, 1 << 31, None, None, None, None,
                                           0, 0) == 5
