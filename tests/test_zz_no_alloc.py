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


def _device_query(ctx, torch, org, dirs, tnear, tfar, mode, dom):
    dev = torch.device("cuda:0")
    n = len(org)
    t_org = torch.from_numpy(org).to(dev).contiguous()
    t_dir = torch.from_numpy(dirs).to(dev).contiguous()
    t_tn = torch.from_numpy(tnear).to(dev)
    t_tf = torch.from_numpy(tfar).to(dev)
    # outputs start as garbage so that nothing unwritten can pass
    t_t = torch.full((n,), -7.0, device=dev)
    t_u = torch.full((n,), -7.0, device=dev)
    t_v = torch.full((n,), -7.0, device=dev)
    t_p = torch.full((n,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    fp = lambda x: C.cast(C.c_void_p(x.data_ptr()), C.POINTER(C.c_float))
    ip = lambda x: C.cast(C.c_void_p(x.data_ptr()), C.POINTER(C.c_int))
    closest = mode == 0
    rc = ctx.lib.spray_trace_rays_device(ctx.h, fp(t_org), fp(t_dir), fp(t_tn), fp(t_tf), n,
                                         fp(t_t) if closest else None, fp(t_u) if closest else None,
                                         fp(t_v) if closest else None, ip(t_p), mode, dom)
    assert rc == 0, ctx.lib.spray_last_error(ctx.h)
    torch.cuda.synchronize()
    return t_t.cpu().numpy(), t_u.cpu().numpy(), t_v.cpu().numpy(), t_p.cpu().numpy()


def test_no_alloc_matches_trace_rays(tmp_path):
    import torch
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
                t, u, v, p = _device_query(ctx, torch, org, dirs, tnear, tfar, mode, dom)
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
