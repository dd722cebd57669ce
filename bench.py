"""Benchmark: SpRay per-rank tracing on one MI355X.

Default workload ("c3", the 256-domain path-tracing scene): 256 domains in an
8x8x4 grid of unit cubes, each a marching-cubes isosurface of
f = sin(a x)sin(a y)sin(a z) + 0.3 sin(b(x+y+z)) with per-domain a in [4, 8],
b in [12, 20] and a log-uniform triangle budget in 10k..2M (seeded, rescaled
to ~120M triangles in total), path tracing with albedo 0.7 (vertex colour
179), one rectangular area light, 1920x1080, 16 spp, depth 6, Russian
roulette from bounce 3, fp32 rays.  A step is one frame (33 blocking tiles of
at most 1,048,576 samples).  Synthetic geometry generated in-process.  The
learned schedule's predictor is trained on 50M (ray, domain, hit) samples
logged by plain-schedule frames.

Other workloads (--workload): "c2" the 64-domain scene (4x4x4 grid, 64^3
cells per domain, a = 6, b = 17, ~5.8M triangles) path traced at 1024x1024,
8 spp, depth 5; "c1" AO on that geometry (512x512, 1 spp, 8 AO rays, max
distance 0.5); "c5" the in-situ stress case:
1024 domains in a 16x8x8 grid, ~500k triangles each (160^3 cells), generated
on the GPU and remeshed on the GPU before every timed frame with the field
phase advancing 0.05 per frame, 3840x2160, 16 spp, depth 8; the learned
schedule fine-tunes the predictor online on every frame's hit log.

Prints one JSON line.  `value` is Mrays/s of useful ray queries (eye +
non-stale bounce + non-stale shadow rays, identical for every schedule) over
the wall time of the timed frames, with the scene resident in HBM.
"""
import argparse
import json
import math
import os
import shutil
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C2_CAM = dict(pos=[2.0, 2.0, 2.0 + 2.0 * math.sqrt(3.0)], lookat=[2.0, 2.0, 2.0], up=[0.0, 1.0, 0.0])
RECT_LIGHT = "light rect 0.5 4.5 0.5 3 0 0 0 0 3 4 4 4"
HBM_PEAK_GBS = 8000.0
TRAIN_SAMPLES = 50_000_000  # (ray, domain, hit) samples the predictor trains on

# name -> scene generator arguments, render settings, camera (None: the
# reference's default framing of the scene bound), the blocking tiles
# [begin, end) the CPU baseline times by default (C3: two tiles across the
# middle of the image, about 2 x 1.01M samples)
WORKLOADS = {
    "c1": dict(grid=(4, 4, 4), gen=dict(res=64, rgb=0xFFFFFF, extra_lines="light point 2 6 2 1 1 1"),
               render=dict(image_w=512, image_h=512, pixel_samples=1, ao_mode=1, ao_samples=8, bounces=1,
                           ao_max_dist=0.5),
               cam=C2_CAM, cpu_tiles=(0, 1),
               title="C1 ambient occlusion: 64 domains (4x4x4 MC isosurfaces, {tris} tris), 512x512, 1 spp, "
                     "8 AO rays, max dist 0.5"),
    "c2": dict(grid=(4, 4, 4), gen=dict(res=64, rgb=0xB3B3B3, extra_lines=RECT_LIGHT),
               render=dict(image_w=1024, image_h=1024, pixel_samples=8, ao_mode=0, ao_samples=1, bounces=5,
                           rr_start=3),
               cam=C2_CAM, cpu_tiles=(0, 8),
               title="C2 path tracing: 64 domains (4x4x4 MC isosurfaces, {tris} tris), {w}x{h}, {spp} spp, "
                     "depth 5, RR>=3, rect light"),
    "c3": dict(grid=(8, 8, 4), gen=dict(res=0, per_domain_random=True, seed=0xC3, tri_total=120e6, rgb=0xB3B3B3,
                                        extra_lines="light rect 1 8.5 0.5 6 0 0 0 0 3 4 4 4"),
               render=dict(image_w=1920, image_h=1080, pixel_samples=16, ao_mode=0, ao_samples=1, bounces=6,
                           rr_start=3),
               cam=None, cpu_tiles=(16, 18),
               title="C3 path tracing: 256 domains (8x8x4 MC isosurfaces, log-uniform 10k-2M tris each, "
                     "{tris} tris), {w}x{h}, {spp} spp, depth 6, RR>=3, rect light"),
    "c5": dict(grid=(16, 8, 8), gen=dict(res=160, rgb=0xB3B3B3,
                                         extra_lines="light rect 2 1 8.5 0 6 0 12 0 0 4 4 4"),
               render=dict(image_w=3840, image_h=2160, pixel_samples=16, ao_mode=0, ao_samples=1, bounces=8,
                           rr_start=3),
               cam=None, cpu_tiles=None, remesh=0.05,
               title="C5 in-situ: 1024 domains (16x8x8 MC isosurfaces remeshed on the GPU every frame, "
                     "phase +0.05 per frame, {tris} tris), {w}x{h}, {spp} spp, depth 8, RR>=3, rect light facing the "
                     "scene's front (+z) face"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=0, help="square image size (0: the workload's)")
    ap.add_argument("--spp", type=int, default=0, help="samples per pixel (0: the workload's)")
    ap.add_argument("--schedules", default="plain,batched,learned")
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--cpu-tiles", type=int, default=0,
                    help="blocking tiles timed on the CPU, from the first (0: the workload's default range)")
    ap.add_argument("--cache-domains", type=int, default=-1,
                    help="domains kept in HBM (< 0: all); a partial cache pages domains (plain schedule)")
    ap.add_argument("--tiles-in-flight", type=int, default=0,
                    help="blocking tiles per batch of the batched forms (0: about 8M samples)")
    ap.add_argument("--bvh", default="lbvh", choices=["ploc", "lbvh"], help="per-domain BVH builder")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default batch parameters")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every pass directly (kernel-level profilers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (rocprof counter calibration)")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "launches": 3}))


def algorithmic_bytes(st):
    """Bytes one frame's trace kernels must move: per (ray, domain) visit a
    32 B ray record, 16 B hit record and 8 B queue entry; per BVH node fetch
    128 B; per triangle test 48 B; per accepted shading vertex 80 B ray write
    + 16 B queue + 16 B history."""
    visits = st["radiance_visits"] + st["shadow_visits"]
    return 56 * visits + 128 * st["node_visits"] + 48 * st["tri_tests"] + 112 * st["shades"]


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU image-stripe "
              "sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import numpy as np
    import spray_amd as sa

    wl = WORKLOADS[args.workload]
    remesh_step = wl.get("remesh")
    tmp = tempfile.mkdtemp(prefix=f"spray_{args.workload}_")
    t0 = time.time()
    nx, ny, nz = wl["grid"]
    g = dict(wl["gen"])
    res = g.pop("res")
    tris = 0
    if not remesh_step:  # in-situ scenes are generated on the GPU below, with no files
        tris = sa.scenegen_write(tmp, args.workload, nx, ny, nz, res, **g)
    gen_s = time.time() - t0
    kw = dict(wl["render"])
    if args.size:
        kw["image_w"] = kw["image_h"] = args.size
    if args.spp:
        kw["pixel_samples"] = args.spp
    W, H = kw["image_w"], kw["image_h"]
    cam = wl["cam"]
    camkw = dict(has_camera=1, camera_pos=cam["pos"], camera_lookat=cam["lookat"],
                 camera_up=cam["up"]) if cam else dict(has_camera=0)
    cfg = sa.default_config(nthreads=min(16, os.cpu_count() or 1), **camkw, **kw)
    ctx = sa.Context(cfg)
    ctx.set_bvh_builder({"lbvh": 0, "ploc": 1}[args.bvh])
    ctx.set_cache_size(args.cache_domains)
    ctx.set_tiles_in_flight(args.tiles_in_flight)
    t0 = time.time()
    if remesh_step:
        ctx.scenegen_load(args.workload, nx, ny, nz, res, **g)
        tris = sum(ctx.mesh_size(i)[1] for i in range(ctx.num_domains))
    else:
        ctx.load_scene(os.path.join(tmp, args.workload + ".spray"), tmp)
    load_s = time.time() - t0
    print(f"# {args.workload}: {tris} triangles, generated in {gen_s:.1f} s, loaded in {load_s:.1f} s",
          file=sys.stderr, flush=True)
    if cam is None:
        cam = default_framing(ctx)
        ctx.set_camera(cam["pos"], cam["lookat"], cam["up"], 90.0)

    schedules = {"plain": 0, "batched": 1, "learned": 2}
    wanted = [s for s in args.schedules.split(",") if s]
    results = {}
    images = {}
    for name in wanted:
        sched = schedules[name]
        try:
            ctx.set_schedule(sched)
            if remesh_step:
                ctx.remesh(0.0)
            if name == "learned":
                # train the predictor on hit logs of plain-schedule frames
                # (in-situ: of a batched frame, the plain one taking minutes)
                ctx.set_schedule(1 if remesh_step else 0)
                ctx.predictor_log(True, TRAIN_SAMPLES)
                for _ in range(4):  # plain frames until the log holds its 50M samples
                    ctx.trace_frame(want_image=False)
                    if ctx.predictor_num_samples() >= TRAIN_SAMPLES:
                        break
                ctx.predictor_log(False)
                print(f"# learned: training on {ctx.predictor_num_samples()} logged samples", file=sys.stderr,
                      flush=True)
                ctx.predictor_train(2, 4096, 2e-3)
                ctx.set_schedule(sched)
            if name != "plain" and not args.no_autotune:
                ctx.autotune()
            ctx.profile_kernels(args.no_graph)
        except sa.SprayError as e:
            results[name] = {"error": str(e)}
            continue
        for k in range(args.warmup):
            ctx.trace_frame(want_image=False)
            print(f"# {name}: warmup frame {k + 1}/{args.warmup}", file=sys.stderr, flush=True)
        frame_ms, walls, remesh_s, train_s = [], [], [], []
        st = None
        t_start = time.perf_counter()
        for step in range(args.steps):
            if remesh_step:  # the in-situ frame: remesh at the next phase, then trace
                tw = time.perf_counter()
                ctx.remesh(remesh_step * (step + 1))
                remesh_s.append(time.perf_counter() - tw)
                print(f"# {name}: remesh {step + 1}/{args.steps} {1000.0 * remesh_s[-1]:.1f} ms",
                      file=sys.stderr, flush=True)
                if name == "learned":
                    ctx.predictor_log(True, 4_000_000)
            tw = time.perf_counter()
            img, st = ctx.trace_frame(want_image=False)
            walls.append(time.perf_counter() - tw)
            frame_ms.append(st["trace_ms"])
            print(f"# {name}: frame {len(walls)}/{args.steps} {1000.0 * walls[-1]:.1f} ms", file=sys.stderr,
                  flush=True)
            if remesh_step and name == "learned":  # online fine-tuning on this frame's hit log
                tw = time.perf_counter()
                ctx.predictor_log(False)
                ctx.predictor_train(1, 4096, 1e-3)
                train_s.append(time.perf_counter() - tw)
        total_s = time.perf_counter() - t_start
        frames_s = sum(walls)
        img, st_img = ctx.trace_frame(want_image=True)
        images[name] = img
        lit = float((img[..., :3].max(axis=2) > 0).mean())
        # kernel timing of the dominant (trace) kernel, on the tracer's stream
        ctx.profile_kernels(True)
        ctx.kernel_time(reset=True)
        _, st_prof = ctx.trace_frame(want_image=False)
        launches, kern_ms = ctx.kernel_time(reset=True)
        ctx.profile_kernels(args.no_graph)
        # in-situ: the frame rate counts the frames; remesh and training are reported beside it
        timed_s = frames_s if remesh_step else total_s
        print(f"# {name}: {st['useful_rays'] * args.steps / timed_s / 1e6:.1f} Mrays/s, "
              f"{1000.0 * timed_s / args.steps:.1f} ms/frame", file=sys.stderr, flush=True)
        results[name] = dict(
            mrays_s=st["useful_rays"] * args.steps / timed_s / 1e6,
            ms_per_frame=1000.0 * timed_s / args.steps,
            device_ms_median=statistics.median(frame_ms),
            useful_rays=st["useful_rays"], radiance_visits=st["radiance_visits"],
            shadow_visits=st["shadow_visits"], shades=st["shades"], final_hits=st["final_hits"],
            queue_entries=st["queue_entries"], pruned=st["pruned_entries"],
            deferred=st["deferred_entries"], passes=st["passes"], regrows=st["regrows"],
            node_visits=st["node_visits"], tri_tests=st["tri_tests"],
            trace_kernel_launches=launches, trace_kernel_ms=kern_ms,
            trace_kernel_bytes=algorithmic_bytes(st_prof),
            cache_hits=st["cache_hits"], cache_misses=st["cache_misses"],
            cache_bytes=st["cache_bytes"],
            cache_h2d_gbs=st["cache_bytes"] / (st["trace_ms"] * 1e6) if st["trace_ms"] > 0 else 0.0,
            image_mean=float(img[..., :3].mean()), image_lit_fraction=lit,
        )
        if remesh_step:
            results[name].update(
                remesh_ms=round(1000.0 * statistics.mean(remesh_s), 1),
                train_ms=round(1000.0 * statistics.mean(train_s), 1) if train_s else 0.0,
                mrays_s_with_remesh=st["useful_rays"] * args.steps / total_s / 1e6,
                final_phase=remesh_step * args.steps)
    ref_name = "plain" if "plain" in images else next(iter(images), None)
    for name, img in images.items():
        if ref_name and name != ref_name:
            diff = np.abs(img[..., :3] - images[ref_name][..., :3])
            results[name]["image_vs"] = ref_name  # plain when it ran, else the first schedule
            results[name]["image_rmse"] = float(np.sqrt((diff ** 2).mean()))
            results[name]["image_max"] = float(diff.max())

    ok = {k: v for k, v in results.items() if "mrays_s" in v}
    best = max(ok, key=lambda k: ok[k]["mrays_s"]) if ok else None
    # roofline of the dominant kernel (the trace kernel) of the reported schedule
    roof = None
    if best:
        r = ok[best]
        avg_ms = r["trace_kernel_ms"] / max(1, r["trace_kernel_launches"])
        bytes_per_launch = r["trace_kernel_bytes"] / max(1, r["trace_kernel_launches"])
        achieved = bytes_per_launch / (avg_ms / 1000.0) / 1e9 if avg_ms > 0 else 0.0
        traffic = measured_traffic(args.workload, best)
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic["bytes_per_dispatch"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                "kernel": "spray::kTraverse",
                "avg_launch_ms": round(avg_ms, 4), "launches_per_frame": r["trace_kernel_launches"],
                "algorithmic_bytes_per_launch": int(bytes_per_launch)}

    cpu = None
    if not args.no_cpu and not remesh_step:
        tiles = (0, args.cpu_tiles) if args.cpu_tiles else wl["cpu_tiles"]
        cpu = cpu_baseline(tmp, args.workload, kw, cam, tiles)

    line = {
        "metric": "Mrays/s (AO + path) on 1 MI355X; learned vs plain schedule at equal image",
        "value": round(ok[best]["mrays_s"], 3) if best else None,
        "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ok[best]["ms_per_frame"], 3) if best else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl["title"].format(tris=tris, w=W, h=H, spp=kw["pixel_samples"]),
                   "schedule": best, "triangles": tris, "scene_gen_s": round(gen_s, 2),
                   "load_build_s": round(load_s, 2)},
        "schedules": results,
        # the metric's comparison: Mrays/s of each schedule over the plain one (equal images above)
        "learned_vs_plain": ratio(ok, "learned", "plain"),
        "batched_vs_plain": ratio(ok, "batched", "plain"),
        "learned_vs_batched": ratio(ok, "learned", "batched"),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()
    shutil.rmtree(tmp, ignore_errors=True)


def ratio(ok, a, b):
    if a in ok and b in ok and ok[b]["mrays_s"] > 0:
        return round(ok[a]["mrays_s"] / ok[b]["mrays_s"], 4)
    return None


def measured_traffic(workload, schedule):
    """HBM bytes per trace-kernel dispatch measured by two rocprofv3 --pmc
    runs (FETCH_SIZE, WRITE_SIZE) of this benchmark, scaled by the same
    counters on a copy of known size; recorded as `key: value` lines."""
    path = os.path.join(ROOT, "profiles", f"kTraverse_{workload}_{schedule}_traffic.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
    try:
        return {"bytes_per_dispatch": int(float(kv["traffic_bytes_per_dispatch"])),
                "source": os.path.relpath(path, ROOT)}
    except (KeyError, ValueError):
        return None


def default_framing(ctx):
    """The reference's camera for a scene without one (spray_renderer.inl:
    95-104): look at the bound's centre from +z at half the diagonal, in the
    library's float32 order of operations."""
    import numpy as np
    f = np.float32
    lo = np.full(3, np.inf, np.float32)
    hi = np.full(3, -np.inf, np.float32)
    for i in range(ctx.num_domains):
        _, _, a, b = ctx.domain_info(i)
        lo, hi = np.minimum(lo, a), np.maximum(hi, b)
    ext = (hi - lo).astype(np.float32)
    look = (0.5 * (lo + hi).astype(np.float64)).astype(np.float32)
    ln = f(np.sqrt(f(f(ext[0] * ext[0]) + f(ext[1] * ext[1])) + f(ext[2] * ext[2])))
    z = f(f(ln * f(0.5)) * f(1.0))
    pos = [float(look[0]), float(look[1]), float(f(look[2] + z))]
    return dict(pos=pos, lookat=[float(x) for x in look], up=[0.0, 1.0, 0.0])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene_dir, name, kw, cam, tiles):
    """The oracle (the CPU restatement of the reference's OpenMP tracer) on
    blocking tiles [tiles[0], tiles[1]) of the same frame, on the host cores
    this process may run on (at most 16: the GPU box's share)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.time()
    ref = orc.Scene(os.path.join(scene_dir, name + ".spray"), scene_dir, nthreads=threads)
    load_s = time.time() - t0
    b, e = tiles
    _, st = ref.render(nthreads=threads, cam_pos=cam["pos"], cam_lookat=cam["lookat"],
                       cam_up=cam["up"], tile_begin=b, tile_end=e, **kw)
    return {"value": round(st["useful_rays"] / st["seconds"] / 1e6, 4), "unit": "Mrays/s",
            "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"blocking tiles {b}..{e - 1} of {orc.num_tiles(**kw)} of the same frame "
                      f"({st['useful_rays']} useful rays in {st['seconds']:.1f} s; scene load and SAH "
                      f"BVH build {load_s:.1f} s, not timed), speculative OpenMP restatement of the "
                      f"reference tracer with its own SAH BVH"}


if __name__ == "__main__":
    main()
