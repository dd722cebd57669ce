"""CPU checks of libspray_gpu.so: it loads, exports every entry point that
include/spray_gpu.h declares, keeps the reference's configuration defaults
(render/config.cc:33-73), reports a missing device as a status code, and its
host-side scene generator writes watertight, consistently oriented meshes."""
import ctypes as C
import math
import os
import re
from collections import Counter

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spray_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spray_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("spray_create", "spray_destroy", "spray_last_error", "spray_load_scene",
                 "spray_set_camera", "spray_trace_frame", "spray_trace_rays", "spray_set_schedule",
                 "spray_predictor_train", "spray_predictor_load", "spray_predictor_save"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from spray_amd import _lib as L
    lib = L.load()
    names = declared_functions()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding types every declared function
    assert set(names) <= set(L.PROTOTYPES), set(names) - set(L.PROTOTYPES)


def test_config_defaults_follow_reference():
    import spray_amd as sa
    cfg = sa.default_config()
    assert (cfg.image_w, cfg.image_h, cfg.fov) == (400, 400, 90.0)
    assert (cfg.pixel_samples, cfg.ao_samples, cfg.ao_mode, cfg.bounces) == (4, 8, 0, 1)
    assert cfg.max_samples_per_rank == 1024 * 1024 and cfg.nthreads == 1
    assert list(cfg.ks) == [0.0, 0.0, 0.0] and cfg.shininess == 1.0
    assert math.isinf(cfg.ao_max_dist) and cfg.rr_start == 0 and cfg.schedule == 0


def test_create_without_device_is_a_status_code():
    """No HIP device in this process (CPU run): spray_create reports
    SPRAY_ERR_DEVICE with a message instead of aborting."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from spray_amd import _lib as L
    import spray_amd as sa
    lib = L.load()
    cfg = sa.default_config()
    h = C.c_void_p()
    rc = lib.spray_create(C.byref(cfg), C.byref(h))
    assert rc == L.SPRAY_ERR_DEVICE and not h.value
    assert b"device" in lib.spray_last_error(None)
    with pytest.raises(sa.SprayError):
        sa.Context(cfg)


def test_create_rejects_invalid_config():
    from spray_amd import _lib as L
    import spray_amd as sa
    lib = L.load()
    for field, bad in (("image_w", 0), ("pixel_samples", 0), ("bounces", -1), ("schedule", 7)):
        cfg = sa.default_config()
        setattr(cfg, field, bad)
        h = C.c_void_p()
        assert lib.spray_create(C.byref(cfg), C.byref(h)) == L.SPRAY_ERR_INVALID, field
    assert lib.spray_create(None, C.byref(C.c_void_p())) == L.SPRAY_ERR_INVALID


def read_ascii_or_binary_ply(path):
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode().split("\n")
    nv = int([l for l in head if l.startswith("element vertex")][0].split()[2])
    nf = int([l for l in head if l.startswith("element face")][0].split()[2])
    props = [l.split()[-1] for l in head if l.startswith("property") and "list" not in l]
    assert "binary_little_endian" in head[1]
    vd = [(p, "<f4" if p in "xyz" else "u1") for p in props]
    v = np.frombuffer(data, dtype=vd, count=nv, offset=end)
    fo = end + v.nbytes
    f = np.frombuffer(data, dtype=[("n", "u1"), ("i", "<i4", 3)], count=nf, offset=fo)
    assert (f["n"] == 3).all()
    xyz = np.stack([v["x"], v["y"], v["z"]], 1)
    return xyz, f["i"].astype(np.int64), v


@pytest.fixture(scope="module")
def generated(tmp_path_factory):
    import spray_amd as sa
    d = str(tmp_path_factory.mktemp("gen"))
    tris = sa.scenegen_write(d, "g", 2, 2, 1, 10, rgb=0xB3B3B3, extra_lines="light point 1 3 1 1 1 1")
    return d, tris


def test_scenegen_declared_counts_and_bounds(generated):
    d, tris = generated
    text = open(os.path.join(d, "g.spray")).read()
    assert "light point 1 3 1 1 1 1" in text
    blocks = text.split("\ndomain\n")[1:]
    assert len(blocks) == 4
    total = 0
    for i, b in enumerate(blocks):
        kv = {l.split()[0]: l.split()[1:] for l in b.strip().split("\n") if l.strip()}
        xyz, faces, v = read_ascii_or_binary_ply(os.path.join(d, kv["file"][0]))
        assert int(kv["vertex"][0]) == len(xyz) and int(kv["face"][0]) == len(faces)
        bound = np.array([float(x) for x in kv["bound"]], np.float32)
        assert np.array_equal(bound[:3], xyz.min(0)) and np.array_equal(bound[3:], xyz.max(0))
        ii, jj = i % 2, i // 2
        assert (xyz >= np.array([ii, jj, 0]) - 1e-6).all() and (xyz <= np.array([ii + 1, jj + 1, 1]) + 1e-6).all()
        assert (v["red"] == 0xB3).all() and (v["blue"] == 0xB3).all()
        total += len(faces)
    assert total == tris


def test_scenegen_meshes_are_watertight_and_oriented(generated):
    """Inside a domain every edge is shared by exactly two triangles that
    traverse it in opposite directions; open edges lie on the domain's cube
    faces and match the neighbour domain's open edges (watertight across
    domains)."""
    d, _ = generated
    open_edges = {}
    for i in range(4):
        xyz, faces, _ = read_ascii_or_binary_ply(os.path.join(d, f"g_d{i:04d}.ply"))
        key = [tuple(np.round(p, 6)) for p in xyz.astype(np.float64)]
        directed = Counter()
        for a, b, c in faces:
            for u, w in ((a, b), (b, c), (c, a)):
                directed[(key[u], key[w])] += 1
        assert max(directed.values()) == 1  # no edge traversed twice the same way
        undirected = Counter()
        for (p, q), n in directed.items():
            undirected[frozenset((p, q))] += n
        assert max(undirected.values()) <= 2
        ii, jj = i % 2, i // 2
        for e, n in undirected.items():
            if n == 1:
                p, q = tuple(e)
                on_face = any(abs(p[k] - q[k]) < 1e-9 and abs(p[k] - s) < 1e-6
                              for k, s in ((0, ii), (0, ii + 1), (1, jj), (1, jj + 1), (2, 0), (2, 1))
                              for _ in [0])
                assert on_face, (i, p, q)
                open_edges.setdefault(e, []).append(i)
    # open edges on the shared faces x=1 and y=1 appear in exactly two domains
    shared = [e for e in open_edges
              if all(abs(p[0] - 1) < 1e-6 for p in e) or all(abs(p[1] - 1) < 1e-6 for p in e)]
    assert shared
    assert all(len(open_edges[e]) == 2 for e in shared)


def test_scenegen_bad_arguments():
    import spray_amd as sa
    with pytest.raises(sa.SprayError):
        sa.scenegen_write("/nonexistent_dir_for_test", "x", 1, 1, 1, 8)
    with pytest.raises(sa.SprayError):
        sa.scenegen_write("/tmp", "x", 0, 1, 1, 8)


def test_write_ppm_matches_reference_format(tmp_path):
    """display/image.h:183-204: P3, max 1023, top row first, values
    clamp(v * 1023, 0, 1023) truncated."""
    import spray_amd as sa
    img = np.zeros((2, 3, 4), np.float32)
    img[0, 0, :3] = [0.0, 0.5, 1.0]        # bottom-left
    img[0, 2, :3] = [2.0, -1.0, 0.9999]     # bottom-right: clamped
    img[1, 1, :3] = [0.25, 1e-4, 0.001]     # top-middle
    img[1, 2, 0] = np.nan
    p = tmp_path / "o.ppm"
    sa.write_ppm(p, img)
    lines = p.read_text().split("\n")
    assert lines[:3] == ["P3", "3 2", "1023"]
    px = [tuple(int(x) for x in l.split()) for l in lines[3:] if l]
    assert len(px) == 6
    # top row (y = 1) first
    assert px[0] == (0, 0, 0) and px[1] == (255, 0, 1) and px[2] == (0, 0, 0)
    assert px[3] == (0, 511, 1023) and px[4] == (0, 0, 0) and px[5] == (1023, 0, 1022)
    with pytest.raises(sa.SprayError):
        sa.write_ppm("/nonexistent_dir_for_test/x.ppm", img)


def test_ply_info_file_format_matches_header_reader(tmp_path):
    """write_ply_info prints what apps/ply_header_reader.cc:90-104 printed
    into examples/wavelet/wavelet_example.spray: counts, and float bounds
    with 16 significant digits (values from the oracle's PLY reader)."""
    import spray_amd as sa
    from test_oracle_geometry import GOLD, ply_read, spray_declared
    for d in spray_declared(os.path.join(GOLD, "wavelet_example.spray")):
        nv, nf, _, lo, hi = ply_read(os.path.join(GOLD, d["file"][0]))
        out = tmp_path / "info.txt"
        sa.write_ply_info(d["file"][0], out, (nv, nf, lo, hi))
        assert out.read_text().splitlines() == [
            f"file {d['file'][0]}", f"vertex {d['vertex'][0]}", f"face {d['face'][0]}",
            "bound " + " ".join(d["bound"])]
