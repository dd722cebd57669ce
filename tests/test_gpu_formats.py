"""GPU tests of the on-disk format path: binary PLY bodies decoded in HBM,
the ply_header_reader query, the GPU marching-cubes generator and remesher,
and the scene writer.

Pinned by the reference's own fixtures (examples/wavelet: the vertex, face
and 16-digit bound lines apps/ply_header_reader.cc:90-104 wrote into
wavelet_example.spray; the binary VTK PLY of wavelets64 under 64
translations), by the oracle's PLY reader and transform (io/ply_loader.cc,
render/trimesh_buffer.cc:139-156), and by the CPU generator
(spray_scenegen_write) whose meshes the GPU generator must reproduce.
"""
import math
import os

import numpy as np
import pytest

import orc
from test_oracle_geometry import GOLD, ply_read, spray_declared

pytestmark = pytest.mark.gpu

LIGHT = "light rect 0.5 2.5 0.5 1 0 0 0 0 1 4 4 4"
CAM = dict(camera_pos=[1.0, 1.0, 1.0 + 2.0 * math.sqrt(3.0)], camera_lookat=[1.0, 1.0, 1.0],
           camera_up=[0.0, 1.0, 0.0])
PT = dict(image_w=96, image_h=96, pixel_samples=2, ao_mode=0, ao_samples=1, bounces=5, rr_start=3,
          max_samples_per_rank=4608, has_camera=1, nthreads=8, **CAM)


def _ctx(**kw):
    import spray_amd as sa
    cfg = dict(PT)
    cfg.update(kw)
    return sa.Context(sa.default_config(**cfg))


def write_ply_layout(path, verts, faces, rgb=None, binary=True, faces_first=False,
                     list_type="uchar", index_type="int", counts=None, truncate=0):
    """A PLY file in any layout the loader accepts; `counts` overrides the
    per-face list counts (a count of 4 writes a fourth index)."""
    verts = np.asarray(verts, np.float32).reshape(-1, 3)
    faces = np.asarray(faces, np.int64).reshape(-1, 3)
    sizes = {"uchar": ("u1", 1), "ushort": ("<u2", 2), "int": ("<i4", 4), "uint": ("<u4", 4)}
    vhead = [f"element vertex {len(verts)}", "property float x", "property float y", "property float z"]
    if rgb is not None:
        vhead += ["property uchar red", "property uchar green", "property uchar blue"]
    fhead = [f"element face {len(faces)}", f"property list {list_type} {index_type} vertex_indices"]
    fmt = "format binary_little_endian 1.0" if binary else "format ascii 1.0"
    head = ["ply", fmt] + (fhead + vhead if faces_first else vhead + fhead) + ["end_header"]
    counts = [3] * len(faces) if counts is None else counts
    if binary:
        if rgb is None:
            vb = verts.astype("<f4").tobytes()
        else:
            rec = np.zeros(len(verts), dtype=[("p", "<f4", 3), ("c", "u1", 3)])
            rec["p"] = verts
            rec["c"] = np.asarray(rgb, np.uint8).reshape(-1, 3)
            vb = rec.tobytes()
        lt, it = sizes[list_type][0], sizes[index_type][0]
        fb = b""
        for n, t in zip(counts, faces):
            idx = list(t) + [0] * (n - 3)
            fb += np.array([n], lt).tobytes() + np.array(idx, it).tobytes()
        body = fb + vb if faces_first else vb + fb
        data = ("\n".join(head) + "\n").encode() + body
        with open(path, "wb") as fp:
            fp.write(data[:len(data) - truncate])
        return
    vl = []
    for i, v in enumerate(verts):
        s = f"{float(v[0])!r} {float(v[1])!r} {float(v[2])!r}"
        if rgb is not None:
            c = np.asarray(rgb).reshape(-1, 3)[i]
            s += f" {int(c[0])} {int(c[1])} {int(c[2])}"
        vl.append(s)
    fl = [" ".join(str(int(x)) for x in [n] + list(t) + [0] * (n - 3)) for n, t in zip(counts, faces)]
    with open(path, "w") as fp:
        fp.write("\n".join(head + (fl + vl if faces_first else vl + fl)) + "\n")


def one_domain_scene(d, ply, verts, nf, extra=""):
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    path = os.path.join(d, ply + ".spray")
    with open(path, "w") as fp:
        fp.write(f"{LIGHT}\ndomain\nfile {ply}\nmtl diffuse 1 1 1\n{extra}"
                 f"bound {' '.join(repr(float(x)) for x in list(lo) + list(hi))}\n"
                 f"vertex {len(v)}\nface {nf}\n")
    return path


def grid_mesh(n=9, seed=0):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(n, dtype=np.float32), np.arange(n, dtype=np.float32))
    z = rng.uniform(-0.2, 0.2, xs.shape).astype(np.float32)
    v = np.stack([xs.ravel() / (n - 1), ys.ravel() / (n - 1), z.ravel()], 1).astype(np.float32)
    f = []
    for j in range(n - 1):
        for i in range(n - 1):
            a = j * n + i
            f += [[a, a + 1, a + n + 1], [a, a + n + 1, a + n]]
    rgb = rng.integers(0, 256, (len(v), 3)).astype(np.uint8)
    return v, np.array(f, np.int64), rgb


# ---------------------------------------------------------------------------
# ply_header_reader

def test_ply_info_reproduces_header_reader_lines(tmp_path):
    """The counts and the 16-digit bound line apps/ply_header_reader.cc wrote
    for the wavelet example (ASCII with colour) come out character for
    character; the binary wavelet gives the oracle's counts and bounds."""
    import spray_amd as sa
    with _ctx() as ctx:
        for d in spray_declared(os.path.join(GOLD, "wavelet_example.spray")):
            info = ctx.ply_info(os.path.join(GOLD, d["file"][0]))
            out = tmp_path / (d["file"][0] + ".txt")
            sa.write_ply_info(d["file"][0], out, info)
            lines = out.read_text().splitlines()
            assert lines == [f"file {d['file'][0]}", f"vertex {d['vertex'][0]}", f"face {d['face'][0]}",
                             "bound " + " ".join(d["bound"])]
        nv, nf, lo, hi = ctx.ply_info(os.path.join(GOLD, "wavelet.ply"))
        rnv, rnf, _, rlo, rhi = ply_read(os.path.join(GOLD, "wavelet.ply"))
        assert (nv, nf) == (rnv, rnf)
        assert np.array_equal(lo, rlo) and np.array_equal(hi, rhi)


# ---------------------------------------------------------------------------
# binary PLY decoded on the GPU

def test_binary_lattice_decodes_to_oracle_meshes():
    """64 translated copies of the binary wavelet PLY: every domain's
    vertices (translated on the GPU), faces and colours equal the oracle's."""
    path = os.path.join(GOLD, "wavelets64.spray")
    ref = orc.Scene(path, GOLD, nthreads=8)
    with _ctx() as ctx:
        ctx.load_scene(path, GOLD)
        assert ctx.num_domains == 64
        for d in range(64):
            v, f, rgb = ctx.domain_mesh(d)
            rv, rf, rrgb, rn = ref.mesh(d)
            assert np.array_equal(v, rv), d
            assert np.array_equal(f, rf) and np.array_equal(rgb, rrgb), d
            assert np.array_equal(ctx.normals(d), rn), d


LAYOUTS = [
    dict(binary=True),
    dict(binary=True, rgb=False),
    dict(binary=True, faces_first=True),
    dict(binary=True, list_type="int", index_type="uint"),
    dict(binary=True, list_type="ushort", index_type="ushort", faces_first=True, rgb=False),
    dict(binary=True, list_type="uchar", index_type="uchar"),
    dict(binary=False),
]


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "-".join(f"{k}={v}" for k, v in l.items()))
def test_ply_layouts_match_oracle(tmp_path, layout):
    """Element order, colour and list/index types: the GPU decode (binary)
    and the host parser (ASCII) give the oracle's mesh, with a rotation and
    translation applied to the vertices."""
    layout = dict(layout)
    with_rgb = layout.pop("rgb", True)
    v, f, rgb = grid_mesh()
    write_ply_layout(tmp_path / "m.ply", v, f, rgb if with_rgb else None, **layout)
    path = one_domain_scene(str(tmp_path), "m.ply", v, len(f),
                            extra="rotate z 30\ntranslate 0.25 -0.5 1\n")
    ref = orc.Scene(path, str(tmp_path), nthreads=2)
    with _ctx() as ctx:
        ctx.load_scene(path, str(tmp_path))
        gv, gf, grgb = ctx.domain_mesh(0)
    rv, rf, rrgb, _ = ref.mesh(0)
    assert np.array_equal(gv, rv) and np.array_equal(gf, rf) and np.array_equal(grgb, rrgb)
    if not with_rgb:
        assert not grgb.any()


@pytest.mark.parametrize("case,message", [
    (dict(counts_at={7: 4}), "non-triangle face"),
    (dict(bad_index_at=3), "face index out of range"),
    (dict(bad_index_at=2, counts_at={5: 4}), "face index out of range"),
    (dict(counts_at={1: 4}, bad_index_at=6), "non-triangle face"),
    (dict(truncate=5), "truncated faces"),
    (dict(truncate=5, faces_first=True), "truncated vertices"),
])
def test_binary_ply_errors(tmp_path, case, message):
    """The host parser's errors (io/ply_loader.cc:289 and the range check):
    the first bad face of the file decides the message; bad input never
    reaches the BVH build."""
    import spray_amd as sa
    v, f, rgb = grid_mesh()
    f = f.copy()
    counts = [3] * len(f)
    for i, n in case.get("counts_at", {}).items():
        counts[i] = n
    if "bad_index_at" in case:
        f[case["bad_index_at"], 1] = len(v) + 7
    write_ply_layout(tmp_path / "m.ply", v, f, rgb, counts=counts, truncate=case.get("truncate", 0),
                     faces_first=case.get("faces_first", False))
    path = one_domain_scene(str(tmp_path), "m.ply", v, len(f))
    with _ctx() as ctx:
        with pytest.raises(sa.SprayError) as ei:
            ctx.load_scene(path, str(tmp_path))
        assert ei.value.code == sa._lib.SPRAY_ERR_FORMAT
        assert message in str(ei.value)
        # the context still loads a good scene afterwards
        write_ply_layout(tmp_path / "g.ply", v, grid_mesh()[1], rgb)
        ctx.load_scene(one_domain_scene(str(tmp_path), "g.ply", v, len(f)), str(tmp_path))
        assert ctx.num_domains == 1


# ---------------------------------------------------------------------------
# GPU marching cubes, remeshing, scene writer

GEN = [
    dict(nx=2, ny=2, nz=2, res=24),
    dict(nx=3, ny=2, nz=1, res=0, per_domain_random=True, seed=11, tri_total=150000.0),
    # one domain of 128^3 cells: the marching-cubes kernels' grid-stride loops
    # (capped grids of 2048 blocks) run several passes
    dict(nx=1, ny=1, nz=1, res=128),
]


def _cpu_scene(d, name, gen, phase=0.0):
    import spray_amd as sa
    g = dict(gen)
    sa.scenegen_write(d, name, g.pop("nx"), g.pop("ny"), g.pop("nz"), g.pop("res"), phase=phase,
                      rgb=0xB3B3B3, extra_lines=LIGHT, **g)
    return os.path.join(d, name + ".spray")


def _gen(ctx, name, gen, phase=0.0):
    g = dict(gen)
    ctx.scenegen_load(name, g.pop("nx"), g.pop("ny"), g.pop("nz"), g.pop("res"), phase=phase,
                      rgb=0xB3B3B3, extra_lines=LIGHT, **g)


def _same_scene(a, b):
    assert a.num_domains == b.num_domains
    for d in range(a.num_domains):
        na, fa, loa, hia = a.domain_info(d)
        nb, fb, lob, hib = b.domain_info(d)
        assert (na, fa) == (nb, fb), d
        assert np.array_equal(loa, lob) and np.array_equal(hia, hib), d
        va, fa_, ca = a.domain_mesh(d)
        vb, fb_, cb = b.domain_mesh(d)
        assert np.array_equal(fa_, fb_) and np.array_equal(ca, cb), d
        assert np.array_equal(va, vb), (d, int((va != vb).sum()))


@pytest.mark.parametrize("gen", GEN, ids=["lattice", "log-uniform", "fine-domain"])
def test_gpu_generator_matches_cpu_generator(tmp_path, gen):
    """Marching cubes on the GPU gives the CPU generator's meshes, vertex and
    face order, domain bounds and counts (then the same frame)."""
    path = _cpu_scene(str(tmp_path), "s", gen)
    with _ctx(schedule=1) as a, _ctx(schedule=1) as b:
        _gen(a, "s", gen)
        b.load_scene(path, str(tmp_path))
        _same_scene(a, b)
        ia, sa_ = a.trace_frame()
        ib, sb = b.trace_frame()
    assert sa_["useful_rays"] == sb["useful_rays"] and sa_["final_hits"] == sb["final_hits"]
    err = np.abs(ia[..., :3] - ib[..., :3])
    assert float(np.sqrt((err ** 2).mean())) <= 1e-6 and float(err.max()) <= 1e-4


@pytest.mark.parametrize("gi,blocks", [(0, 1), (0, 3), (2, 2)], ids=["lattice-1", "lattice-3", "fine-2"])
def test_mc_verts_small_grid(tmp_path, monkeypatch, gi, blocks):
    """kMcVerts on a grid of a few blocks (SPRAY_MC_VERTS_BLOCKS): each wave
    walks several groups of cut-mask words, reusing its LDS slots, in the
    bounds pass and the vertex pass, and the meshes, bounds and counts are
    still the CPU generator's, before and after a remesh."""
    gen = GEN[gi]
    p00 = _cpu_scene(str(tmp_path), "p00", gen, phase=0.0)
    p05 = _cpu_scene(str(tmp_path), "p05", gen, phase=0.05)
    monkeypatch.setenv("SPRAY_MC_VERTS_BLOCKS", str(blocks))
    with _ctx(schedule=1) as a, _ctx(schedule=1) as b, _ctx(schedule=1) as c:
        _gen(a, "s", gen)
        b.load_scene(p00, str(tmp_path))
        _same_scene(a, b)
        a.remesh(0.05)
        c.load_scene(p05, str(tmp_path))
        _same_scene(a, c)


def test_remesh_in_place(tmp_path):
    """Remeshing at a new phase gives the scene generated at that phase
    (meshes, bounds, normals, frame), after a frame whose pass graphs were
    captured on the old pool; remeshing back restores the first scene."""
    gen = GEN[0]
    p05 = _cpu_scene(str(tmp_path), "p05", gen, phase=0.05)
    p00 = _cpu_scene(str(tmp_path), "p00", gen, phase=0.0)
    ref = orc.Scene(p05, str(tmp_path), nthreads=8)
    with _ctx(schedule=1) as a, _ctx(schedule=1) as b, _ctx(schedule=1) as c:
        _gen(a, "s", gen, phase=0.0)
        a.trace_frame()
        a.remesh(0.05)
        b.load_scene(p05, str(tmp_path))
        _same_scene(a, b)
        for d in range(a.num_domains):
            assert np.array_equal(a.normals(d), ref.mesh(d)[3])
        ia, sa_ = a.trace_frame()
        ib, sb = b.trace_frame()
        assert sa_["useful_rays"] == sb["useful_rays"] and sa_["final_hits"] == sb["final_hits"]
        err = np.abs(ia[..., :3] - ib[..., :3])
        assert float(np.sqrt((err ** 2).mean())) <= 1e-6
        a.remesh(0.0)
        c.load_scene(p00, str(tmp_path))
        _same_scene(a, c)


def test_remesh_needs_generated_scene(tmp_path):
    import spray_amd as sa
    path = _cpu_scene(str(tmp_path), "s", GEN[0])
    with _ctx() as ctx:
        ctx.load_scene(path, str(tmp_path))
        with pytest.raises(sa.SprayError):
            ctx.remesh(0.1)


def test_write_scene_round_trip(tmp_path):
    """A generated scene written by the GPU writer: PLY files byte-identical
    to the CPU generator's, and the written .spray loads to the same scene.
    A loaded scene with two lights, ASCII domains and a transform writes
    and reloads to the same meshes, lights and frame."""
    gen = GEN[0]
    cpu = tmp_path / "cpu"
    out = tmp_path / "out"
    cpu.mkdir()
    out.mkdir()
    _cpu_scene(str(cpu), "s", gen)
    with _ctx(schedule=1) as a, _ctx(schedule=1) as b:
        _gen(a, "s", gen)
        a.write_scene(str(out), "s")
        for d in range(a.num_domains):
            name = f"s_d{d:04d}.ply"
            assert (out / name).read_bytes() == (cpu / name).read_bytes(), name
        b.load_scene(str(out / "s.spray"), str(out))
        _same_scene(a, b)
    src = os.path.join(GOLD, "wavelet_example.spray")
    cam = dict(camera_pos=[0.0, 0.0, 40.0], camera_lookat=[0.0, 0.0, 0.0], camera_up=[0.0, 1.0, 0.0])
    with _ctx(**cam) as a, _ctx(**cam) as b:
        a.load_scene(src, GOLD)
        a.write_scene(str(out), "w")
        b.load_scene(str(out / "w.spray"), str(out))
        _same_scene(a, b)
        ia, sa_ = a.trace_frame()
        ib, sb = b.trace_frame()
    assert sa_["useful_rays"] == sb["useful_rays"] and sa_["final_hits"] == sb["final_hits"]
    assert float(np.abs(ia - ib).max()) <= 1e-4
