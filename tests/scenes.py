"""Small scene writers for tests: ASCII PLY meshes and .spray scene files."""
import os

import numpy as np


def write_ply(path, verts, faces, rgb=None, binary=False):
    """Writes a triangle mesh; `rgb` is (n, 3) uint8 per vertex or None."""
    verts = np.asarray(verts, np.float32).reshape(-1, 3)
    faces = np.asarray(faces, np.int64).reshape(-1, 3)
    head = ["ply", "format binary_little_endian 1.0" if binary else "format ascii 1.0",
            f"element vertex {len(verts)}", "property float x", "property float y", "property float z"]
    if rgb is not None:
        head += ["property uchar red", "property uchar green", "property uchar blue"]
    head += [f"element face {len(faces)}", "property list uchar int vertex_indices", "end_header"]
    if binary:
        with open(path, "wb") as f:
            f.write(("\n".join(head) + "\n").encode())
            if rgb is None:
                f.write(verts.astype("<f4").tobytes())
            else:
                rec = np.zeros(len(verts), dtype=[("p", "<f4", 3), ("c", "u1", 3)])
                rec["p"] = verts
                rec["c"] = np.asarray(rgb, np.uint8).reshape(-1, 3)
                f.write(rec.tobytes())
            fr = np.zeros(len(faces), dtype=[("n", "u1"), ("i", "<i4", 3)])
            fr["n"] = 3
            fr["i"] = faces
            f.write(fr.tobytes())
        return
    with open(path, "w") as f:
        f.write("\n".join(head) + "\n")
        for i, v in enumerate(verts):
            line = f"{float(v[0])!r} {float(v[1])!r} {float(v[2])!r}"
            if rgb is not None:
                c = np.asarray(rgb).reshape(-1, 3)[i]
                line += f" {int(c[0])} {int(c[1])} {int(c[2])}"
            f.write(line + "\n")
        for t in faces:
            f.write(f"3 {int(t[0])} {int(t[1])} {int(t[2])}\n")


def write_scene(dirname, name, domains, lights):
    """domains: list of dicts with keys verts, faces, rgb, mtl (optional).
    Bounds are the mesh's own AABB.  Returns the .spray path."""
    lines = list(lights)
    for i, d in enumerate(domains):
        ply = f"{name}_d{i}.ply"
        write_ply(os.path.join(dirname, ply), d["verts"], d["faces"], d.get("rgb"))
        v = np.asarray(d["verts"], np.float32).reshape(-1, 3)
        lo, hi = v.min(0), v.max(0)
        lines += ["domain", f"file {ply}", "mtl " + d.get("mtl", "diffuse 1 1 1"),
                  "bound " + " ".join(repr(float(x)) for x in list(lo) + list(hi)),
                  f"vertex {len(v)}", f"face {len(np.asarray(d['faces']).reshape(-1, 3))}"]
    path = os.path.join(dirname, name + ".spray")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def quad(y, x0, x1, z0, z1, c=(255, 255, 255), up=True):
    """Axis-aligned rectangle in the plane y = const, two triangles, normal
    +y when `up` (counter-clockwise seen from above)."""
    v = [[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]]
    f = [[0, 2, 1], [0, 3, 2]] if up else [[0, 1, 2], [0, 2, 3]]
    return dict(verts=v, faces=f, rgb=[c] * 4)


def cube(lo, hi, c=(255, 255, 255)):
    """Closed axis-aligned box, 12 triangles, outward normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
         [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]]
    f = [[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
         [3, 6, 2], [3, 7, 6], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]]
    return dict(verts=v, faces=f, rgb=[c] * 8)
