"""GPU parity of the DomainStats schedule (ooc/ooc_domain_stats.cc:60-111:
score = sum over queued entries of max(0, 16 - position in the ray's domain
list), domains visited by descending score, ties by id).  A wrong score
leaves the image unchanged, so the frame tests cannot see it: here the plain
schedule's per-pass scores and visit order are read back through the C ABI.

The first pass of every window queues the same rays on both sides (a tile's
eye rays, or the pending rays restarted after validation), so its scores,
order and entry count must equal the oracle's exactly.  Later passes hold
different ray sets by design: the oracle, like the reference, traces rays
spawned in a pass by the domains still to come in that pass, while the GPU
queues every ray spawned in a pass for the next one.  Those passes are
checked for the ordering rule."""
import math
import os

import numpy as np
import pytest

import orc
from test_oracle_schedule import expected_order

pytestmark = pytest.mark.gpu

CAM = dict(pos=[2.0, 2.0, 2.0 + 2.0 * math.sqrt(3.0)], lookat=[2.0, 2.0, 2.0], up=[0.0, 1.0, 0.0])

CASES = {
    "pt": dict(image_w=96, image_h=96, pixel_samples=2, ao_mode=0, ao_samples=1, bounces=5, rr_start=3,
               max_samples_per_rank=4608),
    "ao": dict(image_w=128, image_h=128, pixel_samples=1, ao_mode=1, ao_samples=4, bounces=1,
               ao_max_dist=0.5, max_samples_per_rank=4096),
}


@pytest.fixture(scope="module")
def lattice(tmp_path_factory):
    import spray_amd as sa
    d = str(tmp_path_factory.mktemp("sched"))
    sa.scenegen_write(d, "g", 4, 4, 4, 16, rgb=0xB3B3B3, extra_lines="light rect 0.5 4.5 0.5 3 0 0 0 0 3 4 4 4")
    return d


@pytest.mark.parametrize("case", sorted(CASES))
def test_schedule_matches_oracle(lattice, case):
    import spray_amd as sa
    kw = CASES[case]
    d = lattice
    ref = orc.Scene(os.path.join(d, "g.spray"), d, nthreads=8)
    rimg, rst, rmeta, rscores, rorder = ref.render_schedule(
        512, nthreads=8, cam_pos=CAM["pos"], cam_lookat=CAM["lookat"], cam_up=CAM["up"], **kw)
    cfg = sa.default_config(has_camera=1, camera_pos=CAM["pos"], camera_lookat=CAM["lookat"],
                            camera_up=CAM["up"], nthreads=8, schedule=0, **kw)
    with sa.Context(cfg) as ctx:
        ctx.load_scene(os.path.join(d, "g.spray"), d)
        ctx.schedule_log(512)
        img, st = ctx.trace_frame()
        meta, scores, order = ctx.schedule_read()
    assert st["useful_rays"] == rst["useful_rays"] and st["final_hits"] == rst["final_hits"]
    assert len(meta) == st["passes"]
    for i in range(len(meta)):
        assert np.array_equal(order[i], expected_order(scores[i])), i
        assert scores[i].sum() >= meta[i][3]  # every entry weighs >= 1 (lists of <= 10 domains)
        assert scores[i].sum() <= 16 * meta[i][3]
    gfirst = {(int(m[0]), int(m[1])): i for i, m in enumerate(meta) if m[2] == 0}
    rfirst = {(int(m[0]), int(m[1])): i for i, m in enumerate(rmeta) if m[2] == 0}
    assert gfirst.keys() == rfirst.keys()
    assert len(gfirst) >= (8 if case == "pt" else 4)
    for key in sorted(gfirst):
        g, r = gfirst[key], rfirst[key]
        assert meta[g][3] == rmeta[r][3], key
        assert np.array_equal(scores[g], rscores[r]), key
        assert np.array_equal(order[g], rorder[r]), key
    print(f"{case}: {len(meta)} passes, {len(gfirst)} window starts equal to the oracle's "
          f"(largest score {int(scores.max())})")
