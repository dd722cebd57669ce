"""Pins the oracle's DomainStats schedule (ooc/ooc_domain_stats.cc:60-111,
deviations Q2 and Q12) on the CPU: the first pass of a tile queues exactly the
eye rays, so every domain's score is the sum over eye rays of max(0, 16 -
position of the domain in the ray's (t, id)-sorted domain list), computed
here from the domain lists alone; every pass visits domains by descending
score with ties broken by id."""
import os

import numpy as np

import orc

CAM = dict(cam_pos=[2.0, 2.3, 7.5], cam_lookat=[2.0, 2.0, 2.0], cam_up=[0.0, 1.0, 0.0])


def expected_order(scores):
    nd = len(scores)
    return np.array(sorted(range(nd), key=lambda d: (-int(scores[d]), d)), np.int32)


def first_pass_scores(ref, kw, tile_rows, tile_y):
    """Scores of a tile's eye rays (1 spp: through pixel corners)."""
    w = kw["image_w"]
    ys, xs = np.mgrid[tile_y:tile_y + tile_rows, 0:w]
    xs = xs.ravel().astype(np.float32)
    ys = ys.ravel().astype(np.float32)
    n = len(xs)
    dirs = np.zeros((n, 3), np.float32)
    orc.lib().orc_camera_dirs(orc.p(np.array(CAM["cam_pos"], np.float32)),
                              orc.p(np.array(CAM["cam_lookat"], np.float32)),
                              orc.p(np.array(CAM["cam_up"], np.float32)), 90.0, w, kw["image_h"], n,
                              orc.p(xs), orc.p(ys), orc.p(dirs))
    org = np.tile(np.array(CAM["cam_pos"], np.float32), (n, 1))
    cnt, ids, _ = ref.domain_list(org, dirs, maxk=64)
    sc = np.zeros(ref.num_domains, np.int64)
    for i in range(n):
        for k in range(min(cnt[i], 64)):
            sc[ids[i, k]] += max(0, 16 - k)
    return sc, int(cnt.sum())


def test_schedule_scores_and_order(tmp_path):
    import spray_amd as sa
    d = str(tmp_path)
    # a 6 x 2 x 2 row: eye rays cross up to 8 domains; some cross none
    sa.scenegen_write(d, "s", 6, 2, 2, 8, rgb=0xB3B3B3, extra_lines="light point 2 6 2 1 1 1")
    ref = orc.Scene(os.path.join(d, "s.spray"), d, nthreads=4)
    kw = dict(image_w=40, image_h=32, pixel_samples=1, ao_mode=0, ao_samples=1, bounces=5,
              max_samples_per_rank=40 * 8)
    img, st, meta, scores, order = ref.render_schedule(256, nthreads=4, **CAM, **kw)
    assert len(meta) > 8 and st["tiles"] == 4
    for i in range(len(meta)):
        assert np.array_equal(order[i], expected_order(scores[i])), i
        assert scores[i].sum() >= meta[i][3] or meta[i][3] == 0  # every entry weighs >= 1 here
    firsts = {int(meta[i][0]): i for i in range(len(meta)) if meta[i][1] == 0 and meta[i][2] == 0}
    checked = 0
    for tile in range(4):
        sc, entries = first_pass_scores(ref, kw, 8, 8 * tile)
        if entries == 0:  # a tile whose eye rays meet no domain schedules nothing
            assert tile not in firsts
            continue
        i = firsts[tile]
        assert meta[i][3] == entries
        assert np.array_equal(scores[i], sc), tile
        checked += 1
    assert checked >= 2
