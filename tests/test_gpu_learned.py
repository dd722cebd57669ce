"""GPU tests of the domain-hit predictor (MFMA forward/backward/Adam) and of
the learned schedule."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_ref  # noqa: E402
import orc  # noqa: E402

pytestmark = pytest.mark.gpu

CAM = dict(pos=[2.0, 2.0, 2.0 + 2.0 * math.sqrt(3.0)], lookat=[2.0, 2.0, 2.0], up=[0.0, 1.0, 0.0])
RECT_LIGHT = "light rect 0.5 4.5 0.5 3 0 0 0 0 3 4 4 4"


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    import spray_amd as sa
    d = str(tmp_path_factory.mktemp("learned"))
    sa.scenegen_write(d, "pt", 4, 4, 4, 32, rgb=0xB3B3B3, extra_lines=RECT_LIGHT)
    return d


def make_ctx(scene, schedule=0, size=128, spp=4):
    import spray_amd as sa
    cfg = sa.default_config(image_w=size, image_h=size, pixel_samples=spp, ao_mode=0, ao_samples=1,
                            bounces=5, rr_start=3, has_camera=1, camera_pos=CAM["pos"],
                            camera_lookat=CAM["lookat"], camera_up=CAM["up"], nthreads=8,
                            schedule=schedule, max_samples_per_rank=32768)
    ctx = sa.Context(cfg)
    ctx.load_scene(os.path.join(scene, "pt.spray"), scene)
    return ctx


def random_samples(rng, n, nd):
    s = np.zeros((n, 8), np.float32)
    s[:, 0:3] = rng.uniform(0.0, 4.0, (n, 3))
    d = rng.normal(size=(n, 3))
    s[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 6] = rng.integers(0, nd, n)
    s[:, 7] = rng.integers(0, 2, n)
    return s


def norm_of(scene_dir):
    ref = orc.Scene(os.path.join(scene_dir, "pt.spray"), scene_dir, parse_only=True)
    lo, hi = ref.bound()
    center = np.float32(0.5) * (lo + hi)
    half = np.float32(0.5) * (hi - lo)
    inv = (np.float32(1.0) / half).astype(np.float32)
    return center.astype(np.float64), inv.astype(np.float64)


def test_mlp_gradients_match_oracle(scene):
    ctx = make_ctx(scene)
    nd = ctx.num_domains
    P = ctx.mlp_num_params()
    assert P == 19585 + 16 * nd
    rng = np.random.default_rng(5)
    params = rng.normal(0.0, 0.15, P).astype(np.float32)
    samples = random_samples(rng, 256, nd)
    loss, g = ctx.mlp_grad(samples, params)
    center, inv = norm_of(scene)
    rloss, rg, rz = mlp_ref.forward_backward(params, samples, nd, center, inv)
    assert abs(loss - rloss) <= 1e-4 * abs(rloss) + 1e-4
    blocks = [("W1", 0, 128 * 22), ("b1", 2816, 128), ("W2", 2944, 16384), ("b2", 19328, 128),
              ("w3", 19456, 128), ("b3", 19584, 1), ("emb", 19585, 16 * nd)]
    for name, o, n in blocks:
        a, b = g[o:o + n].astype(np.float64), rg[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"{name}: rel err {rel:.2e}")
        assert rel < 5e-3, name
    z = ctx.mlp_logits(samples, params)
    assert np.max(np.abs(z - rz)) <= 1e-3 * (1 + np.max(np.abs(rz)))
    ctx.close()


def test_learned_schedule_same_image_fewer_visits(scene, tmp_path):
    """Train on hit logs of plain frames, then render with the learned
    schedule: identical image and useful-ray count, fewer radiance visits
    than the batched schedule without the predictor."""
    ctx = make_ctx(scene, schedule=0)
    plain_img, plain = ctx.trace_frame()
    ctx.predictor_log(True, 4_000_000)
    ctx.trace_frame(want_image=False)
    ctx.predictor_log(False)
    n = ctx.predictor_num_samples()
    assert n > 10_000
    loss0 = ctx.predictor_train(1, 1024, 2e-3)
    loss1 = ctx.predictor_train(3, 1024, 2e-3)
    vloss, vacc = ctx.predictor_eval()
    print(f"samples {n} train loss {loss0:.4f} -> {loss1:.4f} held-out loss {vloss:.4f} acc {vacc:.3f}")
    assert loss1 < 0.69 and vacc > 0.6
    ctx.set_schedule(1)
    bimg, batched = ctx.trace_frame()
    ctx.set_schedule(2)
    limg, learned = ctx.trace_frame()
    for name, img, st in (("batched", bimg, batched), ("learned", limg, learned)):
        diff = np.abs(img[..., :3] - plain_img[..., :3])
        print(f"{name}: rmse {np.sqrt((diff ** 2).mean()):.3g} max {diff.max():.3g} "
              f"radiance visits {st['radiance_visits']} deferred {st['deferred_entries']}")
        assert st["useful_rays"] == plain["useful_rays"]
        assert np.sqrt((diff ** 2).mean()) <= 1e-4 and diff.max() <= 1e-3
    assert learned["deferred_entries"] > 0
    assert learned["radiance_visits"] < batched["radiance_visits"]
    # and against the oracle on the same 64-domain scene and rays
    ref = orc.Scene(os.path.join(scene, "pt.spray"), scene, nthreads=8)
    rimg, rst = ref.render(image_w=128, image_h=128, pixel_samples=4, ao_mode=0, ao_samples=1, bounces=5,
                           rr_start=3, max_samples_per_rank=32768, nthreads=8, cam_pos=CAM["pos"],
                           cam_lookat=CAM["lookat"], cam_up=CAM["up"])
    diff = np.abs(limg[..., :3] - rimg[..., :3])
    print(f"learned vs oracle: rmse {np.sqrt((diff ** 2).mean()):.3g} max {diff.max():.3g}")
    assert learned["useful_rays"] == rst["useful_rays"] and learned["final_hits"] == rst["final_hits"]
    assert np.sqrt((diff ** 2).mean()) <= 1e-4 and diff.max() <= 1e-3
    # save / load round trip keeps the network
    path = str(tmp_path / "pred.bin")
    ctx.predictor_save(path)
    ctx.predictor_load(path)
    limg2, _ = ctx.trace_frame()
    assert np.abs(limg2 - limg).max() < 1e-5
    ctx.close()
