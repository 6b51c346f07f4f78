"""GPU parity of scenes with several medium instances (scene intersection over
instance AABBs, instance transforms, intersectTr walks across media) against
the CPU oracle.

The scene: two grids (a 32^3 anisotropic fibre volume and a 16^3 value-noise
volume) placed as four instances:
  * 0: fibre grid, scaled 0.6, shifted left;
  * 1: noise grid rotated by getRotation((1, 0, 0)), anisotropically scaled,
    overlapping instance 0's box;
  * 2: fibre grid rotated by getRotation((0, 0, -1)), scaled 0.4;
  * 3: noise grid rotated by getRotation((-1, 0, 0)), scaled 0.3, above the
    others.
under the procedural sky + sun environment, so that shadow rays toward the sun
cross several boxes.

Bars as for one instance: per-path agreement with the FP32 oracle at the same
seeds, and convergence to the FP64 oracle at independent seeds.
"""
import numpy as np
import pytest

import oracle as O
from renderer2_amd import scenes

pytestmark = pytest.mark.gpu


def relmse(x, r):
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def multi_instance_workload(width=48, height=48, spp=4, max_depth=8):
    fib = scenes.make_workload("c2", n=32, width=width, height=height, spp=spp, max_depth=max_depth)
    noise = scenes.make_workload("c1", n=16)
    grids = [fib["grids"][0], noise["grids"][0]]
    instances = [
        dict(grid=0, pos=(-0.35, 0.0, 0.0), orientation=(0, 1, 0), scaling=(0.6, 0.6, 0.6)),
        dict(grid=1, pos=(0.05, 0.1, 0.05), orientation=(1, 0, 0), scaling=(0.5, 0.7, 0.4)),
        dict(grid=0, pos=(0.35, -0.35, 0.2), orientation=(0, 0, -1), scaling=(0.4, 0.4, 0.4)),
        dict(grid=1, pos=(0.0, 0.45, -0.3), orientation=(-1, 0, 0), scaling=(0.3, 0.3, 0.3)),
    ]
    render = dict(fib["render"], cam_pos=(0.3, 0.4, 2.4), cam_look_at=(0.0, 0.0, 0.0))
    return dict(grids=grids, instances=instances, env=scenes.synth_env("sky", 64, 32), render=render)


@pytest.fixture(scope="module")
def multi():
    return multi_instance_workload()


def _upload(ctx, w):
    return [ctx.upload_grid(g) for g in w["grids"]], ctx.upload_env(w["env"])


@pytest.mark.parametrize("use_brick", [True, False])
def test_multi_instance_paths_match_oracle_fp32(ctx, multi, use_brick):
    w = multi
    grids, env = _upload(ctx, w)
    cfg = dict(w["render"], use_brick_grid=use_brick)
    W, H = cfg["width"], cfg["height"]
    pix = np.repeat(np.arange(W * H, dtype=np.int32), 2)
    smp = np.tile(np.arange(2, dtype=np.int32), W * H)
    gpu, gcnt = ctx.render_paths(cfg, grids, w["instances"], env, pix, smp, counters=True)
    sc = O.Scene(w["grids"], w["instances"], w["env"], "f32")
    ref, rcnt = sc.render_paths(cfg, pix, smp, counters=True)
    close = np.all(np.abs(gpu - ref) <= 1e-3 * np.maximum(np.abs(ref), 1e-3), axis=1)
    same_cnt = np.all(gcnt.astype(np.int64) == rcnt.astype(np.int64), axis=1)
    print(f"4 instances brick={use_brick}: paths agreeing {close.mean():.5f}, counters agreeing {same_cnt.mean():.5f}")
    assert close.mean() >= 0.94
    assert same_cnt.mean() >= 0.95


def _blocks(x, W, H, k=4):
    """Mean over k x k pixel blocks: keeps a bias, divides independent noise by k^2."""
    x = np.asarray(x, np.float64).reshape(H, W, 3)
    return x.reshape(H // k, k, W // k, k, 3).mean(axis=(1, 3))


def test_multi_instance_frame_converges_to_oracle_fp64(ctx):
    """The sun (1e4:1) makes single pixels noisy: two independent FP64 oracle
    frames of this scene at 2048 spp differ by relMSE 2.0e-2.  At 16384 GPU spp
    against 8192 oracle spp the noise floor measured with the FP32 oracle in
    place of the GPU is 3.5e-3 per pixel and 2.3e-4 to 3.1e-4 on 4x4 block means
    (three seeds), so a bias of the multi-instance path above 1e-3 fails."""
    W = H = 32
    w = multi_instance_workload(width=W, height=H, spp=16384)
    grids, env = _upload(ctx, w)
    img = ctx.render(dict(w["render"], seed=101), grids, w["instances"], env)
    ref = O.Scene(w["grids"], w["instances"], w["env"], "f64").render(dict(w["render"], spp=8192, seed=202))
    e = relmse(img, ref)
    eb = relmse(_blocks(img, W, H), _blocks(ref, W, H))
    print(f"4 instances 32^2: relMSE(GPU 16384 spp, oracle FP64 8192 spp) = {e:.3e} per pixel, {eb:.3e} on 4x4 blocks")
    assert e <= 1e-2
    assert eb <= 1e-3


def test_multi_instance_server_matches_standalone(ctx, multi):
    """The batched kernel's multi-instance form against the standalone render, bit for bit."""
    w = multi
    grids, env = _upload(ctx, w)
    cfg = dict(w["render"], spp=16)
    want = ctx.render(cfg, grids, w["instances"], env)
    srv = ctx.server(hbm_budget_bytes=1 << 30, max_inflight=4, slice_paths=1 << 14)
    ids = [srv.submit(cfg, grids, w["instances"], env) for _ in range(2)]
    outs = [srv.wait(i)[0] for i in ids]
    srv.stop()
    for o in outs:
        assert np.array_equal(o, want)
