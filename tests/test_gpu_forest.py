"""GPU parity of scenes with many medium instances: 256 instances of two grids
on the forest ground (the instance hierarchy stands in for the reference's
OptiX instance acceleration structure, raytracer/src/scene/scene.cu:188-256).

* The scene intersection (r2_scene_intersect, built with correctly rounded
  arithmetic) equals the oracle's linear scan in instance order exactly: hit,
  t, instance and surface, for rays from the camera, rays inside the canopy and
  grazing rays along the ground.
* Paths: per-path agreement with the FP32 oracle at the same seeds.
* The render server's batched kernel (the hierarchy in the request's arena)
  equals r2_render bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from renderer2_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def forest():
    return scenes.forest_workload(256, width=48, height=48, spp=4)


def _upload(ctx, w):
    return [ctx.upload_grid(g) for g in w["grids"]], ctx.upload_env(w["env"])


def _oracle(w, precision="f32"):
    return O.Scene(w["grids"], w["instances"], w["env"], precision, surfaces=w["render"].get("surfaces", ()))


def _rays(n, seed=5):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.empty((n, 3))
    k = n // 3
    o[:k] = (0.0, 2.2, 6.5) + rng.normal(scale=0.3, size=(k, 3))          # around the camera
    o[k:2 * k] = rng.uniform((-4, 0.05, -4), (4, 1.5, 4), size=(k, 3))   # inside the canopy
    o[2 * k:] = rng.uniform((-5, 0.01, -5), (5, 0.2, 5), size=(n - 2 * k, 3))  # along the ground
    d[2 * k:, 1] *= 0.05  # grazing
    d[2 * k:] /= np.linalg.norm(d[2 * k:], axis=1, keepdims=True)
    r = np.concatenate([o, d, np.full((n, 1), 1e12)], axis=1).astype(np.float32)
    return r


@pytest.mark.parametrize("ground", [True, False])
def test_forest_intersections_match_oracle(ctx, ground):
    w = scenes.forest_workload(256, ground=ground)
    grids, env = _upload(ctx, w)
    rays = _rays(6000)
    got = ctx.scene_intersect(w["render"], grids, w["instances"], env, rays)
    sc = _oracle(w)
    exp = np.zeros((len(rays), 4), np.float32)
    for i, r in enumerate(rays):
        hit, out = sc.intersect(r.astype(np.float32))
        exp[i] = (hit, out[0] if hit else 0.0, out[4], out[5])
    hits = exp[:, 0] > 0
    print(f"ground={ground}: {hits.mean():.3f} of rays hit, {(exp[:, 2] >= 0).mean():.3f} hit a box")
    assert hits.mean() > 0.3 and (exp[:, 2] >= 0).mean() > 0.2
    assert np.array_equal(got[:, 0], exp[:, 0])
    assert np.array_equal(got[:, 2], exp[:, 2])
    assert np.array_equal(got[:, 3], exp[:, 3])
    assert np.array_equal(got[hits, 1], exp[hits, 1])


def test_forest_paths_match_oracle_fp32(ctx, forest):
    w = forest
    grids, env = _upload(ctx, w)
    cfg = w["render"]
    W, H = cfg["width"], cfg["height"]
    pix = np.repeat(np.arange(W * H, dtype=np.int32), 2)
    smp = np.tile(np.arange(2, dtype=np.int32), W * H)
    gpu, gcnt = ctx.render_paths(cfg, grids, w["instances"], env, pix, smp, counters=True)
    ref, rcnt = _oracle(w).render_paths(cfg, pix, smp, counters=True)
    close = np.all(np.abs(gpu - ref) <= 1e-3 * np.maximum(np.abs(ref), 1e-3), axis=1)
    same_cnt = np.all(gcnt.astype(np.int64) == rcnt.astype(np.int64), axis=1)
    print(f"forest 256: paths agreeing {close.mean():.5f}, counters agreeing {same_cnt.mean():.5f}")
    assert close.mean() >= 0.95
    assert same_cnt.mean() >= 0.95


def test_forest_server_matches_render(ctx, forest):
    w = forest
    grids, env = _upload(ctx, w)
    cfg = dict(w["render"], spp=8)
    ref = ctx.render(cfg, grids, w["instances"], env)
    srv = ctx.server(max_inflight=2, slice_paths=1 << 13)
    try:
        out = srv.wait(srv.submit(cfg, grids, w["instances"], env))[0]
    finally:
        srv.stop()
    assert np.array_equal(out, ref)
    assert np.isfinite(ref).all() and ref.mean() > 0
