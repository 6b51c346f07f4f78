"""GPU parity of the volume integrator against the CPU oracle.

Bars:
  * grid build (NanoVDB-layout lookups, brick ranges / atlas / mips): bit-exact.
  * per-path radiance, same seeds, GPU FP32 vs oracle FP32: >= 99% of paths agree
    to 1e-3 relative (the rest diverge through last-ulp differences of logf /
    sinf / cosf / acosf / atan2f between the ROCm device library and glibc).
  * frames: relMSE(GPU, oracle FP32, same seeds) small; convergence: relMSE of the
    GPU render against the FP64 oracle at independent seeds <= 1e-3 at high spp.
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


@pytest.fixture(scope="module")
def c1_small():
    return scenes.make_workload("c1", width=64, height=64, spp=4)


@pytest.fixture(scope="module")
def c2_small():
    return scenes.make_workload("c2", n=64, width=64, height=64, spp=4, max_depth=16)


def _upload(ctx, w):
    grids = [ctx.upload_grid(g) for g in w["grids"]]
    env = ctx.upload_env(w["env"])
    return grids, env


def test_grid_build_matches_oracle(ctx, c2_small):
    w = c2_small
    grids, env = _upload(ctx, w)
    g = w["grids"][0]
    rec = g["records"]
    dens = rec[..., :4].copy().view(np.float32)[..., 0]
    bg = O.Brickgrid(dens)
    rng, u8, mips = grids[0].download_bricks()
    assert tuple(grids[0].info()["n_bricks"]) == bg.nb
    assert np.array_equal(rng, bg.range)
    # atlas values only matter in non-empty bricks
    nbx, nby, nbz = bg.nb
    ne = bg.nonempty.reshape(nbz, nby, nbx).astype(bool)
    mask = np.repeat(np.repeat(np.repeat(ne, 8, 0), 8, 1), 8, 2).reshape(-1)
    assert np.array_equal(u8[mask], bg.u8[mask])
    for l in range(3):
        assert np.array_equal(mips[l], bg.mips[l])
    sc = O.Scene(w["grids"], w["instances"], w["env"], "f32")
    oi = sc.info()
    gi = grids[0].info()
    for k in ("index_to_model", "model_to_index", "model_to_scaled", "scaled_to_model"):
        assert np.array_equal(np.asarray(gi[k], np.float32), oi[k]), k
    assert np.float32(gi["inv_max_density"]) == oi["inv_max_density"]


def test_voxel_lookup_bit_exact(ctx, c2_small):
    w = c2_small
    grids, _ = _upload(ctx, w)
    rec = w["grids"][0]["records"]
    n = rec.shape[0]
    rng = np.random.default_rng(3)
    coords = rng.integers(-3, n + 3, size=(20000, 3)).astype(np.int32)
    coords[:8] = [[0, 0, 0], [n - 1, n - 1, n - 1], [-1, 0, 0], [n, 0, 0], [0, n - 1, 0], [0, 0, n - 1], [5, 5, 5],
                  [n - 1, 0, n - 1]]
    got = grids[0].lookup(coords)
    dens = rec[..., :4].copy().view(np.float32)[..., 0]
    bg_rec = np.zeros(32, np.uint8)
    bg_rec[28:30] = np.frombuffer(np.uint16(0x3C00).tobytes(), np.uint8)
    for (x, y, z), r in zip(coords, got):
        inside = 0 <= x < n and 0 <= y < n and 0 <= z < n
        corner = x in (0, n - 1) and y in (0, n - 1) and z in (0, n - 1)
        if inside and (dens[z, y, x] > 0 or corner):
            assert np.array_equal(r, rec[z, y, x]), (x, y, z)
        else:
            assert np.array_equal(r, bg_rec), (x, y, z)


@pytest.mark.parametrize("wl,use_brick,bar,bar_cnt", [("c1_small", True, 0.99, 0.99), ("c1_small", False, 0.99, 0.99),
                                                      ("c2_small", True, 0.94, 0.965)])
def test_paths_match_oracle_fp32(ctx, wl, use_brick, bar, bar_cnt, request):
    """Per-path agreement at the same seeds: a GPU path follows the FP32 oracle's
    until a decision falls within rounding of its threshold.  The render kernels
    fuse multiply-adds and divide / take square roots within 2 ulp, while the
    oracle rounds every operation correctly, so such decisions occur.  The
    anisotropic scene under a 1e4:1 sun is the most sensitive: on these paths the
    FP64 oracle agrees with the FP32 oracle on ~94.1% of them, and the GPU on
    ~94.5% (lookup counters, the route a path took, on ~97.0%), hence its bars.
    The smooth scene keeps 99% bars for both.  A second bar is derived rather
    than set: the FP64 oracle's agreement with the FP32 oracle on the same paths
    measures how rounding-sensitive the scene is, and the GPU's agreement must
    stay within 3 points of it."""
    w = request.getfixturevalue(wl)
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
    ref64 = O.Scene(w["grids"], w["instances"], w["env"], "f64").render_paths(cfg, pix, smp)
    o_close = np.all(np.abs(ref64 - ref) <= 1e-3 * np.maximum(np.abs(ref), 1e-3), axis=1)
    print(f"{wl} brick={use_brick}: paths agreeing {close.mean():.5f}, counters agreeing {same_cnt.mean():.5f}; "
          f"FP64 oracle against FP32 oracle {o_close.mean():.5f}")
    assert close.mean() >= bar
    assert same_cnt.mean() >= bar_cnt
    assert close.mean() >= o_close.mean() - 0.03


@pytest.mark.parametrize("wl,tol", [("c1_small", 1e-5), ("c2_small", 2e-2)])
def test_frame_same_seed_vs_oracle_fp32(ctx, wl, tol, request):
    """Same seeds, same precision: the frame differs only through the few paths
    whose decisions flip on a rounding difference.  On c2_small the FP32 and FP64
    oracles differ by relMSE 1.3e-2 in this way, and the GPU by 1.24e-2 from the
    FP32 oracle; the bar is 2e-2."""
    w = request.getfixturevalue(wl)
    grids, env = _upload(ctx, w)
    cfg = w["render"]
    img = ctx.render(cfg, grids, w["instances"], env)
    ref = O.Scene(w["grids"], w["instances"], w["env"], "f32").render(cfg)
    e = relmse(img, ref)
    print(f"{wl}: relMSE(GPU FP32, oracle FP32, same seeds) = {e:.3e}")
    assert np.isfinite(img).all()
    assert e <= tol


@pytest.mark.parametrize("wl,size,spp_gpu,spp_ref", [("c1", 64, 1024, 1024), ("c2", 32, 8192, 4096)])
def test_frame_converges_to_oracle_fp64(ctx, wl, size, spp_gpu, spp_ref):
    """Unbiasedness: independent seeds, GPU FP32 at spp_gpu vs the FP64 oracle at
    spp_ref; relMSE = variance of both estimates, <= 1e-3 at these sample counts."""
    kw = dict(width=size, height=size, spp=spp_gpu)
    if wl == "c2":
        kw["n"] = 64
    w = scenes.make_workload(wl, **kw)
    grids, env = _upload(ctx, w)
    img = ctx.render(dict(w["render"], seed=101), grids, w["instances"], env)
    ref = O.Scene(w["grids"], w["instances"], w["env"], "f64").render(dict(w["render"], spp=spp_ref, seed=202))
    e = relmse(img, ref)
    print(f"{wl} {size}^2: relMSE(GPU {spp_gpu} spp, oracle FP64 {spp_ref} spp) = {e:.3e}")
    assert e <= 1e-3


def test_render_deterministic(ctx, c1_small):
    w = c1_small
    grids, env = _upload(ctx, w)
    a = ctx.render(w["render"], grids, w["instances"], env)
    b = ctx.render(w["render"], grids, w["instances"], env)
    assert np.array_equal(a, b)
