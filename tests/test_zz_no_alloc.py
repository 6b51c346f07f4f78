"""r2_render_noalloc (the render entry point that allocates no device memory)
against r2_render, through the C ABI.

Two calls in a row, the second with a smaller frame: each must equal r2_render's
frame bit for bit (the chunk sums of these frames fit the 64 MiB array, so the
two entry points run the same items and reduce them in the same order).  A third
frame too large for the default chunking runs with longer items and must agree
to FP32 summation order.  Oversized frames and host output pointers are refused
on the host before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from renderer2_amd import _capi as capi
from renderer2_amd import _hip, scenes

pytestmark = pytest.mark.gpu

R2_ERR_INVALID, R2_ERR_UNSUPPORTED = -1, -4


def _noalloc(ctx, cfg, grids, inst, env):
    W, H = cfg["width"], cfg["height"]
    p, keep = ctx.render_params(cfg, grids, inst, env)
    buf = _hip.DeviceBuffer(W * H * 3 * 4)
    try:
        rc = capi.lib().r2_render_noalloc(ctx.h, C.byref(p), buf.addr)
        assert rc == 0, capi.lib().r2_last_error(ctx.h)
        return buf.to_host(np.float32, (H, W, 3))
    finally:
        buf.free()


@pytest.fixture(scope="module")
def scene(ctx):
    w = scenes.make_workload("c2", n=48)
    return w, [ctx.upload_grid(g) for g in w["grids"]], ctx.upload_env(w["env"])


def test_noalloc_matches_render_twice(ctx, scene):
    w, grids, env = scene
    for W, H, spp, seed in ((96, 80, 24, 7), (40, 24, 5, 11)):
        cfg = dict(w["render"], width=W, height=H, spp=spp, seed=seed)
        print(f"noalloc {W}x{H} {spp} spp seed {seed}", flush=True)
        want = ctx.render(cfg, grids, w["instances"], env)
        got = _noalloc(ctx, cfg, grids, w["instances"], env)
        assert np.isfinite(want).all() and want.max() > 0
        assert np.array_equal(got, want)


def test_noalloc_fp64_accumulation(ctx, scene):
    w, grids, env = scene
    cfg = dict(w["render"], width=64, height=48, spp=16, seed=3, fp64_accum=True)
    print("noalloc fp64 64x48 16 spp", flush=True)
    want = ctx.render(cfg, grids, w["instances"], env)
    got = _noalloc(ctx, cfg, grids, w["instances"], env)
    assert np.array_equal(got, want)


def test_noalloc_longer_items_when_chunk_sums_exceed_scratch(ctx, scene):
    # 512^2 x 64 spp: r2_render's chunk sums take 201 MB, the scratch 64 MiB, so
    # items run 4 samples each instead of 1: same paths, other summation order
    w, grids, env = scene
    cfg = dict(w["render"], width=512, height=512, spp=64, seed=5)
    print("noalloc 512x512 64 spp (longer items)", flush=True)
    want = ctx.render(cfg, grids, w["instances"], env)
    got = _noalloc(ctx, cfg, grids, w["instances"], env)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-7)


def test_noalloc_refuses_oversized_frames_and_host_output(ctx, scene):
    w, grids, env = scene
    big = dict(w["render"], width=2400, height=2400, spp=1)  # 5.76 M pixels > 5592405
    p, keep = ctx.render_params(big, grids, w["instances"], env)
    buf = _hip.DeviceBuffer(16)
    try:
        assert capi.lib().r2_render_noalloc(ctx.h, C.byref(p), buf.addr) == R2_ERR_UNSUPPORTED
        assert b"scratch holds" in capi.lib().r2_last_error(ctx.h)
    finally:
        buf.free()
    small = dict(w["render"], width=8, height=8, spp=1)
    p, keep = ctx.render_params(small, grids, w["instances"], env)
    host = np.zeros((8, 8, 3), np.float32)
    assert capi.lib().r2_render_noalloc(ctx.h, C.byref(p), host.ctypes.data) == R2_ERR_INVALID
