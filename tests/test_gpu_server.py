"""Render server: continuous batching must not change results.

Bar: every request's frame from the server equals the frame of a standalone
r2_render with the same parameters bit for bit (same work items, same
chunk-ordered reduction), whatever else is resident; admission respects the HBM
budget; cancellation and oversize rejection behave as documented.
"""
import threading
import time

import numpy as np
import pytest

import renderer2_amd as r2
from renderer2_amd import capi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(ctx):
    w1 = scenes.make_workload("c1", n=48)
    w2 = scenes.make_workload("c2", n=64)
    grids = [ctx.upload_grid(w1["grids"][0]), ctx.upload_grid(w2["grids"][0])]
    env = ctx.upload_env(w2["env"])
    return dict(grids=grids, env=env, base1=w1["render"], base2=w2["render"])


def _requests(world):
    b1, b2 = world["base1"], world["base2"]
    reqs = []
    shapes = [(64, 64, 8), (96, 64, 4), (128, 128, 16), (32, 32, 64), (80, 48, 12), (128, 96, 6), (64, 128, 3),
              (48, 48, 33)]
    for k, (W, H, spp) in enumerate(shapes):
        base = b1 if k % 2 == 0 else b2
        ang = 0.4 * k
        cam = (2.3 * np.sin(ang), 0.3 * (k % 3 - 1), 2.3 * np.cos(ang))
        cfg = dict(base, width=W, height=H, spp=spp, seed=100 + k, cam_pos=cam, max_depth=4 + k % 5,
                   fp64_accum=(k == 5))
        inst = [dict(grid=k % 2)]
        if k == 7:  # two instances of different grids
            inst = [dict(grid=0, pos=(-0.3, 0, 0), scaling=(0.5, 0.5, 0.5)),
                    dict(grid=1, pos=(0.3, 0, 0), scaling=(0.5, 0.5, 0.5))]
        reqs.append((cfg, inst))
    return reqs


def test_server_matches_standalone_renders(ctx, world):
    reqs = _requests(world)
    alone = [ctx.render(cfg, world["grids"], inst, world["env"]) for cfg, inst in reqs]
    srv = ctx.server(hbm_budget_bytes=1 << 30, max_inflight=8, slice_paths=1 << 16)
    try:
        ids = [srv.submit(cfg, world["grids"], inst, world["env"]) for cfg, inst in reqs[:5]]
        time.sleep(0.01)  # the rest arrives while the first ones are on the GPU
        ids += [srv.submit(cfg, world["grids"], inst, world["env"]) for cfg, inst in reqs[5:]]
        for k, rid in enumerate(ids):
            img, lat = srv.wait(rid)
            assert lat > 0
            assert np.array_equal(img, alone[k]), f"request {k}: max diff {np.abs(img - alone[k]).max()}"
        with pytest.raises(r2.R2Error) as e:  # an outcome is collected once, then the id is forgotten
            srv.wait(ids[0])
        assert e.value.code == -1
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["rejected"] == 0
        assert 0 < st["p50_ms"] <= st["p99_ms"]
        assert st["paths_per_sec"] > 0
    finally:
        srv.stop()


def test_server_budget_admission_and_rejection(ctx, world):
    reqs = _requests(world)
    cfg, inst = reqs[2]  # 128x128 x 16 spp
    npix, spp = 128 * 128, 16
    chunk = min(spp, max(1, -(-npix * spp // (1 << 24))))  # samples per work item
    nchunks = -(-spp // chunk)
    # scene copy (< 64 KiB) + chunk sums + frame: room for exactly one request at a time
    one = (1 << 16) + nchunks * npix * 12 + npix * 12
    srv = ctx.server(hbm_budget_bytes=one, max_inflight=4, slice_paths=1 << 15)
    try:
        ref = ctx.render(cfg, world["grids"], inst, world["env"])
        ids = [srv.submit(dict(cfg), world["grids"], inst, world["env"]) for _ in range(4)]
        for rid in ids:
            img, _ = srv.wait(rid)
            assert np.array_equal(img, ref)
        st = srv.stats()
        assert 0 < st["peak_bytes"] <= one
        big = dict(cfg, width=1024, height=1024, spp=64)
        with pytest.raises(r2.R2Error) as e:
            srv.submit(big, world["grids"], inst, world["env"])
        assert e.value.code == -3  # R2_ERR_OOM: larger than the whole budget
        assert srv.stats()["rejected"] == 1
    finally:
        srv.stop()


def test_server_cancel(ctx, world):
    cfg, inst = _requests(world)[2]
    srv = ctx.server(max_inflight=1, slice_paths=1 << 14)
    try:
        a = srv.submit(dict(cfg, spp=256), world["grids"], inst, world["env"])
        b = srv.submit(dict(cfg, spp=8), world["grids"], inst, world["env"])  # queued behind a
        srv.cancel(b)
        with pytest.raises(r2.R2Error) as e:
            srv.wait(b)
        assert e.value.code == -6  # R2_ERR_CANCELLED
        img, _ = srv.wait(a)
        assert np.isfinite(img).all()
        assert srv.stats()["cancelled"] == 1
    finally:
        srv.stop()


def test_server_concurrent_submitters(ctx, world):
    """Submissions from several host threads (the server API is internally synchronised)."""
    reqs = _requests(world)[:6]
    alone = [ctx.render(cfg, world["grids"], inst, world["env"]) for cfg, inst in reqs]
    srv = ctx.server(max_inflight=3, slice_paths=1 << 16)
    out = [None] * len(reqs)

    def worker(k):
        rid = srv.submit(reqs[k][0], world["grids"], reqs[k][1], world["env"])
        out[k] = srv.wait(rid)[0]

    try:
        th = [threading.Thread(target=worker, args=(k,)) for k in range(len(reqs))]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        for k in range(len(reqs)):
            assert out[k] is not None and np.array_equal(out[k], alone[k])
    finally:
        srv.stop()


def test_server_deadlines(ctx, world):
    """Earliest-deadline-first admission, and requests dropped at their deadline
    (R2_ERR_DEADLINE) whether still queued or already resident; requests that make
    their deadline equal their standalone renders."""
    cfg, inst = _requests(world)[2]  # 128x128
    g, env = world["grids"], world["env"]
    cb, cc = dict(cfg, spp=64, seed=11), dict(cfg, spp=64, seed=12)
    ref_b, ref_c = ctx.render(cb, g, inst, env), ctx.render(cc, g, inst, env)
    srv = ctx.server(max_inflight=1, slice_paths=1 << 14)
    try:
        with pytest.raises(r2.R2Error) as e:
            srv.submit(dict(cfg, spp=8), g, inst, env, deadline_ms=-1.0)
        assert e.value.code == -1  # R2_ERR_INVALID
        a = srv.submit(dict(cfg, spp=256), g, inst, env)  # resident first, no deadline
        # b and c must queue behind a: wait until the scheduler has admitted a
        # (else b, having a deadline, would be admitted before a and ahead of c)
        import time
        t_end = time.monotonic() + 10.0
        while srv.stats()["peak_bytes"] == 0 and time.monotonic() < t_end:
            time.sleep(1e-4)
        assert srv.stats()["peak_bytes"] > 0
        b = srv.submit(cb, g, inst, env, deadline_ms=60000.0)
        c = srv.submit(cc, g, inst, env, deadline_ms=30000.0)  # earlier deadline: admitted before b
        d = srv.submit(dict(cfg, spp=8), g, inst, env, deadline_ms=1e-3)  # expires while queued
        with pytest.raises(r2.R2Error) as e:
            srv.wait(d)
        assert e.value.code == -8  # R2_ERR_DEADLINE
        srv.wait(a)
        img_c, lat_c = srv.wait(c)
        img_b, lat_b = srv.wait(b)
        assert lat_c < lat_b, (lat_c, lat_b)
        assert np.array_equal(img_b, ref_b) and np.array_equal(img_c, ref_c)
        # a long request alone on the device, with a deadline far shorter than its run
        f = srv.submit(dict(cfg, width=512, height=512, spp=1024), g, inst, env, deadline_ms=20.0)
        with pytest.raises(r2.R2Error) as e:
            srv.wait(f)
        assert e.value.code == -8
        st = srv.stats()
        assert st["expired"] == 2 and st["completed"] == 3 and st["cancelled"] == 0
        # the server keeps serving after drops
        img, _ = srv.wait(srv.submit(cb, g, inst, env, deadline_ms=60000.0))
        assert np.array_equal(img, ref_b)
    finally:
        srv.stop()
