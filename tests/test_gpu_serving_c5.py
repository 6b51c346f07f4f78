"""The render server at the serving workload's own configuration.

Four grids resident in HBM (the 256^3 anisotropic fibre field filtered to
128^3, 256^3 and 512^3, plus the 64^3 value-noise grid) under the procedural
sky + sun environment; a seeded stream of requests with resolution over
{256, 512, 1024}^2, spp over {16, 64, 256}, depth 16, an orbit camera and a grid
each.  All requests are submitted at once, so they are admitted by the HBM
budget and share slices of the batched kernel.

Bar: every request's frame equals a standalone r2_render of the same request bit
for bit (same work items, same chunk-ordered reduction, whatever else shares the
device), and the server's latency statistics cover every request.
"""
import numpy as np
import pytest

from renderer2_amd import scenes

pytestmark = pytest.mark.gpu

N_REQ = 12


@pytest.fixture(scope="module")
def world(ctx):
    grids = [ctx.upload_grid(g) for g in scenes.serving_grids()]
    env = ctx.upload_env(scenes.synth_env("sky", 512, 256))
    yield grids, env
    for g in grids:
        g.free()
    env.free()


def test_serving_mix_matches_standalone(ctx, world):
    grids, env = world
    reqs = scenes.serving_requests(N_REQ, None, 11)
    shapes = sorted({(c["width"], c["spp"], inst[0]["grid"]) for _, c, inst in reqs})
    print(f"{N_REQ} requests (size, spp, grid): {shapes}", flush=True)
    assert {c["width"] for _, c, _ in reqs} == set(scenes.SERVING_SIZES)
    assert {c["spp"] for _, c, _ in reqs} == set(scenes.SERVING_SPPS)
    srv = ctx.server(hbm_budget_bytes=16 << 30, max_inflight=16, slice_paths=1 << 24)
    try:
        ids = [srv.submit(cfg, grids, inst, env) for _, cfg, inst in reqs]
        outs = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.stop()
    assert st["completed"] == N_REQ and st["rejected"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    print(f"server: p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms, "
          f"{st['paths_per_sec'] / 1e6:.1f} M paths/s, peak {st['peak_bytes'] / 2**30:.2f} GiB", flush=True)
    for k, ((_, cfg, inst), (img, lat)) in enumerate(zip(reqs, outs)):
        alone = ctx.render(cfg, grids, inst, env)
        assert np.isfinite(alone).all() and alone.max() > 0
        assert np.array_equal(img, alone), f"request {k} ({cfg['width']}^2, {cfg['spp']} spp): differs"
