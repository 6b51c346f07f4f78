"""Majorant-DDA cell walks on the GPU, bit-exact against the oracle (FP32).

The walk is the free-flight loop of sampleDDA without its density decisions
(r2_dda_cells / the oracle's r2o_dda_cells): position, stepDDA, the t update and
the mip schedule, with the mip drop of a tentative collision forced every k steps
so that every mip level is walked.  Rays: random, axis-parallel, negative and
near-zero direction components, origins on cell faces.
"""
import numpy as np
import pytest

import oracle as O
from renderer2_amd import scenes

pytestmark = pytest.mark.gpu


def _rays():
    rng = np.random.default_rng(23)
    n = 2000
    d = rng.normal(size=(n, 3))
    d[:200, 1:] = 0.0                      # axis-parallel (+x / -x)
    d[200:300, 0] = 1e-7                   # a near-zero component
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = rng.uniform(0, 256, (n, 3))
    o[300:400] = np.floor(o[300:400] / 64) * 64   # origins on mip-3 cell faces
    o[400:500] = np.floor(o[400:500] / 8) * 8     # and on brick faces
    tmax = rng.uniform(1, 600, (n, 1))
    return np.concatenate([o, d, tmax], 1).astype(np.float32)


@pytest.mark.parametrize("collide_every", [0, 1, 3])
def test_dda_walk_bit_exact(ctx, collide_every):
    rays = _rays()
    got = ctx.dda_cells(rays, sigma_t=1.0, collide_every=collide_every, max_cells=2048)
    w = scenes.make_workload("c1", n=16)
    sc = O.Scene(w["grids"], w["instances"], w["env"], "f32")
    mips = set()
    for r, g in zip(rays, got):
        ref = sc.dda_cells(r, sigma_t=1.0, max_cells=2048, collide_every=collide_every)
        assert np.array_equal(g, ref), (r, len(g), len(ref))
        mips.update(int(m) for m in ref[:, 3])
    # a drop every step walks mips 3 -> 1 -> 0; every third step also reaches 2
    expect = {0: {3}, 1: {0, 1, 3}, 3: {0, 1, 2, 3}}[collide_every]
    assert mips == expect


def test_dda_walk_sigma_t(ctx):
    """t advances by dt / sigma_t, as in the free-flight loop."""
    rays = _rays()[:300]
    got = ctx.dda_cells(rays, sigma_t=3.5, collide_every=2, max_cells=4096)
    w = scenes.make_workload("c1", n=16)
    sc = O.Scene(w["grids"], w["instances"], w["env"], "f32")
    for r, g in zip(rays, got):
        assert np.array_equal(g, sc.dda_cells(r, sigma_t=3.5, max_cells=4096, collide_every=2))
