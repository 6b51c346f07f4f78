"""Reduced-precision form (R2_PREC_PACKED) and compensated accumulation.

Bars:
  * FP32 grid sampling equals the oracle's filteredData restatement.
  * Packed records against the FP32 records at the same positions: density to
    FP16 (relative 1e-3), S entries within 4e-3 of the largest entry (FP16 plus the
    PSD floor), colours exact (u8 both), normals within 1.6 degrees at voxel
    centres (octahedral 8+8) and within 3 degrees for 99% of interpolated
    positions, ior within 1e-3.
  * Packed S is positive definite everywhere (FP64 Jacobi projection).
  * Packed renders converge to the FP64 oracle: relMSE <= 1e-3, the same bar as
    the FP32 form; FP64 compensated accumulation changes the FP32 frame by no
    more than float rounding.
"""
import numpy as np
import pytest

import oracle as O
import renderer2_amd as r2
from renderer2_amd import scenes

pytestmark = pytest.mark.gpu


def relmse(x, r):
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


@pytest.fixture(scope="module")
def c2w():
    return scenes.make_workload("c2", n=48, width=64, height=64, spp=8)


def _positions(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.02, 0.98, (n, 3)).astype(np.float32)


def test_fp32_sample_matches_oracle(ctx, c2w):
    g = ctx.upload_grid(c2w["grids"][0])
    pos = _positions(400)
    got = g.sample(pos)
    sc = O.Scene(c2w["grids"], c2w["instances"], c2w["env"], "f32")
    ref = np.array([sc.filtered_data(p) for p in pos], np.float32)
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-6)


def test_packed_close_to_fp32(ctx, c2w):
    gf = ctx.upload_grid(c2w["grids"][0])
    gp = ctx.upload_grid(dict(c2w["grids"][0], precision="packed"))
    pos = _positions(4000, seed=1)
    a, b = gf.sample(pos), gp.sample(pos)
    dmax = np.abs(a[:, 0]).max()
    assert np.all(np.abs(b[:, 0] - a[:, 0]) <= 1e-3 * np.abs(a[:, 0]) + 1e-3 * dmax)
    smax = np.abs(a[:, 1:7]).max()
    assert np.abs(b[:, 1:7] - a[:, 1:7]).max() <= 4e-3 * smax
    np.testing.assert_allclose(b[:, 7:13], a[:, 7:13], atol=1e-6)
    ang = np.degrees(np.arccos(np.clip(np.sum(a[:, 13:16] * b[:, 13:16], axis=1), -1, 1)))
    # between voxels the renormalised interpolation amplifies the encoding error
    # where neighbouring normals nearly cancel; most samples stay at the code's error
    assert np.percentile(ang, 99) <= 3.0
    assert np.abs(b[:, 16] - a[:, 16]).max() <= 1e-3
    # at voxel centres (no interpolation) the octahedral code bounds the error
    n = 48
    idx = np.stack(np.meshgrid(*[np.arange(2, n - 2, 3)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pc = ((idx + 0.5) / n).astype(np.float32)
    a, b = gf.sample(pc), gp.sample(pc)
    ang = np.degrees(np.arccos(np.clip(np.sum(a[:, 13:16] * b[:, 13:16], axis=1), -1, 1)))
    assert ang.max() <= 1.6


def test_packed_S_positive_definite(ctx, c2w):
    gp = ctx.upload_grid(dict(c2w["grids"][0], precision="packed"))
    n = 48
    idx = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)[::7]
    pos = ((idx + 0.5) / n).astype(np.float32)  # voxel centres: the stored records themselves
    s = gp.sample(pos)[:, 1:7].astype(np.float64)
    S = np.zeros((len(s), 3, 3))
    S[:, 0, 0], S[:, 1, 1], S[:, 2, 2] = s[:, 0], s[:, 1], s[:, 2]
    S[:, 0, 1] = S[:, 1, 0] = s[:, 3]
    S[:, 0, 2] = S[:, 2, 0] = s[:, 4]
    S[:, 1, 2] = S[:, 2, 1] = s[:, 5]
    lam = np.linalg.eigvalsh(S)
    live = lam[:, 2] > 0
    assert live.sum() > 1000
    assert np.all(lam[live, 0] > 0)


@pytest.mark.parametrize("wl,size,spp_gpu,spp_ref", [("c1", 64, 1024, 1024), ("c2", 32, 4096, 2048)])
def test_packed_frame_converges_to_oracle_fp64(ctx, wl, size, spp_gpu, spp_ref):
    kw = dict(width=size, height=size, spp=spp_gpu)
    if wl == "c2":
        kw["n"] = 64
    w = scenes.make_workload(wl, **kw)
    grids = [ctx.upload_grid(dict(g, precision="packed")) for g in w["grids"]]
    env = ctx.upload_env(w["env"])
    img = ctx.render(dict(w["render"], seed=101), grids, w["instances"], env)
    ref = O.Scene(w["grids"], w["instances"], w["env"], "f64").render(dict(w["render"], spp=spp_ref, seed=202))
    e = relmse(img, ref)
    print(f"packed {wl} {size}^2: relMSE(GPU {spp_gpu} spp, oracle FP64 {spp_ref} spp) = {e:.3e}")
    assert e <= 1e-3


def test_fp64_accumulation(ctx):
    """The framebuffer sums against sums of the same paths taken on the host.  A
    48^2 x 64 spp frame has one sample per work item, so its chunk sums are the
    paths' own values L / 64.  Path (pixel, k) of that frame is path (pixel, 0) of
    a 1-spp render that starts at sample k (the random stream is keyed by the
    sample index), and L / 64 is that frame's value scaled by a power of two.
    The FP32 form adds the 64 values in sample order in FP32; the compensated
    form adds them in FP64 and rounds once, i.e. gives the correctly rounded sum
    but for rare near-ties."""
    import math

    w = scenes.make_workload("c1", width=48, height=48, spp=64)
    grids = [ctx.upload_grid(g) for g in w["grids"]]
    env = ctx.upload_env(w["env"])
    a = ctx.render(w["render"], grids, w["instances"], env)
    b = ctx.render(dict(w["render"], fp64_accum=True), grids, w["instances"], env)
    one = np.stack([ctx.render(dict(w["render"], spp=1, sample_begin=k), grids, w["instances"], env)
                    for k in range(64)])
    vals = (one / np.float32(64)).astype(np.float32)
    seq = np.zeros_like(a)
    for k in range(64):
        seq = (seq + vals[k]).astype(np.float32)
    assert np.array_equal(a, seq), float(np.abs(a - seq).max())
    flat = vals.reshape(64, -1).astype(np.float64)
    exact = np.array([math.fsum(flat[:, i]) for i in range(flat.shape[1])]).astype(np.float32).reshape(b.shape)
    ulp = np.spacing(np.abs(exact).astype(np.float32))
    assert np.all(np.abs(b - exact) <= ulp)
    assert np.mean(b == exact) >= 0.999
    # and the FP32 form's rounding is visible against it
    assert not np.array_equal(a, b)
