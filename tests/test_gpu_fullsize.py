"""Parity at the benchmark workloads' full sizes, through checks that do not
depend on the oracle rendering the whole frame.

  * 64^3 isotropic volume, 256x256: the full GPU frame at 4096 spp (256
    launches of the workload's 16 spp) against the committed full-frame FP64
    oracle image (4096 spp): relMSE <= 1e-3.
  * 256^3 anisotropic volume, 1024x1024: the GPU frame at 2048 spp (32 launches
    of 64 spp) against the committed FP64 oracle image (4096 spp, every 8th
    pixel in x and y): relMSE <= 1e-3, the bar of the render metric.
  * 512^3 isotropic volume in packed records, 1024x1024, compensated FP64
    framebuffer sums: 4096 spp (16 launches of 256 spp) against its FP64 oracle
    image: relMSE <= 1e-3.
  * 2.4 M-triangle mesh filtered to its 512^3 LOD: the dense output of
    r2_filter_lod bit-exact with the filter oracle on 10,000 uniformly random
    voxels and 10,000 random voxels that the filter found occupied.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import renderer2_amd as r2
from renderer2_amd import scenes

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def relmse(x, r):
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def _converged(ctx, name, launches, precision=None):
    kw = {"precision": precision} if precision else {}
    w = scenes.make_workload(name, seed=1, **kw)
    grids = [ctx.upload_grid(g) for g in w["grids"]]
    env = ctx.upload_env(w["env"])
    cfg = w["render"]
    acc = np.zeros((cfg["height"], cfg["width"], 3), np.float64)
    for k in range(launches):
        acc += ctx.render(dict(cfg, sample_begin=k * cfg["spp"]), grids, w["instances"], env)
    for g in grids:
        g.free()
    z = np.load(GOLDEN / f"{name}_oracle_fp64.npz", allow_pickle=False)
    st = int(z["stride"])
    return relmse(acc[::st, ::st] / launches, z["image"])


def test_c1_full_frame_converges(ctx):
    e = _converged(ctx, "c1", 256)
    print(f"64^3 isotropic, 256^2, 4096 spp: relMSE vs FP64 oracle (full frame) = {e:.3e}")
    assert e <= 1e-3


def test_c2_full_frame_converges(ctx):
    e = _converged(ctx, "c2", 32)
    print(f"256^3 aniso, 1024^2, 2048 spp: relMSE vs FP64 oracle = {e:.3e}")
    assert e <= 1e-3


def test_c4_packed_full_frame_converges(ctx):
    assert scenes.WORKLOADS["c4"]["fp64_accum"]
    e = _converged(ctx, "c4", 16, precision="packed")
    print(f"512^3 packed, 1024^2, 4096 spp: relMSE vs FP64 oracle = {e:.3e}")
    assert e <= 1e-3


def test_c3_finest_lod_bit_exact(ctx):
    w = scenes.FILTER_WORKLOADS["c3"]
    mesh = scenes.make_filter_mesh("c3", seed=1)
    gm = ctx.upload_mesh(mesh)
    vs = scenes.finest_voxel_size(mesh, w["finest"])
    params = r2.filter_params(vs, w["samples"], w["iterations"], w["sigma_t"], w["alpha"], w["clip_rays"], seed=1,
                              lod=0)
    rec, g = ctx.filter_lod(gm, params)
    flat = rec.reshape(-1, 32)
    dens = flat[:, :4].copy().view(np.float32)[:, 0]
    rng = np.random.default_rng(99)
    occupied = np.flatnonzero(dens > 0)
    assert occupied.size > 1_000_000
    ids = np.concatenate([rng.integers(0, flat.shape[0], 10_000), rng.choice(occupied, 10_000, replace=False)])
    of = O.MeshFilter(mesh)
    og = of.setup(vs)
    assert tuple(int(v) for v in og.nv) == tuple(int(v) for v in g.num_voxels)
    ref = of.run(og, samples=w["samples"], iterations=w["iterations"], sigma_t=w["sigma_t"], alpha=w["alpha"],
                 clip_rays=w["clip_rays"], seed=1, lod=0, ids=ids.astype(np.int64), threads=16)
    same = np.all(flat[ids] == ref, axis=1)
    print(f"512^3 LOD: {same.mean():.6f} of {len(ids)} voxels bit-exact ({occupied.size} occupied voxels)")
    assert same.all()
