"""Sparse grids: the sparse form of r2_grid_desc (8^3 leaves with their origins and
active masks, as a NanoVDB tree holds a medium, raytracer/src/medium/nvdb_medium.cu:130-173).

* The leaves of a dense grid, uploaded sparse (any order, with masks or with the
  dense active rule), give the device grid of the dense upload: every voxel
  lookup, the brick grid, the grid info, and the rendered frame, bit for bit.
* A 2048^3 index extent holding a small volume far from the origin (dense
  records would take 275 GB) uploads from its leaves.  Its lookups, active box,
  extrema and brick grid equal those of the same volume as a small dense grid
  (translated by a multiple of 64 voxels, so that bricks and mips align), and
  its converged frame matches that grid's placed at the same world position.
"""
import numpy as np
import pytest

from renderer2_amd import scenes

pytestmark = pytest.mark.gpu


def _leaves(rec, masks=True, keep_empty=False, seed=0):
    """(origins, leaf records, masks) of a dense (nz, ny, nx, 32) record array."""
    nz, ny, nx = rec.shape[:3]
    dens = rec[..., :4].copy().view(np.float32)[..., 0]
    active = dens > 0
    for z in (0, nz - 1):
        for y in (0, ny - 1):
            for x in (0, nx - 1):
                active[z, y, x] = True
    org, lrec, lmask = [], [], []
    for oz in range(0, nz, 8):
        for oy in range(0, ny, 8):
            for ox in range(0, nx, 8):
                blk = np.zeros((8, 8, 8, 32), np.uint8)
                act = np.zeros((8, 8, 8), bool)
                sz, sy, sx = min(8, nz - oz), min(8, ny - oy), min(8, nx - ox)
                blk[:sz, :sy, :sx] = rec[oz:oz + sz, oy:oy + sy, ox:ox + sx]
                act[:sz, :sy, :sx] = active[oz:oz + sz, oy:oy + sy, ox:ox + sx]
                if not act.any() and not keep_empty:
                    continue
                org.append((ox, oy, oz))
                lrec.append(blk.reshape(512, 32))
                lmask.append(np.packbits(act.reshape(512), bitorder="little"))
    org, lrec, lmask = np.array(org, np.int32), np.array(lrec), np.array(lmask, np.uint8)
    perm = np.random.default_rng(seed).permutation(len(org))
    out = dict(leaf_origins=org[perm], records=lrec[perm])
    if masks:
        out["leaf_masks"] = lmask[perm]
    return out


def _all_lookups(grid, shape):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(-1, nz + 1), np.arange(-1, ny + 1), np.arange(-1, nx + 1), indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.int32)
    return grid.lookup(xyz)


@pytest.mark.parametrize("masks,keep_empty", [(True, False), (False, True), (True, True)])
def test_sparse_upload_equals_dense(ctx, masks, keep_empty):
    w = scenes.make_workload("c2", n=56, width=48, height=48, spp=4)
    g = w["grids"][0]
    dense = ctx.upload_grid(g)
    sp = dict(g, **_leaves(g["records"], masks=masks, keep_empty=keep_empty))
    sparse = ctx.upload_grid(sp)
    assert np.array_equal(_all_lookups(dense, g["shape"]), _all_lookups(sparse, g["shape"]))
    for a, b in zip(dense.download_bricks(), sparse.download_bricks()):
        if isinstance(a, list):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        else:
            assert np.array_equal(a, b)
    di, si = dense.info(), sparse.info()
    for k in di:
        if k in ("device_bytes", "n_leaves"):
            continue
        assert np.array_equal(np.asarray(di[k]), np.asarray(si[k])), k
    env = ctx.upload_env(w["env"])
    a = ctx.render(w["render"], [dense], w["instances"], env)
    b = ctx.render(w["render"], [sparse], w["instances"], env)
    assert np.array_equal(a, b)


def test_sparse_large_extent(ctx):
    """A 56^3 volume placed at index (1024, 1536, 512) of a 2048^3 extent."""
    w = scenes.make_workload("c2", n=56, width=48, height=48, spp=8, max_depth=8)
    g = w["grids"][0]
    T = np.array([1024, 1536, 512], np.int32)  # multiples of 64 voxels
    lv = _leaves(g["records"])
    vs = np.float32(g["map_mat"][0])
    big = dict(g, shape=(2048, 2048, 2048), leaf_origins=lv["leaf_origins"] + T, records=lv["records"],
               leaf_masks=lv["leaf_masks"],
               map_vec=tuple(float(np.float32(v) - np.float32(t) * vs) for v, t in zip(g["map_vec"], T)))
    small = ctx.upload_grid(g)
    large = ctx.upload_grid(big)
    si, li = small.info(), large.info()
    assert np.array_equal(li["ibb_min"], np.asarray(si["ibb_min"]) + T)
    assert np.array_equal(li["size"], si["size"])
    assert li["max_density"] == si["max_density"] and li["n_leaves"] == si["n_leaves"]
    assert li["min_density"] == 0.0
    # lookups: the volume's voxels and their neighbours, and far away
    rng = np.random.default_rng(4)
    xyz = rng.integers(-2, 58, size=(50000, 3)).astype(np.int32)
    assert np.array_equal(large.lookup(xyz + T), small.lookup(xyz))
    far = rng.integers(0, 2048, size=(20000, 3)).astype(np.int32)
    bg = np.zeros(32, np.uint8)
    bg[28:30] = np.frombuffer(np.uint16(0x3C00).tobytes(), np.uint8)
    inside = np.all((far >= T) & (far < T + 56), axis=1)
    assert (large.lookup(far)[~inside] == bg).all()
    # brick grid: ranges and mips of the aligned region equal the small grid's
    rs, _, ms = small.download_bricks(with_atlas=False)
    rl, _, ml = large.download_bricks(with_atlas=False)
    nbs = tuple(int(v) for v in si["n_bricks"])
    nbl = tuple(int(v) for v in li["n_bricks"])
    rs = rs.reshape(nbs[2], nbs[1], nbs[0])
    rl = rl.reshape(nbl[2], nbl[1], nbl[0])
    bt = T // 8
    assert np.array_equal(rl[bt[2]:bt[2] + nbs[2], bt[1]:bt[1] + nbs[1], bt[0]:bt[0] + nbs[0]], rs)
    # far from the volume every brick is empty
    assert (rl[:bt[2] - 1] == rl[0, 0, 0]).all() and (rl[bt[2] + nbs[2] + 1:] == rl[0, 0, 0]).all()
    assert li["n_nonempty_bricks"] >= si["n_nonempty_bricks"]
    for lvl in range(3):
        a = ms[lvl].reshape(nbs[2] >> (lvl + 1), nbs[1] >> (lvl + 1), nbs[0] >> (lvl + 1))
        b = ml[lvl].reshape(nbl[2] >> (lvl + 1), nbl[1] >> (lvl + 1), nbl[0] >> (lvl + 1))
        o = bt >> (lvl + 1)
        assert np.array_equal(b[o[2]:o[2] + a.shape[0], o[1]:o[1] + a.shape[1], o[0]:o[0] + a.shape[2]], a)
    # the same world volume: the two index-to-model maps differ in rounding, so paths
    # part ways early; the converged frames agree
    env = ctx.upload_env(w["env"])
    cfg = dict(w["render"], width=32, height=32, spp=4096)
    fa = ctx.render(cfg, [small], w["instances"], env).astype(np.float64)
    fb = ctx.render(dict(cfg, seed=77), [large], w["instances"], env).astype(np.float64)
    e = float(np.mean((fa - fb) ** 2 / (fa * fa + 1e-2)))
    print(f"2048^3 sparse extent: relMSE against the dense 56^3 grid {e:.2e}, "
          f"{np.mean(np.all(fa == fb, axis=2)):.4f} of pixels identical")
    assert e < 2e-3
