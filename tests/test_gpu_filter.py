"""GPU parity of the mesh -> SGGX filter against the CPU oracle.

Bar: the 32-byte voxel records of the GPU filter equal the oracle's bit for bit
(same seeds, same watertight triangle test, same closest-hit tie rule, the same
double-evaluated sin/cos/exp/log, sums in ray order).  Grid set-up transforms are
compared bit for bit as well.  This is stricter than the 1e-4 relative error per
coefficient allowed for filtered SGGX grids.
"""
import numpy as np
import pytest

import oracle as O
import renderer2_amd as r2
from renderer2_amd import scenes

pytestmark = pytest.mark.gpu

GRID_FIELDS = ("index_to_model", "model_to_index", "model_to_world", "world_to_model", "world_bb_min", "world_bb_max")


def _mesh(res=10, quads=60, size=0.08, seed=5, **xf):
    m = scenes.synth_mesh(res, quads, size, seed=seed)
    m.update(xf)
    return m


def _same_grid(g_gpu, g_or):
    assert tuple(g_gpu.num_voxels) == tuple(g_or.nv)
    for f in GRID_FIELDS:
        a = np.array(getattr(g_gpu, f), np.float32)
        b = np.array(getattr(g_or, f), np.float32)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f


def _diff_report(gpu, ref):
    bad = np.nonzero(np.any(gpu.reshape(-1, 32) != ref.reshape(-1, 32), axis=1))[0]
    return f"{bad.size} of {gpu.reshape(-1, 32).shape[0]} records differ; first ids {bad[:8].tolist()}"


XFORMS = [
    dict(),
    dict(pos=(0.3, -1.2, 2.0), orientation=(1.0, 0.0, 0.0), scaling=(0.5, 0.5, 0.5)),
    dict(pos=(-4.0, 0.5, 0.25), orientation=(0.0, 0.0, -1.0), scaling=(2.0, 2.0, 2.0)),
]


@pytest.mark.parametrize("xf", range(len(XFORMS)))
def test_filter_setup_matches_oracle(ctx, xf):
    m = _mesh(**XFORMS[xf])
    gm = ctx.upload_mesh(m)
    of = O.MeshFilter(m)
    s = float(np.abs(np.asarray(m.get("scaling", (1, 1, 1)))).max())
    for n in (7, 16, 33, 128):
        vs = scenes.finest_voxel_size(m, n) * s
        _same_grid(gm.setup(vs), of.setup(vs))


@pytest.mark.parametrize(
    "case",
    [
        dict(n=24, samples=64, iterations=8, clip=True, seed=1, xf=0),
        dict(n=32, samples=64, iterations=8, clip=True, seed=7, xf=1),
        dict(n=20, samples=100, iterations=3, clip=False, seed=3, xf=0),
        dict(n=16, samples=150, iterations=12, clip=True, seed=9, xf=2),
        dict(n=28, samples=1, iterations=0, clip=True, seed=2, xf=0),
    ],
    ids=lambda c: f"n{c['n']}-N{c['samples']}-K{c['iterations']}-clip{int(c['clip'])}-xf{c['xf']}",
)
def test_filter_lod_bitexact(ctx, case):
    m = _mesh(**XFORMS[case["xf"]])
    s = float(np.abs(np.asarray(m.get("scaling", (1, 1, 1)))).max())
    vs = scenes.finest_voxel_size(m, case["n"]) * s
    gm = ctx.upload_mesh(m)
    p = r2.filter_params(vs, case["samples"], case["iterations"], clip_rays=case["clip"], seed=case["seed"], lod=1)
    rec, g = ctx.filter_lod(gm, p)
    st = ctx.filter_stats()
    assert st["voxels"] == rec.shape[0] * rec.shape[1] * rec.shape[2]
    assert 0 < st["active_voxels"] < st["voxels"]
    of = O.MeshFilter(m)
    og = of.setup(vs)
    _same_grid(g, og)
    ref = of.run(og, samples=case["samples"], iterations=case["iterations"], clip_rays=case["clip"],
                 seed=case["seed"], lod=1).reshape(rec.shape)
    assert np.array_equal(rec, ref), _diff_report(rec, ref)
    d = O.decode_records(ref)
    assert (d["density"] > 0).sum() > 50  # the case exercises hits


def test_filter_voxels_matches_lod(ctx):
    m = _mesh(res=14, quads=200, size=0.05, seed=11)
    vs = scenes.finest_voxel_size(m, 40)
    gm = ctx.upload_mesh(m)
    p = r2.filter_params(vs, 64, 8, seed=4)
    rec, g = ctx.filter_lod(gm, p)
    flat = rec.reshape(-1, 32)
    rng = np.random.default_rng(0)
    ids = rng.choice(flat.shape[0], 3000, replace=False).astype(np.uint32)
    ids[:5] = [0, flat.shape[0] - 1, 1, 2, 3]
    got = ctx.filter_voxels(gm, p, ids)
    assert np.array_equal(got, flat[ids])
    # and the oracle at the same ids
    of = O.MeshFilter(m)
    ref = of.run(of.setup(vs), samples=64, iterations=8, seed=4, ids=ids.astype(np.int64))
    assert np.array_equal(got, ref), _diff_report(got, ref)


def test_filter_lods_doubling(ctx):
    """The LOD loop of CFilter::runFiltering: voxel size doubles per LOD."""
    m = _mesh(res=12, quads=100, size=0.06, seed=2)
    gm = ctx.upload_mesh(m)
    of = O.MeshFilter(m)
    vs0 = scenes.finest_voxel_size(m, 32)
    for lod in range(3):
        vs = vs0 * (2.0 ** lod)
        rec, g = ctx.filter_lod(gm, r2.filter_params(vs, 32, 8, seed=5, lod=lod))
        og = of.setup(vs)
        ref = of.run(og, samples=32, iterations=8, seed=5, lod=lod).reshape(rec.shape)
        assert np.array_equal(rec, ref), f"lod {lod}: " + _diff_report(rec, ref)


def test_filter_device_output(ctx):
    """A filtered LOD written straight to device memory equals the host copy."""
    from renderer2_amd import _hip

    m = _mesh(res=12, quads=100, size=0.06, seed=2)
    gm = ctx.upload_mesh(m)
    vs = scenes.finest_voxel_size(m, 32)
    g = gm.setup(vs)
    nx, ny, nz = list(g.num_voxels)
    buf = _hip.DeviceBuffer(nx * ny * nz * 32)
    try:
        _, g2 = ctx.filter_lod(gm, r2.filter_params(vs, 64, 8, seed=1), out_dev_ptr=buf.addr)
        host, _ = ctx.filter_lod(gm, r2.filter_params(vs, 64, 8, seed=1))
        dev = buf.to_host(np.uint8).reshape(host.shape)
    finally:
        buf.free()
    assert np.array_equal(dev, host)


def test_filtered_lod_render_converges_to_oracle_fp64(ctx):
    """A filtered LOD as a medium (sigma_a 0, sigma_s = SigmaT, colours and SGGX from
    the records): the GPU frame at 8192 spp converges to the FP64 oracle's at 4096."""
    m = scenes.make_filter_mesh("c3", seed=3, res=12, n_quads=150, quad_size=0.08)
    gm = ctx.upload_mesh(m)
    vs = scenes.finest_voxel_size(m, 32)
    p = r2.filter_params(vs, 64, 8, 10.0, 0.1, True, seed=1, lod=0)
    rec, fg = ctx.filter_lod(gm, p)
    gd = scenes.filtered_grid(fg, vs, records=rec)
    gd["sigma_a"], gd["sigma_s"] = (0.0,) * 3, (10.0,) * 3
    ext = float((m["positions"].max(axis=0) - m["positions"].min(axis=0)).max())
    inst = [dict(grid=0, pos=(0, 0, 0), orientation=(0, 1, 0), scaling=(1.0 / ext,) * 3)]
    w = scenes.make_workload("c1", width=32, height=32, spp=8192)
    g = ctx.upload_grid(gd)
    env = ctx.upload_env(w["env"])
    img = ctx.render(dict(w["render"], seed=101), [g], inst, env)
    ref = O.Scene([gd], inst, w["env"], "f64").render(dict(w["render"], spp=4096, seed=202))
    x, r = img.astype(np.float64), ref.astype(np.float64)
    e = float(np.mean((x - r) ** 2 / (r * r + 1e-2)))
    print(f"filtered 32^3 LOD, 32^2: relMSE(GPU 8192 spp, oracle FP64 4096 spp) = {e:.3e}")
    assert ref.max() > 0 and e <= 1e-3


def _textured_foliage(n_quads=90, seed=11, with_uvs=True):
    """Randomly placed and oriented quads with three materials: a textured diffuse
    colour with a disk-shaped alpha mask and a bumpy normal map; a textured glossy
    colour with a striped alpha mask; constant colours."""
    from test_oracle_filter import quad_mesh
    from test_oracle_textures import quad_uvs

    rng = np.random.default_rng(seed)
    quads = []
    for _ in range(n_quads):
        n = rng.normal(size=3)
        quads.append((rng.uniform(-0.8, 0.8, 3), n / np.linalg.norm(n), float(rng.uniform(0.15, 0.45))))
    m = quad_mesh(quads, material=rng.integers(0, 3, n_quads))
    m["mat_diffuse"] = np.array([[0.5, 0.25, 0.75], [0.9, 0.1, 0.2], [0.3, 0.6, 0.1]], np.float32)
    m["mat_specular"] = np.array([[0.04, 0.5, 1.0], [0.3, 0.3, 0.3], [0.1, 0.2, 0.3]], np.float32)
    m["mat_roughness"] = np.array([0.2, 0.7, 0.4], np.float32)
    m["mat_ior"] = np.array([1.5, 1.25, 1.33], np.float32)
    yy, xx = np.mgrid[0:32, 0:32]
    disk = ((xx - 15.5) ** 2 + (yy - 15.5) ** 2 < 13.0 ** 2).astype(np.float32)
    stripes = ((xx // 4) % 2 == 0).astype(np.float32)
    j, i = np.mgrid[0:16, 0:16].astype(np.float32)
    bumps = np.stack([0.5 + 0.3 * np.sin(i / 2.5), 0.5 + 0.3 * np.cos(j / 3.0), 0.8 + 0.1 * np.sin((i + j) / 4.0)], -1)
    textures = [
        rng.uniform(0.05, 0.95, (16, 16, 4)).astype(np.float32),                          # 0 diffuse
        np.dstack([disk, disk, disk, np.ones_like(disk)]),                                # 1 alpha disk
        np.dstack([bumps, np.ones((16, 16), np.float32)]).astype(np.float32),             # 2 normal map
        rng.uniform(0.0, 0.6, (8, 8, 4)).astype(np.float32),                              # 3 glossy
        np.dstack([stripes, stripes, stripes, np.ones_like(stripes)]),                    # 4 alpha stripes
    ]
    m["textures"] = textures
    m["mat_textures"] = np.array([[0, -1, 2, 1], [-1, 3, -1, 4], [-1, -1, -1, -1]], np.int32)
    if with_uvs:
        m["uvs"] = quad_uvs(n_quads)
    return m


@pytest.mark.parametrize("with_uvs", [True, False])
def test_filter_textured_materials_bit_exact(ctx, with_uvs):
    """Alpha-masked, textured and normal-mapped foliage (shaders.optix.cu:48-152,
    material.hpp:111-156): GPU records equal the oracle's bit for bit.  Without
    texture coordinates every hit reads its textures at tc = (0, 0) and no alpha
    test or normal map applies."""
    m = _textured_foliage(with_uvs=with_uvs)
    gm = ctx.upload_mesh(m)
    of = O.MeshFilter(m)
    for n, lod in ((24, 0), (12, 1)):
        vs = scenes.finest_voxel_size(m, n)
        rec, g = ctx.filter_lod(gm, r2.filter_params(vs, 64, 8, 10.0, 0.1, True, seed=9, lod=lod))
        og = of.setup(vs)
        ref = of.run(og, samples=64, iterations=8, sigma_t=10.0, alpha=0.1, clip_rays=True, seed=9,
                     lod=lod).reshape(rec.shape)
        d = O.decode_records(ref.reshape(-1, 32))
        print(f"textured foliage uvs={with_uvs} {n}^3: {int((d['density'] > 0).sum())} voxels with hits", flush=True)
        assert (d["density"] > 0).sum() > 50
        assert np.array_equal(rec, ref), f"{n}^3: " + _diff_report(rec, ref)
