"""Benchmark workloads of the volumetric hot path, built from the seeded
synthetic generators of the native library (r2_synth_volume / r2_synth_env).

A workload is a plain dict: grid (records + NanoVDB map + optical parameters),
instance transform, environment map and render settings, consumable by both
the product path (renderer2_amd.Renderer) and the CPU checker.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi

CAMERA = dict(cam_pos=(0.0, 0.0, 2.5), cam_look_at=(0.0, 0.0, 0.0), cam_up=(0.0, 1.0, 0.0), fov_deg=40.0)
# The anisotropic workload's camera is not fixed by its definition; at distance
# 2.0 the unit cube fills the 40 degree frame, so every path meets the volume.
CAMERA_CLOSE = dict(CAMERA, cam_pos=(0.0, 0.0, 2.0))

# name -> generator parameters and render settings
WORKLOADS = {
    # 64^3 isotropic value-noise volume in [0,4], albedo 0.8, constant white env, 256^2, 16 spp, depth 8
    "c1": dict(kind=1, n=64, scale=4.0, octaves=3, albedo=0.8, env="white", env_w=32, env_h=16,
               width=256, height=256, spp=16, max_depth=8),
    # 256^3 anisotropic fibre SGGX volume in [0,8], albedo 0.9, procedural sky + sun, 1024^2, 64 spp, depth 16
    "c2": dict(kind=2, n=256, scale=8.0, octaves=4, albedo=0.9, env="sky", env_w=512, env_h=256,
               width=1024, height=1024, spp=64, max_depth=16, camera="close"),
    # the config-1 field at 512^3 (packed records), 1024^2, 256 spp, compensated
    # FP64 framebuffer sums
    "c4": dict(kind=1, n=512, scale=4.0, octaves=3, albedo=0.8, env="white", env_w=32, env_h=16,
               width=1024, height=1024, spp=256, max_depth=8, precision="packed", fp64_accum=True),
}


def synth_volume(kind: int, n: int, seed: int = 1, scale: float = 4.0, octaves: int = 3) -> np.ndarray:
    rec = np.zeros((n, n, n, 32), np.uint8)
    capi.check(capi.lib().r2_synth_volume(kind, n, seed, scale, octaves, capi.ptr(rec)))
    return rec


def synth_env(kind: str, w: int, h: int) -> np.ndarray:
    out = np.zeros((h, w, 3), np.float32)
    capi.check(capi.lib().r2_synth_env(0 if kind == "white" else 1, w, h, capi.ptr(out)))
    return out


# materials of the synthetic mesh: 0 = displaced sphere, 1 = foliage quads
# (diffuse rgb, specular rgb, SGGX roughness sigma, ior)
MESH_MATERIALS = dict(
    mat_diffuse=np.array([[0.60, 0.50, 0.40], [0.20, 0.50, 0.10]], np.float32),
    mat_specular=np.array([[0.04, 0.04, 0.04], [0.10, 0.10, 0.10]], np.float32),
    mat_roughness=np.array([0.30, 0.60], np.float32),
    mat_ior=np.array([1.50, 1.33], np.float32),
)

# mesh -> SGGX filter workload: ~2 M triangle displaced sphere (12 res^2) plus
# 200 k thin quads of side 0.01 in a shell around it, filtered to LOD grids of
# 512^3, 256^3 and 128^3 (voxel size doubling from the finest), 64 rays per
# voxel, 8 density iterations, rays clipped to the voxel's bounding sphere
FILTER_WORKLOADS = {
    # filter settings of the reference's example config (SigmaT 10, Alpha 0.1,
    # ClipRays true); one constant material for sphere and quads: the quad
    # generator's base colour (0.8, 0.443, 0.231), roughness 0.5, ior 1.5
    "c3": dict(res=408, n_quads=200_000, quad_size=0.01, finest=512, lods=3, samples=64, iterations=8,
               sigma_t=10.0, alpha=0.1, clip_rays=True,
               material=dict(diffuse=(0.800000, 0.442748, 0.231366), specular=(0.04, 0.04, 0.04),
                             roughness=0.5, ior=1.5)),
}


def make_filter_mesh(name: str = "c3", seed: int = 1, **overrides) -> dict:
    """The mesh of a filter workload with its constant material."""
    w = dict(FILTER_WORKLOADS[name])
    w.update(overrides)
    m = synth_mesh(w["res"], w["n_quads"], w["quad_size"], seed=seed)
    mat = w["material"]
    m["mat_diffuse"] = np.array([mat["diffuse"]] * 2, np.float32)
    m["mat_specular"] = np.array([mat["specular"]] * 2, np.float32)
    m["mat_roughness"] = np.array([mat["roughness"]] * 2, np.float32)
    m["mat_ior"] = np.array([mat["ior"]] * 2, np.float32)
    return m


def synth_mesh(res: int, n_quads: int = 0, quad_size: float = 0.01, seed: int = 1) -> dict:
    """Seeded noise-displaced cube-sphere (12 res^2 triangles) plus thin quads."""
    L = capi.lib()
    nv, nt = C.c_int32(0), C.c_int32(0)
    capi.check(L.r2_synth_mesh(res, n_quads, quad_size, seed, C.byref(nv), C.byref(nt), None, None, None, None))
    pos = np.zeros((nv.value, 3), np.float32)
    nrm = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.int32)
    mat = np.zeros(nt.value, np.int32)
    capi.check(L.r2_synth_mesh(res, n_quads, quad_size, seed, C.byref(nv), C.byref(nt), capi.ptr(pos), capi.ptr(nrm),
                               capi.ptr(idx), capi.ptr(mat)))
    mesh = dict(positions=pos, normals=nrm, indices=idx, material=mat, pos=(0.0, 0.0, 0.0),
                orientation=(0.0, 1.0, 0.0), scaling=(1.0, 1.0, 1.0))
    mesh.update({k: v.copy() for k, v in MESH_MATERIALS.items()})
    return mesh


def finest_voxel_size(mesh: dict, n: int) -> float:
    """Voxel size that gives about n voxels along the mesh's longest world axis."""
    p = mesh["positions"]
    ext = (p.max(axis=0) - p.min(axis=0)) * np.abs(np.asarray(mesh.get("scaling", (1, 1, 1)), np.float32))
    return float(np.float32(ext.max()) / np.float32(n))


def make_workload(name: str, seed: int = 1, **overrides) -> dict:
    """Build the inputs of one workload.  `overrides` may change n, width,
    height, spp, max_depth (smaller parity-test sizes)."""
    w = dict(WORKLOADS[name])
    w.update(overrides)
    n = w["n"]
    rec = synth_volume(w["kind"], n, seed, w["scale"], w["octaves"])
    vs = np.float32(1.0) / np.float32(n)
    albedo = w["albedo"]
    grid = dict(
        records=rec,
        shape=(n, n, n),
        map_mat=(float(vs), 0.0, 0.0, 0.0, float(vs), 0.0, 0.0, 0.0, float(vs)),
        map_vec=(-0.5, -0.5, -0.5),
        sigma_a=(1.0 - albedo,) * 3,
        sigma_s=(albedo,) * 3,
        voxel_size_filtering=float(vs),
        precision=w.get("precision", "fp32"),
    )
    env = synth_env(w["env"], w["env_w"], w["env_h"])
    cam = CAMERA_CLOSE if w.get("camera") == "close" else CAMERA
    render = dict(width=w["width"], height=w["height"], spp=w["spp"], max_depth=w["max_depth"], seed=seed,
                  stream=0, use_brick_grid=True, fp64_accum=bool(w.get("fp64_accum", False)), **cam)
    return dict(name=name, grids=[grid], instances=[dict(grid=0, pos=(0, 0, 0), orientation=(0, 1, 0),
                                                         scaling=(1, 1, 1))],
                env=env, render=render, params=w)


# The ground of the reference's forest scenes: a circle of radius FLT_MAX through
# the origin with a dark rough diffuse lobe under a rough glossy one
# (raytracer/src/filtering/common_functions.py:94-106)
GROUND = dict(shape="circle", pos=(0.0, 0.0, 0.0), radius=3.402823466e38, normal=(0.0, 1.0, 0.0),
              diffuse=(0.07, 0.07, 0.01), diffuse_roughness=0.999, specular=(0.9, 0.9, 0.9), alpha_x=0.99,
              alpha_y=0.99, eta_i=1.00029, eta_t=1.2)


def mixed_workload(name: str = "c1", ground_y: float = -0.4, seed: int = 1, **overrides) -> dict:
    """A volume workload standing on the forest ground, with a glossy sphere half
    inside the volume's box: the ground cuts the box's bottom, so paths meet
    surfaces both outside and inside the medium.  The camera looks down at the
    scene from above the ground."""
    w = make_workload(name, seed=seed, **overrides)
    ground = dict(GROUND, pos=(0.0, ground_y, 0.0))
    sphere = dict(shape="sphere", pos=(0.45, -0.15, 0.35), radius=0.25, normal=(0.0, 1.0, 0.0),
                  diffuse=(0.6, 0.3, 0.1), diffuse_roughness=0.5, specular=(0.9, 0.9, 0.9), alpha_x=0.2,
                  alpha_y=0.35, eta_i=1.0, eta_t=1.5)
    w["render"].update(surfaces=[ground, sphere], cam_pos=(0.0, 0.7, 2.3), cam_look_at=(0.0, -0.15, 0.0))
    w["name"] = name + "+surfaces"
    return w


def forest_workload(n_trees: int = 256, seed: int = 3, ground: bool = True, width: int = 48, height: int = 48,
                    spp: int = 4, max_depth: int = 8, extent: float = 4.0) -> dict:
    """Many instances of two small grids (a 32^3 fibre volume and a 16^3 value-noise
    volume) on a jittered lattice over [-extent, extent]^2, standing on the forest
    ground: random scales and axis-aligned orientations (getRotation keeps them
    exact), a few neighbouring boxes overlapping, seen from above the canopy under
    the sky + sun map.  (A path whose box meets another box ends, as in the
    reference's Li, so overlaps are kept rare.)"""
    rng = np.random.default_rng(seed)
    fib = make_workload("c2", n=32, width=width, height=height, spp=spp, max_depth=max_depth)
    noise = make_workload("c1", n=16)
    grids = [fib["grids"][0], noise["grids"][0]]
    side = int(np.ceil(np.sqrt(n_trees)))
    cell = 2.0 * extent / side
    orients = [(0, 1, 0), (1, 0, 0), (0, 0, 1), (-1, 0, 0), (0, 0, -1)]
    instances = []
    for k in range(n_trees):
        i, j = k % side, k // side
        sc = float(rng.uniform(0.45, 0.8)) * cell
        sy = sc * float(rng.uniform(0.8, 1.6))
        x = -extent + (i + 0.5) * cell + float(rng.uniform(-0.12, 0.12)) * cell
        z = -extent + (j + 0.5) * cell + float(rng.uniform(-0.12, 0.12)) * cell
        instances.append(dict(grid=int(rng.integers(0, 2)), pos=(x, 0.5 * sy, z),
                              orientation=orients[int(rng.integers(0, len(orients)))], scaling=(sc, sy, sc)))
    render = dict(fib["render"], cam_pos=(0.0, 2.2, extent + 2.5), cam_look_at=(0.0, 0.2, 0.0), fov_deg=55.0)
    if ground:
        render["surfaces"] = [dict(GROUND)]
    return dict(name=f"forest{n_trees}", grids=grids, instances=instances, env=synth_env("sky", 64, 32),
                render=render)


def filtered_grid(fgrid, voxel_size: float, albedo: float = 0.9, records=None, records_dev: int | None = None,
                  precision: str = "fp32") -> dict:
    """Grid description of one filtered LOD (r2_filter_grid): the index -> model map
    is the filter's indexToModel, so the LOD renders in the mesh's model space."""
    i2m = np.array(fgrid.index_to_model, np.float32).reshape(4, 3)
    nx, ny, nz = (int(v) for v in fgrid.num_voxels)
    g = dict(shape=(nz, ny, nx), map_mat=tuple(float(v) for v in i2m[:3].T.reshape(-1)),
             map_vec=tuple(float(v) for v in i2m[3]), sigma_a=(1.0 - albedo,) * 3, sigma_s=(albedo,) * 3,
             voxel_size_filtering=float(voxel_size), precision=precision)
    if records_dev is not None:
        g["records_dev"] = records_dev
    else:
        g["records"] = records
    return g


# serving workload: device-resident LOD grids of the 256^3 anisotropic field at
# 128^3, 256^3 and 512^3 plus the 64^3 value-noise grid, the procedural sky + sun
# environment; requests of resolution {256, 512, 1024}^2, spp {16, 64, 256},
# depth 16, an orbit camera and one LOD grid each
SERVING_SIZES = (256, 512, 1024)
SERVING_SPPS = (16, 64, 256)
SERVING_DEPTH = 16


def serving_grids() -> list:
    """Host descriptions of the serving workload's four grids (128^3, 256^3 and
    512^3 fibre field, 64^3 value noise), in that order."""
    out = [make_workload("c2", n=n)["grids"][0] for n in (128, 256, 512)]
    out.append(make_workload("c1")["grids"][0])
    return out


def serving_requests(n_req: int, rate: float | None, seed: int, n_grids: int = 4, sizes=SERVING_SIZES,
                     spps=SERVING_SPPS, max_depth: int = SERVING_DEPTH, radius: float = 2.5) -> list:
    """Seeded Poisson request stream: (arrival time s, render config, instances) per
    request, each with a resolution, an spp, a grid (LOD) and an orbit camera drawn
    at random.  rate None: all arrive at t = 0."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req)) if rate else np.zeros(n_req)
    reqs = []
    for k in range(n_req):
        res = int(rng.choice(sizes))
        spp = int(rng.choice(spps))
        g = int(rng.integers(n_grids))
        az, el = rng.uniform(0, 2 * np.pi), rng.uniform(-0.3, 0.6)
        cam = (radius * np.cos(el) * np.sin(az), radius * np.sin(el), radius * np.cos(el) * np.cos(az))
        cfg = dict(width=res, height=res, spp=spp, max_depth=max_depth, cam_pos=cam, cam_look_at=(0, 0, 0),
                   cam_up=(0, 1, 0), fov_deg=40.0, seed=1000 + k, stream=0, use_brick_grid=True)
        inst = [dict(grid=g, pos=(0, 0, 0), orientation=(0, 1, 0), scaling=(1, 1, 1))]
        reqs.append((float(arrivals[k]), cfg, inst))
    return reqs
