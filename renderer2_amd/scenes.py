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
    # the config-1 field at 512^3 (packed records), 1024^2, 256 spp
    "c4": dict(kind=1, n=512, scale=4.0, octaves=3, albedo=0.8, env="white", env_w=32, env_h=16,
               width=1024, height=1024, spp=256, max_depth=8, precision="packed"),
}


def synth_volume(kind: int, n: int, seed: int = 1, scale: float = 4.0, octaves: int = 3) -> np.ndarray:
    rec = np.zeros((n, n, n, 32), np.uint8)
    capi.check(capi.lib().r2_synth_volume(kind, n, seed, scale, octaves, capi.ptr(rec)))
    return rec


def synth_env(kind: str, w: int, h: int) -> np.ndarray:
    out = np.zeros((h, w, 3), np.float32)
    capi.check(capi.lib().r2_synth_env(0 if kind == "white" else 1, w, h, capi.ptr(out)))
    return out


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
                  stream=0, use_brick_grid=True, **cam)
    return dict(name=name, grids=[grid], instances=[dict(grid=0, pos=(0, 0, 0), orientation=(0, 1, 0),
                                                         scaling=(1, 1, 1))],
                env=env, render=render, params=w)
