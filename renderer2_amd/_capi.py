"""ctypes binding of include/r2.h (librenderer2.so).

The library is loaded from renderer2_amd/lib; there is no fallback path: if the
native library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import os

LIB_PATH = Path(os.environ.get("R2_LIB_PATH", Path(__file__).resolve().parent / "lib" / "librenderer2.so"))

R2_OK = 0
R2_ERR = {
    -1: "R2_ERR_INVALID",
    -2: "R2_ERR_HIP",
    -3: "R2_ERR_OOM",
    -4: "R2_ERR_UNSUPPORTED",
    -5: "R2_ERR_IO",
    -6: "R2_ERR_CANCELLED",
    -7: "R2_ERR_BUSY",
    -8: "R2_ERR_DEADLINE",
}
R2_ERR_CANCELLED = -6
R2_PREC_FP32 = 0
R2_PREC_PACKED = 1
R2_IMAGE_PNG = 0
R2_IMAGE_JPG = 1
R2_FLAG_BRICK_GRID = 1
R2_FLAG_FP64_ACCUM = 2

f3 = C.c_float * 3
f9 = C.c_float * 9
f12 = C.c_float * 12
i3 = C.c_int32 * 3


class GridDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("records", C.c_void_p), ("records_on_device", C.c_int32),
        ("map_mat", f9), ("map_vec", f3),
        ("sigma_a", f3), ("sigma_s", f3),
        ("voxel_size_filtering", C.c_float),
        ("precision", C.c_int32),
        ("n_leaves", C.c_int64), ("leaf_origins", C.c_void_p), ("leaf_masks", C.c_void_p),
    ]


class GridInfo(C.Structure):
    _fields_ = [
        ("index_to_model", f12), ("model_to_index", f12),
        ("model_to_scaled", f12), ("scaled_to_model", f12),
        ("ibb_min", i3), ("size", i3), ("n_bricks", i3),
        ("n_leaves", C.c_int64), ("n_nonempty_bricks", C.c_int64),
        ("max_density", C.c_float), ("inv_max_density", C.c_float),
        ("bbox_min", f3), ("bbox_max", f3),
        ("device_bytes", C.c_int64),
        ("min_density", C.c_float),
    ]


class NvdbInfo(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("ibb_min", i3),
        ("active_voxels", C.c_int64), ("map_mat", f9), ("map_vec", f3),
        ("leaf_count", C.c_int32), ("lower_count", C.c_int32), ("upper_count", C.c_int32),
    ]


class Instance(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("pos", f3), ("orientation", f3), ("scaling", f3)]


R2_SHAPE_CIRCLE, R2_SHAPE_SPHERE = 0, 1
R2_MAX_SURFACES = 8


class Surface(C.Structure):
    _fields_ = [
        ("shape", C.c_int32), ("pos", f3), ("radius", C.c_float), ("normal", f3),
        ("diffuse", f3), ("diffuse_roughness", C.c_float), ("specular", f3),
        ("alpha_x", C.c_float), ("alpha_y", C.c_float), ("eta_i", C.c_float), ("eta_t", C.c_float),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
        ("cam_pos", f3), ("cam_look_at", f3), ("cam_up", f3), ("fov_deg", C.c_float),
        ("seed", C.c_uint32), ("stream", C.c_uint32),
        ("n_instances", C.c_int32), ("instances", C.POINTER(Instance)),
        ("env", C.c_void_p),
        ("flags", C.c_uint32),
        ("sample_begin", C.c_int32),
        ("n_surfaces", C.c_int32), ("surfaces", C.POINTER(Surface)),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("last_kernel_ms", C.c_double), ("last_total_ms", C.c_double),
        ("last_paths", C.c_int64), ("last_launches", C.c_int32),
    ]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgba", C.c_void_p)]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int32), ("n_triangles", C.c_int32),
        ("positions", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p),
        ("material", C.c_void_p), ("n_materials", C.c_int32),
        ("mat_diffuse", C.c_void_p), ("mat_specular", C.c_void_p),
        ("mat_roughness", C.c_void_p), ("mat_ior", C.c_void_p),
        ("pos", f3), ("orientation", f3), ("scaling", f3),
        ("uvs", C.c_void_p), ("n_textures", C.c_int32), ("textures", C.c_void_p), ("mat_textures", C.c_void_p),
    ]


class FilterParams(C.Structure):
    _fields_ = [
        ("voxel_size", C.c_float), ("samples_per_voxel", C.c_int32),
        ("estimation_iterations", C.c_int32), ("sigma_t", C.c_float), ("alpha", C.c_float),
        ("clip_rays", C.c_int32), ("seed", C.c_uint32), ("lod", C.c_int32),
    ]


class FilterGrid(C.Structure):
    _fields_ = [
        ("num_voxels", i3),
        ("index_to_model", f12), ("model_to_index", f12),
        ("model_to_world", f12), ("world_to_model", f12),
        ("world_bb_min", f3), ("world_bb_max", f3),
    ]


class FilterStats(C.Structure):
    _fields_ = [
        ("voxels", C.c_int64), ("active_voxels", C.c_int64), ("pairs", C.c_int64), ("rays", C.c_int64),
        ("filter_kernel_ms", C.c_double), ("total_ms", C.c_double),
    ]


class BrickgridInfo(C.Structure):
    _fields_ = [("n_bricks", C.c_int32 * 3), ("min_density", C.c_float), ("max_density", C.c_float),
                ("n_nonempty_bricks", C.c_int64)]


class PipelineReport(C.Structure):
    _fields_ = [
        ("lods_filtered", C.c_int32), ("filter_ms", C.c_double),
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
        ("render_ms", C.c_double), ("mpaths_per_s", C.c_double),
        ("media_rendered", C.c_int32), ("objects_skipped", C.c_int32),
        ("surfaces_rendered", C.c_int32),
        ("brickgrids_matched", C.c_int32), ("brickgrids_differing", C.c_int32),
    ]


class ConfigSummary(C.Structure):
    _fields_ = [
        ("frame_width", C.c_int32), ("frame_height", C.c_int32), ("samples", C.c_int32),
        ("gamma", C.c_float), ("use_brick_grid", C.c_int32), ("filter_active", C.c_int32),
        ("min_voxel_size", C.c_float), ("lods", C.c_int32), ("samples_per_voxel", C.c_int32),
        ("sigma_t", C.c_float), ("estimation_iterations", C.c_int32), ("alpha", C.c_float),
        ("clip_rays", C.c_int32), ("fov", C.c_float), ("cam_pos", f3), ("cam_look_at", f3), ("cam_up", f3),
        ("n_media", C.c_int32), ("n_meshes", C.c_int32), ("n_surfaces", C.c_int32), ("n_filter_objects", C.c_int32),
        ("max_depth", C.c_int32), ("n_render_objects", C.c_int32),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_inflight", C.c_int32), ("slice_paths", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("params", RenderParams), ("out_rgb", C.c_void_p), ("deadline_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64), ("cancelled", C.c_int64),
        ("paths_per_sec", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
        ("peak_bytes", C.c_int64), ("expired", C.c_int64),
    ]


# (name, restype, argtypes) of every entry point declared in include/r2.h
SIGNATURES = [
    ("r2_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("r2_destroy", C.c_int, [C.c_void_p]),
    ("r2_last_error", C.c_char_p, [C.c_void_p]),
    ("r2_version", C.c_int, []),
    ("r2_stream", C.c_void_p, [C.c_void_p]),
    ("r2_grid_upload", C.c_int, [C.c_void_p, C.POINTER(GridDesc), C.POINTER(C.c_void_p)]),
    ("r2_grid_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("r2_grid_get_info", C.c_int, [C.c_void_p, C.POINTER(GridInfo)]),
    ("r2_grid_download_bricks", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("r2_brickgrid_read", C.c_int, [C.c_char_p, C.POINTER(BrickgridInfo), C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_char_p, C.c_int32]),
    ("r2_grid_lookup", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("r2_grid_sample", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("r2_grid_write_brickgrid", C.c_int, [C.c_void_p, C.c_void_p, C.c_char_p, C.c_char_p, C.c_int32]),
    ("r2_nvdb_write", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_char_p, C.c_int32]),
    ("r2_nvdb_read", C.c_int, [C.c_char_p, C.POINTER(NvdbInfo), C.c_void_p, C.c_int64, C.c_char_p, C.c_int32]),
    ("r2_nvdb_read_leaves", C.c_int, [C.c_char_p, C.POINTER(NvdbInfo), C.c_void_p, C.POINTER(C.c_int64), C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_char_p, C.c_int32]),
    ("r2_dda_cells", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_float, C.c_int32, C.c_int32, C.c_void_p,
                               C.c_void_p]),
    ("r2_env_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("r2_env_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("r2_render", C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_void_p, C.c_int32]),
    ("r2_render_progressive", C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_void_p, C.c_int32,
                                        C.c_void_p, C.c_void_p]),
    ("r2_render_noalloc", C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_void_p]),
    ("r2_render_paths", C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_int32, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    ("r2_scene_intersect", C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_int32, C.c_void_p, C.c_void_p]),
    ("r2_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("r2_mesh_upload", C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    ("r2_mesh_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("r2_filter_setup", C.c_int, [C.c_void_p, C.c_float, C.POINTER(FilterGrid)]),
    ("r2_filter_lod", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FilterParams), C.c_void_p, C.c_int32,
                                C.POINTER(FilterGrid)]),
    ("r2_filter_voxels", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FilterParams), C.c_int32, C.c_void_p,
                                   C.c_void_p]),
    ("r2_filter_get_stats", C.c_int, [C.c_void_p, C.POINTER(FilterStats)]),
    ("r2_server_start", C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    ("r2_server_submit", C.c_int64, [C.c_void_p, C.POINTER(Request)]),
    ("r2_server_poll", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    ("r2_server_wait", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    ("r2_server_cancel", C.c_int, [C.c_void_p, C.c_int64]),
    ("r2_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    ("r2_server_stop", C.c_int, [C.c_void_p]),
    ("r2_postprocess", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p, C.c_void_p,
                                 C.c_int32]),
    ("r2_pipeline_run", C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(PipelineReport), C.c_char_p,
                                  C.c_int32]),
    ("r2_image_write", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_char_p,
                                 C.c_int32]),
    ("r2_config_summary", C.c_int, [C.c_char_p, C.POINTER(ConfigSummary), C.c_char_p, C.c_int32]),
    ("r2_synth_volume", C.c_int, [C.c_int32, C.c_int32, C.c_uint32, C.c_float, C.c_int32, C.c_void_p]),
    ("r2_synth_env", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    ("r2_synth_mesh", C.c_int, [C.c_int32, C.c_int32, C.c_float, C.c_uint32, C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
]


# r2_progress_fn: int (*)(void* user, int32_t samples_done, int32_t samples_total)
PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_int32)


class R2Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{R2_ERR.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load librenderer2.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: run `python -m renderer2_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, ctx=None) -> int:
    if rc < 0:
        msg = ""
        if ctx is not None:
            m = lib().r2_last_error(ctx)
            msg = m.decode() if m else ""
        raise R2Error(rc, msg)
    return rc


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
