"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  It is the checker and the CPU baseline, never the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libr2oracle.so"
MAX_GRIDS = 16
NCNT = 7
CNT_NAMES = ["vox", "maj", "brick", "env", "dda", "flight", "scatter"]


class GridDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("records", C.c_void_p),
        ("map_mat", C.c_float * 9), ("map_vec", C.c_float * 3),
        ("sigma_a", C.c_float * 3), ("sigma_s", C.c_float * 3),
        ("voxel_size_filtering", C.c_float),
    ]


class InstanceDesc(C.Structure):
    _fields_ = [("grid", C.c_int), ("pos", C.c_float * 3), ("orientation", C.c_float * 3), ("scaling", C.c_float * 3)]


MAX_SURF = 8
SHAPES = {"circle": 0, "sphere": 1}


class SurfaceDesc(C.Structure):
    _fields_ = [
        ("shape", C.c_int), ("pos", C.c_float * 3), ("radius", C.c_float), ("normal", C.c_float * 3),
        ("diffuse", C.c_float * 3), ("diffuse_roughness", C.c_float), ("specular", C.c_float * 3),
        ("alpha_x", C.c_float), ("alpha_y", C.c_float), ("eta_i", C.c_float), ("eta_t", C.c_float),
    ]


def surface_desc(sf: dict) -> SurfaceDesc:
    """Surface dict (shape, pos, radius, normal, diffuse, diffuse_roughness,
    specular, alpha_x, alpha_y, eta_i, eta_t) -> SurfaceDesc."""
    d = SurfaceDesc()
    d.shape = SHAPES[sf.get("shape", "circle")]
    d.pos[:] = list(sf.get("pos", (0, 0, 0)))
    d.radius = sf.get("radius", 1.0)
    d.normal[:] = list(sf.get("normal", (0, 1, 0)))
    d.diffuse[:] = list(sf.get("diffuse", (0.5, 0.5, 0.5)))
    d.diffuse_roughness = sf.get("diffuse_roughness", 0.0)
    d.specular[:] = list(sf.get("specular", (0, 0, 0)))
    d.alpha_x = sf.get("alpha_x", 0.5)
    d.alpha_y = sf.get("alpha_y", 0.5)
    d.eta_i = sf.get("eta_i", 1.0)
    d.eta_t = sf.get("eta_t", 1.5)
    return d


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_grids", C.c_int), ("grids", GridDesc * MAX_GRIDS),
        ("n_instances", C.c_int), ("instances", C.POINTER(InstanceDesc)),
        ("env_rgb", C.c_void_p), ("env_w", C.c_int), ("env_h", C.c_int),
        ("n_surfaces", C.c_int), ("surfaces", SurfaceDesc * MAX_SURF),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("max_depth", C.c_int),
        ("cam_pos", C.c_float * 3), ("cam_look_at", C.c_float * 3), ("cam_up", C.c_float * 3),
        ("fov_deg", C.c_float),
        ("seed", C.c_uint32), ("stream", C.c_uint32),
        ("use_brick_grid", C.c_int), ("sample_begin", C.c_int), ("threads", C.c_int),
    ]


class Bricks(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nb", C.c_int * 3),
        ("range", C.POINTER(C.c_uint32)), ("nonempty", C.POINTER(C.c_uint8)), ("u8", C.POINTER(C.c_uint8)),
        ("mip", C.POINTER(C.c_uint32) * 3), ("n_nonempty", C.c_size_t),
    ]


class Texture(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("rgba", C.c_void_p)]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int), ("n_triangles", C.c_int),
        ("positions", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p),
        ("material", C.c_void_p), ("n_materials", C.c_int),
        ("mat_diffuse", C.c_void_p), ("mat_specular", C.c_void_p),
        ("mat_roughness", C.c_void_p), ("mat_ior", C.c_void_p),
        ("pos", C.c_float * 3), ("orientation", C.c_float * 3), ("scaling", C.c_float * 3),
        ("uvs", C.c_void_p), ("n_textures", C.c_int), ("textures", C.c_void_p), ("mat_textures", C.c_void_p),
    ]


def mesh_textures(mesh: dict, keep: dict, tex_struct):
    """(uvs pointer, n_textures, textures array, mat_textures pointer) of a mesh dict
    with optional uvs (V,2), textures [list of (H,W,4) float32 RGBA] and
    mat_textures (M,4) int32 (diffuse, glossy, normal map, alpha; -1 none)."""
    uv = mesh.get("uvs")
    keep["uvs"] = None if uv is None else np.ascontiguousarray(uv, np.float32)
    texs = [np.ascontiguousarray(t, np.float32) for t in mesh.get("textures", [])]
    keep["tex_data"] = texs
    arr = (tex_struct * max(1, len(texs)))()
    for i, t in enumerate(texs):
        assert t.ndim == 3 and t.shape[2] == 4
        arr[i].height, arr[i].width = t.shape[0], t.shape[1]
        arr[i].rgba = t.ctypes.data
    keep["tex_arr"] = arr
    mt = mesh.get("mat_textures")
    keep["mat_textures"] = None if mt is None else np.ascontiguousarray(mt, np.int32).reshape(-1, 4)
    return (None if keep["uvs"] is None else keep["uvs"].ctypes.data, len(texs),
            C.cast(arr, C.c_void_p).value if texs else None,
            None if keep["mat_textures"] is None else keep["mat_textures"].ctypes.data)


class FilterGrid(C.Structure):
    _fields_ = [
        ("nv", C.c_int * 3),
        ("index_to_model", C.c_float * 12), ("model_to_index", C.c_float * 12),
        ("model_to_world", C.c_float * 12), ("world_to_model", C.c_float * 12),
        ("world_bb_min", C.c_float * 3), ("world_bb_max", C.c_float * 3),
    ]


class FilterParams(C.Structure):
    _fields_ = [
        ("samples_per_voxel", C.c_int), ("estimation_iterations", C.c_int), ("clip_rays", C.c_int),
        ("lod", C.c_int), ("threads", C.c_int),
        ("sigma_t", C.c_float), ("alpha", C.c_float), ("seed", C.c_uint32),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        vp, u32p = C.c_void_p, C.POINTER(C.c_uint32)
        L.r2o_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.r2o_half_from_float_glm.argtypes = [C.c_float]
        L.r2o_half_from_float_glm.restype = C.c_uint16
        L.r2o_half_from_float_rne.argtypes = [C.c_float]
        L.r2o_half_from_float_rne.restype = C.c_uint16
        L.r2o_half_to_float.argtypes = [C.c_uint16]
        L.r2o_half_to_float.restype = C.c_float
        L.r2o_brick_build.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(Bricks)]
        L.r2o_brick_free.argtypes = [C.POINTER(Bricks)]
        L.r2o_brick_majorant.argtypes = [C.POINTER(Bricks), C.c_int, C.c_int, C.c_int, C.c_int]
        L.r2o_brick_majorant.restype = C.c_float
        L.r2o_brick_density.argtypes = [C.POINTER(Bricks), C.c_int, C.c_int, C.c_int]
        L.r2o_brick_density.restype = C.c_float
        for sfx in ("f32", "f64"):
            R = C.c_float if sfx == "f32" else C.c_double
            g = lambda n: getattr(L, f"{n}_{sfx}")  # noqa: E731
            g("r2o_scene_new").argtypes = [C.POINTER(SceneDesc)]
            g("r2o_scene_new").restype = vp
            g("r2o_scene_free").argtypes = [vp]
            g("r2o_render").argtypes = [vp, C.POINTER(RenderParams), vp, vp]
            g("r2o_render_paths").argtypes = [vp, C.POINTER(RenderParams), C.c_int, vp, vp, vp, vp]
            g("r2o_camera_view_to_world").argtypes = [C.POINTER(RenderParams), vp]
            g("r2o_camera_ray").argtypes = [C.POINTER(RenderParams), C.c_int, C.c_int, C.c_uint32, vp]
            g("r2o_scene_info").argtypes = [vp, C.c_int, vp]
            g("r2o_instance_info").argtypes = [vp, C.c_int, vp]
            g("r2o_filtered_data").argtypes = [vp, C.c_int, vp, vp]
            g("r2o_decode_record").argtypes = [vp, vp]
            g("r2o_dda_cells").argtypes = [vp, C.c_int, vp, R, C.c_int, C.c_int, vp]
            g("r2o_dda_cells").restype = C.c_int
            g("r2o_fresnel").argtypes = [R, R]
            g("r2o_fresnel").restype = R
            g("r2o_sggx_D").argtypes = [vp, vp]
            g("r2o_sggx_D").restype = R
            g("r2o_sggx_sigma").argtypes = [vp, vp]
            g("r2o_sggx_sigma").restype = R
            g("r2o_sggx_vndf").argtypes = [vp, vp, R, R, vp]
            g("r2o_phase_sample").argtypes = [vp, vp, R, vp, C.c_uint32, C.c_uint32, vp]
            g("r2o_phase_sample").restype = R
            g("r2o_phase_p").argtypes = [vp, vp, R, vp, vp, C.c_uint32, C.c_uint32]
            g("r2o_phase_p").restype = R
            g("r2o_env_sample").argtypes = [vp, C.c_uint32, C.c_uint32, vp]
            g("r2o_env_le").argtypes = [vp, vp, vp]
            g("r2o_env_tables").argtypes = [vp, vp, vp, vp]
            g("r2o_tr_segment").argtypes = [vp, vp, C.c_int, C.c_uint32, C.c_uint32]
            g("r2o_tr_segment").restype = R
            g("r2o_sample_segment").argtypes = [vp, vp, C.c_int, C.c_uint32, C.c_uint32, vp]
            g("r2o_sample_segment").restype = C.c_int
            g("r2o_move_to_voxel_border").argtypes = [vp, vp, vp]
            g("r2o_intersect").argtypes = [vp, vp, vp]
            g("r2o_intersect").restype = C.c_int
            g("r2o_direct_mean").argtypes = [vp, C.c_int, vp, vp, vp, R, vp, C.c_int, C.c_uint32, C.c_uint32,
                                             C.c_int64, vp, vp]
            g("r2o_phase_p_mean").argtypes = [vp, vp, R, vp, vp, C.c_int64, C.c_uint32, C.c_uint32, C.c_int, vp]
            g("r2o_bsdf_eval").argtypes = [C.POINTER(SurfaceDesc), vp, vp, vp]
            g("r2o_ggx_eval").argtypes = [C.POINTER(SurfaceDesc), vp, vp, vp, vp]
            g("r2o_ggx_sample_h").argtypes = [C.POINTER(SurfaceDesc), vp, R, R, vp]
            g("r2o_bsdf_sample").argtypes = [C.POINTER(SurfaceDesc), vp, C.c_uint32, C.c_uint32, vp]
            g("r2o_fresnel_ie").argtypes = [R, R, R]
            g("r2o_fresnel_ie").restype = R
            g("r2o_surface_hit").argtypes = [C.POINTER(SurfaceDesc), vp, vp]
            g("r2o_surface_hit").restype = C.c_int
        L.r2o_filter_setup.argtypes = [C.POINTER(MeshDesc), C.c_float, C.POINTER(FilterGrid)]
        L.r2o_filter_new.argtypes = [C.POINTER(MeshDesc)]
        L.r2o_filter_new.restype = vp
        L.r2o_filter_free.argtypes = [vp]
        L.r2o_filter_run.argtypes = [vp, C.POINTER(FilterGrid), C.POINTER(FilterParams), C.c_int64, vp, vp]
        L.r2o_filter_probe.argtypes = [vp, C.POINTER(FilterGrid), C.POINTER(FilterParams), C.c_int64, vp, vp, vp, vp,
                                       vp, vp]
        L.r2o_filter_closest_hit.argtypes = [vp, vp, vp, C.c_float, vp]
        L.r2o_tex_fetch.argtypes = [vp, C.c_float, C.c_float, vp]
        L.r2o_det_expf.argtypes = [C.c_float]
        L.r2o_det_expf.restype = C.c_float
        L.r2o_det_logf.argtypes = [C.c_float]
        L.r2o_det_logf.restype = C.c_float
        L.r2o_det_sincosf.argtypes = [C.c_float, vp, vp]
        L.r2o_det_powf.argtypes = [C.c_float, C.c_float]
        L.r2o_det_powf.restype = C.c_float
        L.r2o_postprocess.argtypes = [vp, C.c_int, C.c_int, C.c_float, vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32).copy()
    k = np.asarray(key, dtype=np.uint32).copy()
    out = np.zeros(4, np.uint32)
    u32p = C.POINTER(C.c_uint32)
    lib().r2o_philox4x32_10(c.ctypes.data_as(u32p), k.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
    return out


class Brickgrid:
    """CPU restatement of the brick majorant grid (host_grid_brick.cpp:61-152)."""

    def __init__(self, dens: np.ndarray):
        dens = np.ascontiguousarray(dens, dtype=np.float32)
        nz, ny, nx = dens.shape
        self.b = Bricks()
        rc = lib().r2o_brick_build(_p(dens), nx, ny, nz, C.byref(self.b))
        if rc != 0:
            raise RuntimeError(f"r2o_brick_build rc={rc}")
        self.nb = tuple(self.b.nb)
        nbx, nby, nbz = self.nb
        self.range = np.ctypeslib.as_array(self.b.range, shape=(nbz * nby * nbx,)).copy()
        self.nonempty = np.ctypeslib.as_array(self.b.nonempty, shape=(nbz * nby * nbx,)).copy()
        self.u8 = np.ctypeslib.as_array(self.b.u8, shape=(nbz * 8 * nby * 8 * nbx * 8,)).copy()
        self.mips = []
        for l in range(3):
            n = (nbx >> (l + 1)) * (nby >> (l + 1)) * (nbz >> (l + 1))
            self.mips.append(np.ctypeslib.as_array(self.b.mip[l], shape=(n,)).copy())

    def majorant(self, bx, by, bz, mip):
        return lib().r2o_brick_majorant(C.byref(self.b), bx, by, bz, mip)

    def density(self, ix, iy, iz):
        return lib().r2o_brick_density(C.byref(self.b), ix, iy, iz)

    def __del__(self):
        try:
            lib().r2o_brick_free(C.byref(self.b))
        except Exception:
            pass


class Scene:
    """Oracle scene: grids (records) + instances + env map, FP32 or FP64."""

    def __init__(self, grids, instances, env_rgb: np.ndarray, precision: str = "f64", surfaces=()):
        assert precision in ("f32", "f64")
        assert len(surfaces) <= MAX_SURF
        self.sfx = precision
        self.R = np.float32 if precision == "f32" else np.float64
        self._keep = []
        d = SceneDesc()
        d.n_grids = len(grids)
        for i, g in enumerate(grids):
            rec = np.ascontiguousarray(g["records"])
            self._keep.append(rec)
            gd = d.grids[i]
            gd.nz, gd.ny, gd.nx = g["shape"]
            gd.records = _p(rec)
            gd.map_mat[:] = list(g["map_mat"])
            gd.map_vec[:] = list(g["map_vec"])
            gd.sigma_a[:] = list(g["sigma_a"])
            gd.sigma_s[:] = list(g["sigma_s"])
            gd.voxel_size_filtering = g["voxel_size_filtering"]
        assert len(grids) <= MAX_GRIDS
        d.n_instances = len(instances)
        iarr = (InstanceDesc * max(1, len(instances)))()
        self._keep.append(iarr)
        d.instances = iarr
        for k, inst in enumerate(instances):
            idd = iarr[k]
            idd.grid = inst.get("grid", 0)
            idd.pos[:] = list(inst.get("pos", (0, 0, 0)))
            idd.orientation[:] = list(inst.get("orientation", (0, 1, 0)))
            idd.scaling[:] = list(inst.get("scaling", (1, 1, 1)))
        env = np.ascontiguousarray(env_rgb, dtype=np.float32)
        self._keep.append(env)
        d.env_h, d.env_w = env.shape[0], env.shape[1]
        d.env_rgb = _p(env)
        d.n_surfaces = len(surfaces)
        for j, sf in enumerate(surfaces):
            d.surfaces[j] = surface_desc(sf)
        self.desc = d
        self.h = self._fn("r2o_scene_new")(C.byref(d))

    def _fn(self, name):
        return getattr(lib(), f"{name}_{self.sfx}")

    def __del__(self):
        try:
            self._fn("r2o_scene_free")(self.h)
        except Exception:
            pass

    @staticmethod
    def params(cfg: dict, threads: int = 0) -> RenderParams:
        p = RenderParams()
        p.width, p.height = cfg["width"], cfg["height"]
        p.spp, p.max_depth = cfg["spp"], cfg["max_depth"]
        p.cam_pos[:] = list(cfg["cam_pos"])
        p.cam_look_at[:] = list(cfg["cam_look_at"])
        p.cam_up[:] = list(cfg["cam_up"])
        p.fov_deg = cfg["fov_deg"]
        p.seed, p.stream = cfg.get("seed", 1), cfg.get("stream", 0)
        p.use_brick_grid = int(cfg.get("use_brick_grid", True))
        p.sample_begin = cfg.get("sample_begin", 0)
        p.threads = threads
        return p

    def render(self, cfg: dict, threads: int = 0, counters: bool = False):
        p = self.params(cfg, threads)
        out = np.zeros((p.height, p.width, 3), self.R)
        cnt = np.zeros(8, np.uint64) if counters else None
        self._fn("r2o_render")(self.h, C.byref(p), _p(out), _p(cnt) if counters else None)
        return (out, cnt) if counters else out

    def render_paths(self, cfg: dict, pix: np.ndarray, smp: np.ndarray, threads: int = 0, counters: bool = False):
        p = self.params(cfg, threads)
        pix = np.ascontiguousarray(pix, np.int32)
        smp = np.ascontiguousarray(smp, np.int32)
        n = len(pix)
        out = np.zeros((n, 3), self.R)
        cnt = np.zeros((n, NCNT), np.uint64) if counters else None
        self._fn("r2o_render_paths")(self.h, C.byref(p), n, _p(pix), _p(smp), _p(out), _p(cnt) if counters else None)
        return (out, cnt) if counters else out

    def info(self, medium: int = 0) -> dict:
        out = np.zeros(64, self.R)
        self._fn("r2o_scene_info")(self.h, medium, _p(out))
        m = {}
        names = ["index_to_model", "model_to_index", "model_to_scaled", "scaled_to_model"]
        for q, nme in enumerate(names):
            m[nme] = out[12 * q:12 * q + 12].copy()
        m["ibb_min"] = out[48:51].copy()
        m["size"] = out[51:54].copy()
        m["inv_max_density"] = out[54]
        m["bb_min"] = out[55:58].copy()
        m["bb_max"] = out[58:61].copy()
        m["ratio"] = out[61]
        return m

    def instance_info(self, k: int = 0) -> dict:
        """Instance k: model_to_world / world_to_model as 3x4 [R | t] matrices,
        world bbox and filterRenderRatio."""
        out = np.zeros(31, self.R)
        self._fn("r2o_instance_info")(self.h, k, _p(out))
        m2w = out[0:12].reshape(4, 3).T.astype(np.float64)  # column major -> rows
        w2m = out[12:24].reshape(4, 3).T.astype(np.float64)
        return {"model_to_world": m2w, "world_to_model": w2m, "bb_min": out[24:27].copy(),
                "bb_max": out[27:30].copy(), "ratio": float(out[30])}

    def camera_view_to_world(self, cfg: dict) -> np.ndarray:
        p = self.params(cfg)
        out = np.zeros(12, self.R)
        self._fn("r2o_camera_view_to_world")(C.byref(p), _p(out))
        return out

    def filtered_data(self, p3, medium: int = 0) -> np.ndarray:
        p = np.ascontiguousarray(p3, self.R)
        out = np.zeros(17, self.R)
        self._fn("r2o_filtered_data")(self.h, medium, _p(p), _p(out))
        return out

    def env_sample(self, key: int, c0: int) -> np.ndarray:
        out = np.zeros(7, self.R)
        self._fn("r2o_env_sample")(self.h, key, c0, _p(out))
        return out

    def env_le(self, d) -> np.ndarray:
        dd = np.ascontiguousarray(d, self.R)
        out = np.zeros(4, self.R)
        self._fn("r2o_env_le")(self.h, _p(dd), _p(out))
        return out

    def env_tables(self):
        w, h = self.desc.env_w, self.desc.env_h
        marg = np.zeros(h + 1, self.R)
        rows = np.zeros((h, w + 1), self.R)
        ints = np.zeros(h + 1, self.R)
        self._fn("r2o_env_tables")(self.h, _p(marg), _p(rows), _p(ints))
        return marg, rows, ints

    def direct_mean(self, inst, pos, S6, n, ior, wo, use_dda: bool, key: int, c0: int, count: int):
        """(mean, variance) per channel of direct() at a medium interaction."""
        a = [np.ascontiguousarray(x, self.R) for x in (pos, S6, n, wo)]
        mean, var = np.zeros(3), np.zeros(3)
        self._fn("r2o_direct_mean")(self.h, inst, _p(a[0]), _p(a[1]), _p(a[2]), ior, _p(a[3]), int(use_dda), key, c0,
                                    count, _p(mean), _p(var))
        return mean, var

    def tr_segment(self, ray7, use_dda: bool, key: int, c0: int) -> float:
        r = np.ascontiguousarray(ray7, self.R)
        return self._fn("r2o_tr_segment")(self.h, _p(r), int(use_dda), key, c0)

    def sample_segment(self, ray7, use_dda: bool, key: int, c0: int):
        r = np.ascontiguousarray(ray7, self.R)
        t = np.zeros(1, self.R)
        hit = self._fn("r2o_sample_segment")(self.h, _p(r), int(use_dda), key, c0, _p(t))
        return hit, float(t[0])

    def move_to_voxel_border(self, ray7) -> np.ndarray:
        r = np.ascontiguousarray(ray7, self.R)
        out = np.zeros(7, self.R)
        self._fn("r2o_move_to_voxel_border")(self.h, _p(r), _p(out))
        return out

    def intersect(self, ray7):
        """(hit, [t, pos xyz, instance, surface]) of the scene intersection."""
        r = np.ascontiguousarray(ray7, self.R)
        out = np.zeros(6, self.R)
        hit = self._fn("r2o_intersect")(self.h, _p(r), _p(out))
        return hit, out

    def dda_cells(self, ray7, sigma_t: float = 1.0, max_cells: int = 4096, medium: int = 0,
                  collide_every: int = 0) -> np.ndarray:
        """(brick x, y, z, mip) of every majorant lookup of a DDA walk; ray7 = origin,
        direction, tmax in voxel space."""
        r = np.ascontiguousarray(ray7, self.R)
        cells = np.zeros((max_cells, 4), np.int32)
        n = self._fn("r2o_dda_cells")(self.h, medium, _p(r), sigma_t, collide_every, max_cells, _p(cells))
        return cells[:n].copy()


class Surface:
    """One analytic surface with its material, for the BSDF and shape probes."""

    def __init__(self, precision: str = "f64", **sf):
        self.sfx = precision
        self.R = np.float32 if precision == "f32" else np.float64
        self.desc = surface_desc(sf)

    def _fn(self, name):
        return getattr(lib(), f"{name}_{self.sfx}")

    def _a(self, x):
        return np.ascontiguousarray(x, self.R)

    def eval(self, wo, wi) -> dict:
        """Material f (rgb) and mixture pdf, both lobes' f and pdf, Fresnel weight."""
        out = np.zeros(9, self.R)
        o, i = self._a(wo), self._a(wi)
        self._fn("r2o_bsdf_eval")(C.byref(self.desc), _p(o), _p(i), _p(out))
        return {"f": out[0:3].copy(), "pdf": out[3], "on_f": out[4], "on_pdf": out[5], "mf_f": out[6],
                "mf_pdf": out[7], "weight": out[8]}

    def ggx(self, wo, wi, h) -> dict:
        out = np.zeros(4, self.R)
        o, i, hh = self._a(wo), self._a(wi), self._a(h)
        self._fn("r2o_ggx_eval")(C.byref(self.desc), _p(o), _p(i), _p(hh), _p(out))
        return {"D": out[0], "G1": out[1], "G": out[2], "pdf": out[3]}

    def sample_h(self, wo, u1: float, u2: float) -> np.ndarray:
        out = np.zeros(3, self.R)
        o = self._a(wo)
        self._fn("r2o_ggx_sample_h")(C.byref(self.desc), _p(o), u1, u2, _p(out))
        return out

    def sample(self, wo, key: int, c0: int):
        """(wi, f rgb, pdf) of CMaterial::sampleF on stream (key, c0)."""
        out = np.zeros(7, self.R)
        o = self._a(wo)
        self._fn("r2o_bsdf_sample")(C.byref(self.desc), _p(o), key, c0, _p(out))
        return out[0:3].copy(), out[3:6].copy(), out[6]

    def hit(self, ray7):
        """(hit, t, normal) of the shape test against ray (o, d, tmax)."""
        r = self._a(ray7)
        out = np.zeros(4, self.R)
        ok = self._fn("r2o_surface_hit")(C.byref(self.desc), _p(r), _p(out))
        return bool(ok), float(out[0]), out[1:4].copy()


def fresnel_ie(cos_i: float, eta_i: float, eta_t: float, precision: str = "f64") -> float:
    return fn("r2o_fresnel_ie", precision)(cos_i, eta_i, eta_t)


def fn(name: str, precision: str = "f64"):
    return getattr(lib(), f"{name}_{precision}")


def as_arr(x, precision="f64"):
    return np.ascontiguousarray(x, np.float32 if precision == "f32" else np.float64)


class MeshFilter:
    """CPU restatement of CMeshFilter over one mesh (BVH built once).

    mesh: dict with positions (V,3) f32, normals (V,3) f32, indices (T,3) i32,
    material (T,) i32 or None, mat_diffuse (M,3), mat_specular (M,3),
    mat_roughness (M,), mat_ior (M,), pos/orientation/scaling (3,); optional
    uvs (V,2), textures (list of (H,W,4) RGBA float32) and mat_textures (M,4).
    """

    def __init__(self, mesh: dict):
        self._keep = {}
        for k in ("positions", "normals", "mat_diffuse", "mat_specular", "mat_roughness", "mat_ior"):
            self._keep[k] = np.ascontiguousarray(mesh[k], np.float32)
        self._keep["indices"] = np.ascontiguousarray(mesh["indices"], np.int32)
        mat = mesh.get("material")
        self._keep["material"] = None if mat is None else np.ascontiguousarray(mat, np.int32)
        d = MeshDesc()
        d.n_vertices = self._keep["positions"].shape[0]
        d.n_triangles = self._keep["indices"].shape[0]
        for k in ("positions", "normals", "indices", "mat_diffuse", "mat_specular", "mat_roughness", "mat_ior"):
            setattr(d, k, _p(self._keep[k]))
        d.material = None if self._keep["material"] is None else _p(self._keep["material"])
        d.n_materials = self._keep["mat_ior"].shape[0]
        for k, dflt in (("pos", (0, 0, 0)), ("orientation", (0, 1, 0)), ("scaling", (1, 1, 1))):
            setattr(d, k, (C.c_float * 3)(*[float(v) for v in mesh.get(k, dflt)]))
        d.uvs, d.n_textures, d.textures, d.mat_textures = mesh_textures(mesh, self._keep, Texture)
        self.desc = d
        self._h = lib().r2o_filter_new(C.byref(d))
        if not self._h:
            raise ValueError("empty mesh")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().r2o_filter_free(self._h)
            self._h = None

    def setup(self, voxel_size: float) -> FilterGrid:
        g = FilterGrid()
        if lib().r2o_filter_setup(C.byref(self.desc), C.c_float(voxel_size), C.byref(g)) != 0:
            raise ValueError("r2o_filter_setup failed")
        return g

    def run(self, grid: FilterGrid, samples: int = 64, iterations: int = 8, sigma_t: float = 1.0,
            alpha: float = 1.0, clip_rays: bool = True, seed: int = 0, lod: int = 0, ids=None,
            threads: int = 0) -> np.ndarray:
        """32-byte records (n, 32) u8 for `ids` (voxel ids x + y*nx + z*nx*ny) or every voxel."""
        p = FilterParams(samples, iterations, int(bool(clip_rays)), lod, threads, sigma_t, alpha, seed)
        if ids is None:
            n = int(grid.nv[0]) * int(grid.nv[1]) * int(grid.nv[2])
            idp = None
        else:
            ids = np.ascontiguousarray(ids, np.int64)
            n = ids.shape[0]
            idp = _p(ids)
        out = np.zeros((n, 32), np.uint8)
        if lib().r2o_filter_run(self._h, C.byref(grid), C.byref(p), n, idp, _p(out)) != 0:
            raise ValueError("r2o_filter_run failed")
        return out

    def probe(self, grid: FilterGrid, vid: int, samples: int = 64, iterations: int = 8, sigma_t: float = 1.0,
              alpha: float = 1.0, clip_rays: bool = True, seed: int = 0, lod: int = 0) -> dict:
        """Per-ray view of one voxel: origins, directions, ray tmax, hit triangle (-1: miss), hit t."""
        p = FilterParams(samples, iterations, int(bool(clip_rays)), lod, 1, sigma_t, alpha, seed)
        rec = np.zeros(32, np.uint8)
        o = np.zeros((samples, 3), np.float32)
        d = np.zeros((samples, 3), np.float32)
        tm = np.zeros(samples, np.float32)
        tri = np.zeros(samples, np.int32)
        t = np.zeros(samples, np.float32)
        hits = lib().r2o_filter_probe(self._h, C.byref(grid), C.byref(p), int(vid), _p(rec), _p(o), _p(d), _p(tm),
                                      _p(tri), _p(t))
        if hits < 0:
            raise ValueError("r2o_filter_probe failed")
        return {"hits": hits, "record": rec, "o": o, "d": d, "tmax": tm, "tri": tri, "t": t}

    def closest_hit(self, o, d, tmax: float):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        out = np.zeros(3, np.float32)
        tri = lib().r2o_filter_closest_hit(self._h, _p(o), _p(d), C.c_float(tmax), _p(out))
        return tri, out


def tex_fetch(rgba: np.ndarray, x: float, y: float) -> np.ndarray:
    """The filter's texture fetch of (H, W, 4) RGBA float32 texels at (x, y)."""
    t = np.ascontiguousarray(rgba, np.float32)
    desc = Texture(t.shape[1], t.shape[0], t.ctypes.data)
    out = np.zeros(4, np.float32)
    lib().r2o_tex_fetch(C.byref(desc), C.c_float(x), C.c_float(y), _p(out))
    return out


def grid_dict(g: FilterGrid) -> dict:
    return {
        "nv": tuple(g.nv),
        "index_to_model": np.array(g.index_to_model, np.float32),
        "model_to_index": np.array(g.model_to_index, np.float32),
        "model_to_world": np.array(g.model_to_world, np.float32),
        "world_to_model": np.array(g.world_to_model, np.float32),
        "world_bb_min": np.array(g.world_bb_min, np.float32),
        "world_bb_max": np.array(g.world_bb_max, np.float32),
    }


def decode_records(rec: np.ndarray) -> dict:
    """Field view of (n, 32) SFilteredDataCompact records."""
    rec = np.ascontiguousarray(rec).reshape(-1, 32)
    u16 = rec.view(np.uint16)
    return {
        "density": rec[:, 0:4].copy().view(np.float32)[:, 0],
        "sigma": u16[:, 2:5],
        "r": u16[:, 5:8],
        "diffuse": rec[:, 16:19],
        "specular": rec[:, 19:22],
        "normal": u16[:, 11:14],
        "ior": u16[:, 14].view(np.float16).astype(np.float32),
    }


def postprocess(hdr: np.ndarray, gamma: float = 2.0):
    """Tonemapped + gamma-corrected frame (float) and its bytes, like SFrame."""
    hdr = np.ascontiguousarray(hdr, np.float32)
    h, w = hdr.shape[:2]
    ldr = np.zeros_like(hdr)
    by = np.zeros(hdr.shape, np.uint8)
    lib().r2o_postprocess(_p(hdr), w, h, C.c_float(gamma), _p(ldr), _p(by))
    return ldr, by
