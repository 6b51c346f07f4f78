"""MI355X-native volumetric hot path of FrederikBoehm/Renderer2.

The product is the native library renderer2_amd/lib/librenderer2.so (HIP
kernels for gfx950 behind the C ABI of include/r2.h).  This package is a thin
Python binding over that ABI for tests and benchmarks; it has no compute of its
own and no fallback when the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from ._capi import R2Error

__all__ = ["Context", "Grid", "Env", "Mesh", "Server", "R2Error", "capi", "filter_params", "image_write", "nvdb_write",
           "nvdb_read"]


def filter_params(voxel_size: float, samples: int = 64, iterations: int = 8, sigma_t: float = 1.0,
                  alpha: float = 1.0, clip_rays: bool = True, seed: int = 0, lod: int = 0) -> "capi.FilterParams":
    return capi.FilterParams(voxel_size, samples, iterations, sigma_t, alpha, int(bool(clip_rays)), seed, lod)


def image_write(path, rgb: np.ndarray, fmt: str = "png", quality: int = 100) -> None:
    """Write an (h, w, 3) u8 frame, rows bottom to top as r2_postprocess gives them, as
    the reference's output.png (fmt "png") or output.jpg (fmt "jpg", quality 1..100);
    the file is flipped to top-down.  Host only."""
    img = np.ascontiguousarray(rgb, np.uint8)
    if img.ndim != 3 or img.shape[2] != 3:
        raise ValueError("rgb must be (h, w, 3) uint8")
    f = {"png": capi.R2_IMAGE_PNG, "jpg": capi.R2_IMAGE_JPG}.get(fmt)
    if f is None:
        raise ValueError("fmt must be 'png' or 'jpg'")
    err = C.create_string_buffer(512)
    rc = capi.lib().r2_image_write(str(path).encode(), img.shape[1], img.shape[0], img.ctypes.data, f, int(quality),
                                   err, 512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))


def nvdb_write(path, records: np.ndarray, map_mat=None, map_vec=None) -> None:
    """NanoVDB file (one Vec4d grid) of (nz, ny, nx, 32) u8 records: the voxels with
    density > 0 and the 8 grid corners are active; map_mat (3x3, row major) and map_vec
    are the index -> model map (default: the identity).  Host only."""
    rec = np.ascontiguousarray(records, np.uint8)
    if rec.ndim != 4 or rec.shape[3] != 32:
        raise ValueError("records must be (nz, ny, nx, 32) uint8")
    nz, ny, nx = rec.shape[:3]
    mm = None if map_mat is None else np.ascontiguousarray(np.asarray(map_mat, np.float32).reshape(9))
    mv = None if map_vec is None else np.ascontiguousarray(np.asarray(map_vec, np.float32).reshape(3))
    err = C.create_string_buffer(512)
    rc = capi.lib().r2_nvdb_write(str(path).encode(), nx, ny, nz, rec.ctypes.data,
                                  None if mm is None else mm.ctypes.data, None if mv is None else mv.ctypes.data,
                                  err, 512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))


def nvdb_read(path, with_records: bool = True):
    """(info dict, (nz, ny, nx, 32) u8 records over the index box or None).  Host only."""
    info = capi.NvdbInfo()
    err = C.create_string_buffer(512)
    rc = capi.lib().r2_nvdb_read(str(path).encode(), C.byref(info), None, 0, err, 512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))
    d = {k: (np.array(getattr(info, k)) if not isinstance(getattr(info, k), (int, float)) else getattr(info, k))
         for k, _ in info._fields_}
    if not with_records:
        return d, None
    rec = np.empty((info.nz, info.ny, info.nx, 32), np.uint8)
    rc = capi.lib().r2_nvdb_read(str(path).encode(), C.byref(info), rec.ctypes.data, rec.nbytes, err, 512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))
    return d, rec


def brickgrid_read(path):
    """(info dict, level-0 ranges, atlas densities in brick layout, [mip 1, 2, 3]) of a .brickgrid
    file, in the layout of Grid.download_bricks.  Host only."""
    info = capi.BrickgridInfo()
    err = C.create_string_buffer(512)
    rc = capi.lib().r2_brickgrid_read(str(path).encode(), C.byref(info), None, None, None, err, 512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))
    nbx, nby, nbz = info.n_bricks
    rng = np.zeros(nbx * nby * nbz, np.uint32)
    u8 = np.zeros(nbx * nby * nbz * 512, np.uint8)
    cnt = [(nbx >> (l + 1)) * (nby >> (l + 1)) * (nbz >> (l + 1)) for l in range(3)]
    mips = np.zeros(sum(cnt), np.uint32)
    rc = capi.lib().r2_brickgrid_read(str(path).encode(), C.byref(info), rng.ctypes.data, u8.ctypes.data,
                                      mips.ctypes.data, err, 512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))
    d = dict(n_bricks=tuple(info.n_bricks), min_density=info.min_density, max_density=info.max_density,
             n_nonempty_bricks=info.n_nonempty_bricks)
    return d, rng, u8, [mips[sum(cnt[:l]):sum(cnt[:l + 1])] for l in range(3)]


def nvdb_read_leaves(path):
    """Sparse read of a NanoVDB file: (info dict with origin_shift, leaf_origins (n, 3) int32,
    records (n, 512, 32) u8 x fastest, leaf_masks (n, 64) u8).  Host only."""
    info = capi.NvdbInfo()
    err = C.create_string_buffer(512)
    shift = (C.c_int32 * 3)()
    n = C.c_int64(0)
    rc = capi.lib().r2_nvdb_read_leaves(str(path).encode(), C.byref(info), shift, C.byref(n), None, None, None, err,
                                        512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))
    org = np.empty((n.value, 3), np.int32)
    rec = np.empty((n.value, 512, 32), np.uint8)
    msk = np.empty((n.value, 64), np.uint8)
    rc = capi.lib().r2_nvdb_read_leaves(str(path).encode(), C.byref(info), shift, C.byref(n), org.ctypes.data,
                                        rec.ctypes.data, msk.ctypes.data, err, 512)
    if rc != 0:
        raise R2Error(rc, err.value.decode(errors="replace"))
    d = {k: (np.array(getattr(info, k)) if not isinstance(getattr(info, k), (int, float)) else getattr(info, k))
         for k, _ in info._fields_}
    d["origin_shift"] = np.array(shift[:], np.int32)
    return d, org, rec, msk


class Context:
    """Owns an r2_ctx (one device, one HIP stream)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        capi.check(capi.lib().r2_create(C.byref(self.h), device))
        self._grids = []

    def close(self):
        if self.h:
            capi.lib().r2_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        return capi.check(rc, self.h)

    # ---------------------------------------------------------------- grid
    def upload_grid(self, g: dict) -> "Grid":
        """g: records (host array) or records_dev (device address), shape (nz, ny, nx), map, optics.
        Sparse form: leaf_origins (n, 3) int32 (x, y, z, multiples of 8) with records (n, 512, 32)
        u8 and optionally leaf_masks (n, 64) u8 of active voxels; shape is then the index extent."""
        d = capi.GridDesc()
        d.nz, d.ny, d.nx = g["shape"]
        keep = []
        if g.get("records_dev"):
            d.records = g["records_dev"]
            d.records_on_device = 1
        else:
            rec = np.ascontiguousarray(g["records"])
            keep.append(rec)
            d.records = rec.ctypes.data
            d.records_on_device = 0
        if g.get("leaf_origins") is not None:
            org = np.ascontiguousarray(g["leaf_origins"], np.int32).reshape(-1, 3)
            keep.append(org)
            d.n_leaves = len(org)
            d.leaf_origins = org.ctypes.data
            if g.get("leaf_masks") is not None:
                msk = np.ascontiguousarray(g["leaf_masks"], np.uint8).reshape(-1, 64)
                assert len(msk) == len(org)
                keep.append(msk)
                d.leaf_masks = msk.ctypes.data
        d.map_mat[:] = list(g["map_mat"])
        d.map_vec[:] = list(g["map_vec"])
        d.sigma_a[:] = list(g["sigma_a"])
        d.sigma_s[:] = list(g["sigma_s"])
        d.voxel_size_filtering = g["voxel_size_filtering"]
        d.precision = capi.R2_PREC_PACKED if g.get("precision") == "packed" else capi.R2_PREC_FP32
        h = C.c_void_p()
        self._check(capi.lib().r2_grid_upload(self.h, C.byref(d), C.byref(h)))
        return Grid(self, h)

    def upload_env(self, rgb: np.ndarray) -> "Env":
        rgb = np.ascontiguousarray(rgb, np.float32)
        h = C.c_void_p()
        self._check(capi.lib().r2_env_upload(self.h, rgb.ctypes.data, rgb.shape[1], rgb.shape[0], C.byref(h)))
        return Env(self, h)

    # -------------------------------------------------------------- render
    @staticmethod
    def render_params(cfg: dict, grids, instances, env, flags: int = 0):
        p = capi.RenderParams()
        p.width, p.height = cfg["width"], cfg["height"]
        p.spp, p.max_depth = cfg["spp"], cfg["max_depth"]
        p.cam_pos[:] = list(cfg["cam_pos"])
        p.cam_look_at[:] = list(cfg["cam_look_at"])
        p.cam_up[:] = list(cfg["cam_up"])
        p.fov_deg = cfg["fov_deg"]
        p.seed, p.stream = cfg.get("seed", 1), cfg.get("stream", 0)
        arr = (capi.Instance * max(1, len(instances)))()
        for k, inst in enumerate(instances):
            arr[k].grid = grids[inst.get("grid", 0)].h
            arr[k].pos[:] = list(inst.get("pos", (0, 0, 0)))
            arr[k].orientation[:] = list(inst.get("orientation", (0, 1, 0)))
            arr[k].scaling[:] = list(inst.get("scaling", (1, 1, 1)))
        p.n_instances = len(instances)
        p.instances = arr
        surfs = cfg.get("surfaces", ())
        sarr = (capi.Surface * max(1, len(surfs)))()
        for j, sf in enumerate(surfs):
            sarr[j].shape = {"circle": capi.R2_SHAPE_CIRCLE, "sphere": capi.R2_SHAPE_SPHERE}[sf.get("shape", "circle")]
            sarr[j].pos[:] = list(sf.get("pos", (0, 0, 0)))
            sarr[j].radius = sf.get("radius", 1.0)
            sarr[j].normal[:] = list(sf.get("normal", (0, 1, 0)))
            sarr[j].diffuse[:] = list(sf.get("diffuse", (0.5, 0.5, 0.5)))
            sarr[j].diffuse_roughness = sf.get("diffuse_roughness", 0.0)
            sarr[j].specular[:] = list(sf.get("specular", (0, 0, 0)))
            sarr[j].alpha_x = sf.get("alpha_x", 0.5)
            sarr[j].alpha_y = sf.get("alpha_y", 0.5)
            sarr[j].eta_i = sf.get("eta_i", 1.0)
            sarr[j].eta_t = sf.get("eta_t", 1.5)
        p.n_surfaces = len(surfs)
        p.surfaces = sarr
        p.env = env.h
        f = flags
        if cfg.get("use_brick_grid", True):
            f |= capi.R2_FLAG_BRICK_GRID
        if cfg.get("fp64_accum", False):
            f |= capi.R2_FLAG_FP64_ACCUM
        p.flags = f
        p.sample_begin = cfg.get("sample_begin", 0)
        return p, (arr, sarr)

    def render(self, cfg: dict, grids, instances, env) -> np.ndarray:
        p, keep = self.render_params(cfg, grids, instances, env)
        out = np.zeros((p.height, p.width, 3), np.float32)
        self._check(capi.lib().r2_render(self.h, C.byref(p), out.ctypes.data, 0))
        return out

    def render_progressive(self, cfg: dict, grids, instances, env, callback):
        """Render with callback(samples_done, samples_total) -> bool after each sample
        chunk; True stops the frame.  Returns (frame, completed): a stopped frame holds
        the sum of L / spp over the completed samples (renderFrame's keyCallback)."""
        p, keep = self.render_params(cfg, grids, instances, env)
        out = np.zeros((p.height, p.width, 3), np.float32)
        fn = capi.PROGRESS_FN(lambda _u, done, total: 1 if callback(int(done), int(total)) else 0)
        rc = capi.lib().r2_render_progressive(self.h, C.byref(p), out.ctypes.data, 0, C.cast(fn, C.c_void_p), None)
        if rc == capi.R2_ERR_CANCELLED:
            return out, False
        self._check(rc)
        return out, True

    def render_device(self, cfg: dict, grids, instances, env, out_dev_ptr: int):
        """Render into a caller-owned device buffer (inputs and output stay in HBM)."""
        p, keep = self.render_params(cfg, grids, instances, env)
        self._check(capi.lib().r2_render(self.h, C.byref(p), out_dev_ptr, 1))

    def render_noalloc(self, cfg: dict, grids, instances, env, out_dev_ptr: int):
        """r2_render_noalloc: render into a caller-owned device buffer with no device
        allocation (chunk sums in the library's fixed 64 MiB per-GPU array)."""
        p, keep = self.render_params(cfg, grids, instances, env)
        self._check(capi.lib().r2_render_noalloc(self.h, C.byref(p), out_dev_ptr))

    def render_paths(self, cfg: dict, grids, instances, env, pix, smp, counters: bool = False):
        p, keep = self.render_params(cfg, grids, instances, env)
        pix = np.ascontiguousarray(pix, np.int32)
        smp = np.ascontiguousarray(smp, np.int32)
        n = len(pix)
        out = np.zeros((n, 3), np.float32)
        cnt = np.zeros((n, 8), np.uint32) if counters else None
        self._check(capi.lib().r2_render_paths(self.h, C.byref(p), n, pix.ctypes.data, smp.ctypes.data,
                                               out.ctypes.data, cnt.ctypes.data if counters else None))
        return (out, cnt[:, :7]) if counters else out

    def scene_intersect(self, cfg: dict, grids, instances, env, rays7) -> np.ndarray:
        """r2_scene_intersect (parity probe): n x (hit, t, instance, surface) of rays
        (origin, direction, tmax) against the scene's instance boxes and surfaces."""
        p, keep = self.render_params(cfg, grids, instances, env)
        r = np.ascontiguousarray(rays7, np.float32)
        out = np.zeros((len(r), 4), np.float32)
        self._check(capi.lib().r2_scene_intersect(self.h, C.byref(p), len(r), r.ctypes.data, out.ctypes.data))
        return out

    def dda_cells(self, rays7, sigma_t: float = 1.0, collide_every: int = 0, max_cells: int = 4096):
        """Majorant-DDA walks (parity probe): for each ray (origin, direction, tmax in
        voxel space) the (brick x, y, z, mip) of every majorant lookup."""
        rays7 = np.ascontiguousarray(rays7, np.float32).reshape(-1, 7)
        n = rays7.shape[0]
        cells = np.zeros((n, max_cells, 4), np.int32)
        counts = np.zeros(n, np.int32)
        self._check(capi.lib().r2_dda_cells(self.h, n, rays7.ctypes.data, C.c_float(sigma_t), collide_every,
                                            max_cells, cells.ctypes.data, counts.ctypes.data))
        return [cells[i, :counts[i]].copy() for i in range(n)]

    # -------------------------------------------------------------- filter
    def upload_mesh(self, mesh: dict) -> "Mesh":
        keep = {}
        for k in ("positions", "normals", "mat_diffuse", "mat_specular", "mat_roughness", "mat_ior"):
            keep[k] = np.ascontiguousarray(mesh[k], np.float32)
        keep["indices"] = np.ascontiguousarray(mesh["indices"], np.int32)
        mat = mesh.get("material")
        keep["material"] = None if mat is None else np.ascontiguousarray(mat, np.int32)
        d = capi.MeshDesc()
        d.n_vertices = keep["positions"].shape[0]
        d.n_triangles = keep["indices"].shape[0]
        for k in ("positions", "normals", "indices", "mat_diffuse", "mat_specular", "mat_roughness", "mat_ior"):
            setattr(d, k, keep[k].ctypes.data)
        d.material = None if keep["material"] is None else keep["material"].ctypes.data
        d.n_materials = keep["mat_ior"].shape[0]
        d.pos[:] = [float(v) for v in mesh.get("pos", (0, 0, 0))]
        d.orientation[:] = [float(v) for v in mesh.get("orientation", (0, 1, 0))]
        d.scaling[:] = [float(v) for v in mesh.get("scaling", (1, 1, 1))]
        # textured materials: uvs (V,2), textures [(H,W,4) RGBA float32], mat_textures (M,4)
        uv = mesh.get("uvs")
        keep["uvs"] = None if uv is None else np.ascontiguousarray(uv, np.float32)
        keep["textures"] = [np.ascontiguousarray(t, np.float32) for t in mesh.get("textures", [])]
        tex = (capi.TextureDesc * max(1, len(keep["textures"])))()
        for i, t in enumerate(keep["textures"]):
            if t.ndim != 3 or t.shape[2] != 4:
                raise ValueError("textures are (height, width, 4) RGBA arrays")
            tex[i].height, tex[i].width, tex[i].rgba = t.shape[0], t.shape[1], t.ctypes.data
        mt = mesh.get("mat_textures")
        keep["mat_textures"] = None if mt is None else np.ascontiguousarray(mt, np.int32).reshape(-1, 4)
        d.uvs = None if keep["uvs"] is None else keep["uvs"].ctypes.data
        d.n_textures = len(keep["textures"])
        d.textures = C.cast(tex, C.c_void_p).value if keep["textures"] else None
        d.mat_textures = None if keep["mat_textures"] is None else keep["mat_textures"].ctypes.data
        h = C.c_void_p()
        self._check(capi.lib().r2_mesh_upload(self.h, C.byref(d), C.byref(h)))
        return Mesh(self, h)

    def filter_lod(self, mesh: "Mesh", params, out_dev_ptr: int | None = None):
        """Filter one LOD.  Returns (records (nz, ny, nx, 32) u8 or None when written
        to device memory, r2_filter_grid)."""
        g = capi.FilterGrid()
        capi.check(capi.lib().r2_filter_setup(mesh.h, C.c_float(params.voxel_size), C.byref(g)), self.h)
        nx, ny, nz = list(g.num_voxels)
        if out_dev_ptr is None:
            out = np.zeros((nz, ny, nx, 32), np.uint8)
            self._check(capi.lib().r2_filter_lod(self.h, mesh.h, C.byref(params), out.ctypes.data, 0, C.byref(g)))
            return out, g
        self._check(capi.lib().r2_filter_lod(self.h, mesh.h, C.byref(params), C.c_void_p(out_dev_ptr), 1,
                                             C.byref(g)))
        return None, g

    def filter_voxels(self, mesh: "Mesh", params, ids) -> np.ndarray:
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros((ids.shape[0], 32), np.uint8)
        self._check(capi.lib().r2_filter_voxels(self.h, mesh.h, C.byref(params), ids.shape[0], ids.ctypes.data,
                                                out.ctypes.data))
        return out

    def filter_stats(self) -> dict:
        s = capi.FilterStats()
        self._check(capi.lib().r2_filter_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    # -------------------------------------------------------------- server
    def server(self, hbm_budget_bytes: int = 8 << 30, max_inflight: int = 16, slice_paths: int = 1 << 24) -> "Server":
        cfg = capi.ServerConfig(hbm_budget_bytes, max_inflight, slice_paths)
        h = C.c_void_p()
        self._check(capi.lib().r2_server_start(self.h, C.byref(cfg), C.byref(h)))
        return Server(self, h)

    def stats(self) -> dict:
        s = capi.Stats()
        self._check(capi.lib().r2_get_stats(self.h, C.byref(s)))
        return dict(kernel_ms=s.last_kernel_ms, total_ms=s.last_total_ms, paths=s.last_paths)

    def stream(self) -> int:
        return capi.lib().r2_stream(self.h) or 0


class Grid:
    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def free(self):
        if self.h:
            capi.lib().r2_grid_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self) -> dict:
        i = capi.GridInfo()
        capi.check(capi.lib().r2_grid_get_info(self.h, C.byref(i)))
        return {k: (np.array(getattr(i, k)) if not isinstance(getattr(i, k), (int, float)) else getattr(i, k))
                for k, _ in i._fields_}

    def write_brickgrid(self, path: str) -> None:
        """The reference's .brickgrid file of this grid's brick majorant grid."""
        err = C.create_string_buffer(512)
        rc = capi.lib().r2_grid_write_brickgrid(self.ctx.h, self.h, str(path).encode(), err, 512)
        if rc != 0:
            raise capi.R2Error(rc, err.value.decode(errors="replace"))

    def download_bricks(self, with_atlas: bool = True):
        """(level-0 ranges, the atlas as dense u8 densities (None without with_atlas), mips)."""
        inf = self.info()
        nbx, nby, nbz = [int(x) for x in inf["n_bricks"]]
        rng = np.zeros(nbx * nby * nbz, np.uint32)
        u8 = np.zeros(nbx * nby * nbz * 512, np.uint8) if with_atlas else None
        nm = sum((nbx >> (l + 1)) * (nby >> (l + 1)) * (nbz >> (l + 1)) for l in range(3))
        mips = np.zeros(nm, np.uint32)
        capi.check(capi.lib().r2_grid_download_bricks(self.ctx.h, self.h, rng.ctypes.data,
                                                      u8.ctypes.data if with_atlas else None, mips.ctypes.data),
                   self.ctx.h)
        out_m, off = [], 0
        for l in range(3):
            c = (nbx >> (l + 1)) * (nby >> (l + 1)) * (nbz >> (l + 1))
            out_m.append(mips[off:off + c])
            off += c
        return rng, u8, out_m

    def sample(self, pos: np.ndarray) -> np.ndarray:
        """Filtered (interpolated, decoded) records at normalised positions: (n, 17) floats."""
        pos = np.ascontiguousarray(pos, np.float32)
        n = pos.shape[0]
        out = np.zeros((n, 17), np.float32)
        capi.check(capi.lib().r2_grid_sample(self.ctx.h, self.h, n, pos.ctypes.data, out.ctypes.data), self.ctx.h)
        return out

    def lookup(self, coords: np.ndarray) -> np.ndarray:
        coords = np.ascontiguousarray(coords, np.int32)
        n = coords.shape[0]
        out = np.zeros((n, 32), np.uint8)
        capi.check(capi.lib().r2_grid_lookup(self.ctx.h, self.h, n, coords.ctypes.data, out.ctypes.data), self.ctx.h)
        return out


class Env:
    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def free(self):
        if self.h:
            capi.lib().r2_env_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Mesh:
    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def free(self):
        if self.h:
            capi.lib().r2_mesh_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def setup(self, voxel_size: float) -> "capi.FilterGrid":
        g = capi.FilterGrid()
        capi.check(capi.lib().r2_filter_setup(self.h, C.c_float(voxel_size), C.byref(g)))
        return g


class Server:
    """Render server on a context: submit() returns a request id, wait() the frame."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h
        self._req = {}

    def submit(self, cfg: dict, grids, instances, env, deadline_ms: float = 0.0) -> int:
        """deadline_ms > 0: admitted earliest deadline first, dropped (R2_ERR_DEADLINE)
        if unfinished that long after submission."""
        p, keep = Context.render_params(cfg, grids, instances, env)
        out = np.zeros((p.height, p.width, 3), np.float32)
        r = capi.Request(p, out.ctypes.data, float(deadline_ms))
        rid = capi.lib().r2_server_submit(self.h, C.byref(r))
        if rid < 0:
            raise capi.R2Error(int(rid), "r2_server_submit")
        self._req[rid] = (out, keep)
        return int(rid)

    def poll(self, rid: int):
        lat = C.c_double(0.0)
        rc = capi.lib().r2_server_poll(self.h, rid, C.byref(lat))
        if rc < 0:
            raise capi.R2Error(rc, f"request {rid}")
        return (self._req[rid][0], lat.value) if rc == 1 else None

    def wait(self, rid: int):
        lat = C.c_double(0.0)
        rc = capi.lib().r2_server_wait(self.h, rid, C.byref(lat))
        if rc < 0:
            raise capi.R2Error(rc, f"request {rid}")
        return self._req[rid][0], lat.value

    def cancel(self, rid: int):
        capi.check(capi.lib().r2_server_cancel(self.h, rid))

    def stats(self) -> dict:
        s = capi.ServerStats()
        capi.check(capi.lib().r2_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def stop(self):
        if self.h:
            capi.lib().r2_server_stop(self.h)
            self.h = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
