"""Front end and post-process on the GPU.

* r2_postprocess equals the oracle's restatement of the reference's tonemap /
  gamma / bytes kernels bit for bit.
* r2_pipeline_run with a config that filters a mesh (OBJ + MTL written here) to
  two LODs and renders the finer LOD: the grid files hold exactly the records of
  r2_filter_lod for the same mesh and parameters, the volume description lists
  both LODs, and output.png is the post-processed output.pfm, flipped (output.jpg
  the same image at quality 100).
"""
import ctypes as C
import json
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import renderer2_amd as r2
from renderer2_amd import capi, scenes

from brickgrid_io import check_against_device

pytestmark = pytest.mark.gpu


def test_postprocess_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    for (h, w) in ((37, 53), (128, 96)):
        hdr = (rng.lognormal(0.0, 2.0, (h, w, 3)) * (rng.random((h, w, 1)) > 0.05)).astype(np.float32)
        ldr = np.zeros_like(hdr)
        by = np.zeros(hdr.shape, np.uint8)
        capi.check(capi.lib().r2_postprocess(ctx.h, hdr.ctypes.data, w, h, C.c_float(2.2), ldr.ctypes.data,
                                             by.ctypes.data, 0), ctx.h)
        ref_ldr, ref_by = O.postprocess(hdr, 2.2)
        assert np.array_equal(ldr.view(np.uint32), ref_ldr.view(np.uint32))
        assert np.array_equal(by, ref_by)
        assert 0 < by.mean() < 255


def write_obj(path: Path, mesh: dict):
    with open(path.with_suffix(".mtl"), "w") as f:
        for k in range(len(mesh["mat_ior"])):
            d, s = mesh["mat_diffuse"][k], mesh["mat_specular"][k]
            # roughness = (2 / (Ns + 2))^0.5  =>  Ns = 2 / roughness^2 - 2
            ns = 2.0 / float(mesh["mat_roughness"][k]) ** 2 - 2.0
            f.write(f"newmtl m{k}\nKd {d[0]} {d[1]} {d[2]}\nKs {s[0]} {s[1]} {s[2]}\nNs {ns!r}\n"
                    f"Ni {float(mesh['mat_ior'][k])!r}\n")
    with open(path, "w") as f:
        f.write(f"mtllib {path.with_suffix('.mtl').name}\n")
        for p in mesh["positions"]:
            f.write(f"v {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")
        for n in mesh["normals"]:
            f.write(f"vn {float(n[0])!r} {float(n[1])!r} {float(n[2])!r}\n")
        cur = None
        for t, tri in enumerate(mesh["indices"]):
            m = int(mesh["material"][t])
            if m != cur:
                f.write(f"usemtl m{m}\n")
                cur = m
            a, b, c = (int(i) + 1 for i in tri)
            f.write(f"f {a}//{a} {b}//{b} {c}//{c}\n")


def read_png(path: Path):
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        assert zlib.crc32(typ + body) == struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        if typ == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 3)


def read_pfm(path: Path):
    data = path.read_bytes()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    assert float(parts[2]) < 0  # little endian
    return np.frombuffer(parts[3], np.float32).reshape(h, w, 3)


def test_pipeline_filters_and_renders(ctx, tmp_path):
    mesh = scenes.synth_mesh(10, 80, 0.08, seed=7)
    # roughness values that survive the Ns round trip of the MTL file exactly
    mesh["mat_roughness"] = np.array([0.5, 1.0], np.float32)
    obj = tmp_path / "assets" / "blob.obj"
    obj.parent.mkdir()
    write_obj(obj, mesh)
    vs = np.float32(scenes.finest_voxel_size(mesh, 24))
    vs_name = "%f" % vs
    cfg = {
        "General": {"FrameWidth": 96, "FrameHeight": 64, "ChannelsPerPixel": 3, "Samples": 8, "Gamma": 2.0,
                    "UseBrickGrid": True, "DebugMode": "NONE", "MaxDepth": 6, "Seed": 3},
        "Filtering": {"Active": True, "MinVoxelSize": float(vs), "LODs": 2, "SamplesPerVoxel": 32, "Debug": False,
                      "DebugSamples": 1, "SigmaT": 10.0, "EstimationIterations": 4, "Alpha": 0.1, "ClipRays": True},
        "Camera": {"Fov": 40.0, "Pos": [0.0, 0.5, 4.0], "LookAt": [0, 0, 0], "Up": [0, 1, 0]},
        "Scene": {"Envmap": "synth:sky", "Scenedescription": "scene.json"},
    }
    scene = [
        {"Type": "Mesh", "Directory": "assets", "Filename": "blob.obj", "Pos": [0, 0, 0], "Orientation": [0, 1, 0],
         "Scaling": [1, 1, 1], "Mask": ["FILTER"]},
        {"Type": "Medium", "Path": f"filtering/blob/filtered_mesh_{vs_name}", "SigmaA": [0.1, 0.1, 0.1],
         "SigmaS": [0.9, 0.9, 0.9], "Pos": [0, 0, 0], "Orientation": [0, 1, 0], "Scaling": [1, 1, 1]},
    ]
    (tmp_path / "config.json").write_text("// test config\n" + json.dumps(cfg))
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    rep = capi.PipelineReport()
    err = C.create_string_buffer(1024)
    rc = capi.lib().r2_pipeline_run(str(tmp_path / "config.json").encode(), str(tmp_path).encode(), 0, C.byref(rep),
                                    err, 1024)
    assert rc == 0, err.value.decode()
    assert rep.lods_filtered == 2 and rep.media_rendered == 1 and rep.objects_skipped == 0
    # the rendered LOD's .brickgrid (written by the filter step) equals its rebuilt brick grid
    assert rep.brickgrids_matched == 1 and rep.brickgrids_differing == 0
    assert (rep.width, rep.height, rep.spp) == (96, 64, 8)
    # grid files equal the library's own filter of the same mesh and parameters
    gm = ctx.upload_mesh(dict(mesh, pos=(0, 0, 0)))
    for lod in range(2):
        v = np.float32(vs * 2 ** lod)
        path = tmp_path / "filtering" / "blob" / f"filtered_mesh_{'%f' % v}.r2grid"
        data = path.read_bytes()
        assert data[:8] == b"R2GRID01"
        nx, ny, nz = struct.unpack("<3i", data[8:20])
        rec = np.frombuffer(data[20 + 48 + 4:], np.uint8).reshape(nz, ny, nx, 32)
        ref, g = ctx.filter_lod(gm, r2.filter_params(float(v), 32, 4, 10.0, 0.1, True, seed=3, lod=lod))
        assert tuple(g.num_voxels) == (nx, ny, nz)
        assert np.array_equal(rec, ref)
        # the LOD's .brickgrid beside it equals the brick grid of the same records
        grid = ctx.upload_grid(scenes.filtered_grid(g, float(v), records=ref))
        check_against_device(grid, path.with_suffix(".brickgrid"), ref)
        grid.free()
        # and its .nvdb: the active voxels (density > 0, the 8 corners) hold their records
        info, nv = r2.nvdb_read(path.with_suffix(".nvdb"))
        assert (info["nx"], info["ny"], info["nz"]) == (nx, ny, nz)
        dens = ref[..., :4].copy().view(np.float32)[..., 0]
        act = dens > 0
        act[::nz - 1, ::ny - 1, ::nx - 1] = True
        bg = np.zeros(32, np.uint8)
        bg[29] = 0x3C
        assert np.array_equal(nv, np.where(act[..., None], ref, bg))
        # its map is the grid's index -> model transform (row-major, as NanoVDB's mMatF)
        i2m = np.frombuffer(data[20:68], np.float32)
        assert np.array_equal(info["map_mat"], i2m[:9].reshape(3, 3).T.ravel())
        assert np.array_equal(info["map_vec"], i2m[9:12])
    desc = json.loads((tmp_path / "filtering" / "volume_description.json").read_text())
    assert set(desc["blob"].keys()) == {"%f" % vs, "%f" % np.float32(vs * 2)}
    # output.png = post-process(output.pfm), rows flipped
    hdr = read_pfm(tmp_path / "output.pfm")
    png = read_png(tmp_path / "output.png")
    assert np.isfinite(hdr).all() and hdr.mean() > 0
    _, by = O.postprocess(np.ascontiguousarray(hdr), 2.0)
    assert np.array_equal(png, by[::-1])
    # output.jpg: the same bytes as a quality-100 JPEG (decoded by libjpeg through PIL),
    # no further from them than libjpeg's own quality-100 encoding of the same bytes
    from PIL import Image
    from test_image import libjpeg_q100_error
    jpg = np.asarray(Image.open(tmp_path / "output.jpg").convert("RGB")).astype(np.int64)
    lib = libjpeg_q100_error(png)
    assert jpg.shape == png.shape and np.abs(jpg - png.astype(np.int64)).max() <= max(lib.max(), 3)
    # the same scene rendered from the .nvdb files alone (the reference's own filter
    # output format): the .r2grid files removed, filtering off -> the same image
    for f in (tmp_path / "filtering" / "blob").glob("*.r2grid"):
        f.unlink()
    cfg["Filtering"]["Active"] = False
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    out2 = tmp_path / "from_nvdb"
    out2.mkdir()
    rc = capi.lib().r2_pipeline_run(str(tmp_path / "config.json").encode(), str(out2).encode(), 0, C.byref(rep),
                                    err, 1024)
    assert rc == 0, err.value.decode()
    assert rep.lods_filtered == 0 and rep.media_rendered == 1
    assert np.array_equal(read_pfm(out2 / "output.pfm"), hdr)


def test_pipeline_renders_surfaces(ctx, tmp_path):
    """A scene description with the forest scenes' ground Circle and a Sphere
    beside a Medium read from a NanoVDB file: the front end renders both surfaces
    with their materials, and its frame equals r2_render of the same scene."""
    w = scenes.mixed_workload("c1", width=64, height=48, spp=16, max_depth=8)
    g = w["grids"][0]
    vs = 1.0 / 64
    path = tmp_path / "blob" / ("filtered_mesh_%f" % vs)
    path.parent.mkdir()
    r2.nvdb_write(str(path) + ".nvdb", g["records"], np.array(g["map_mat"], np.float32),
                  np.array(g["map_vec"], np.float32))
    rc_ = w["render"]
    cfg = {
        "General": {"FrameWidth": 64, "FrameHeight": 48, "ChannelsPerPixel": 3, "Samples": 16, "Gamma": 2.0,
                    "UseBrickGrid": True, "DebugMode": "NONE", "MaxDepth": 8, "Seed": 3},
        "Filtering": {"Active": False, "MinVoxelSize": vs, "LODs": 1, "SamplesPerVoxel": 32, "Debug": False,
                      "DebugSamples": 1, "SigmaT": 10.0, "EstimationIterations": 4, "Alpha": 0.1, "ClipRays": True},
        "Camera": {"Fov": rc_["fov_deg"], "Pos": list(rc_["cam_pos"]), "LookAt": list(rc_["cam_look_at"]),
                   "Up": list(rc_["cam_up"])},
        "Scene": {"Envmap": "synth:white", "Scenedescription": "scene.json"},
    }
    shape_key = {"circle": "Circle", "sphere": "Sphere"}
    scene = [{"Type": "Medium", "Path": "blob/filtered_mesh_%f" % vs, "SigmaA": list(g["sigma_a"]),
              "SigmaS": list(g["sigma_s"]), "Pos": [0, 0, 0], "Orientation": [0, 1, 0], "Scaling": [1, 1, 1]}]
    for sf in rc_["surfaces"]:
        scene.append({"Type": shape_key[sf["shape"]], "Pos": list(sf["pos"]), "Radius": sf["radius"],
                      "Normal": list(sf["normal"]), "DiffuseReflection": list(sf["diffuse"]),
                      "DiffuseRoughness": sf["diffuse_roughness"], "SpecularReflection": list(sf["specular"]),
                      "AlphaX": sf["alpha_x"], "AlphaY": sf["alpha_y"], "EtaI": sf["eta_i"], "EtaT": sf["eta_t"]})
    (tmp_path / "config.json").write_text("// surfaces\n" + json.dumps(cfg))
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    rep = capi.PipelineReport()
    err = C.create_string_buffer(1024)
    rc = capi.lib().r2_pipeline_run(str(tmp_path / "config.json").encode(), str(tmp_path).encode(), 0, C.byref(rep),
                                    err, 1024)
    assert rc == 0, err.value.decode()
    assert rep.media_rendered == 1 and rep.surfaces_rendered == 2 and rep.objects_skipped == 0
    hdr = read_pfm(tmp_path / "output.pfm")
    grid = ctx.upload_grid(dict(g, voxel_size_filtering=float(np.float32(vs))))
    env = ctx.upload_env(scenes.synth_env("white", 32, 16))
    ref = ctx.render(dict(rc_, width=64, height=48, spp=16, max_depth=8, seed=3, stream=0), [grid], w["instances"],
                     env)
    assert np.array_equal(hdr, ref)
    # the ground is visible: the bottom rows see the dark ground, not the white sky
    assert hdr[:4].mean() < 0.5 * hdr[-4:].mean()
