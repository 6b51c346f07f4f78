"""NanoVDB grid files (.nvdb) of filtered voxel records, host only (no GPU).

The reference writes each filtered LOD with openToNanoVDB + nanovdb::io::writeGrid
(raytracer/src/filtering/openvdb_backend.cpp:108-157) and reads it back in
CNVDBMedium (raytracer/src/medium/nvdb_medium.cu:130-173, 308-323) through the index
bounding box and voxel lookups.  The NanoVDB library is not in this image and the
reference ships no .nvdb file, so the byte layout (node sizes and field offsets of
NanoVDB 32.3) is parity unpinned.  What these tests pin:

  * active set: density > 0 or one of the 8 grid corners (setValues,
    openvdb_backend.cpp:113-126); every other voxel reads as the background
    record (zeros, ior = half 1.0; openvdb_backend.cpp:30-42);
  * tree semantics, through a top-down accessor written here that follows the
    root tile key -> upper child -> lower child -> leaf value path as NanoVDB's
    ReadAccessor does; the library's reader scans the leaf array instead, so a
    wrong key, child index, child offset or mask bit fails one of the two;
  * counts that follow from the input alone: active voxels, leaves (distinct
    8^3 blocks holding an active voxel), lower (128^3) and upper (4096^3) nodes;
  * the file header fields (magic, version 32.3.3, one grid, no codec) and
    error handling (truncated file, wrong magic, short buffer).
"""
import struct

import numpy as np
import pytest

import renderer2_amd as r2
from renderer2_amd import capi

BG = np.zeros(32, np.uint8)
BG[28:30] = (0x00, 0x3C)

# NanoVDB 32.3 node layout for Vec4d values (see the module docstring)
GRID_DATA, TREE_DATA, ROOT_DATA, ROOT_TILE = 672, 64, 160, 64
UPPER_TABLE, LOWER_TABLE, LEAF_TABLE = 8320, 1152, 160


def records(shape, seed, occupancy):
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    rec = rng.integers(0, 256, (nz, ny, nx, 32), dtype=np.uint8)
    dens = np.where(rng.random((nz, ny, nx)) < occupancy, rng.uniform(0.1, 5.0, (nz, ny, nx)),
                    rng.choice([0.0, -0.0, -1.0], (nz, ny, nx))).astype(np.float32)
    rec[..., :4] = dens.view(np.uint8).reshape(nz, ny, nx, 4)
    return rec, dens


def active_mask(dens):
    nz, ny, nx = dens.shape
    act = dens > 0
    for z in (0, nz - 1):
        for y in (0, ny - 1):
            for x in (0, nx - 1):
                act[z, y, x] = True
    return act


class Accessor:
    """Top-down voxel lookup over the grid buffer of a .nvdb file."""

    def __init__(self, data: bytes):
        name_size = struct.unpack_from("<I", data, 16 + 136)[0]
        self.d = data
        self.base = 16 + 176 + name_size
        tree = self.base + GRID_DATA
        self.root = tree + struct.unpack_from("<Q", data, tree + 24)[0]
        self.table = struct.unpack_from("<I", data, self.root + 24)[0]
        self.bg = np.frombuffer(data, np.uint8, 32, self.root + 32)

    def _bit(self, off, n):
        return (self.d[off + n // 8] >> (n % 8)) & 1

    def _val(self, off):
        return np.frombuffer(self.d, np.uint8, 32, off)

    def get(self, x, y, z):
        d = self.d
        key = ((z & 0xFFFFFFFF) >> 12) | (((y & 0xFFFFFFFF) >> 12) << 21) | (((x & 0xFFFFFFFF) >> 12) << 42)
        for k in range(self.table):
            t = self.root + ROOT_DATA + ROOT_TILE * k
            if struct.unpack_from("<Q", d, t)[0] == key:
                child = struct.unpack_from("<q", d, t + 8)[0]
                if child == 0:
                    return self._val(t + 24)
                up = self.root + child
                break
        else:
            return self.bg
        n = (((x & 4095) >> 7) << 10) | (((y & 4095) >> 7) << 5) | ((z & 4095) >> 7)
        if not self._bit(up + 32 + 4096, n):
            return self._val(up + UPPER_TABLE + 32 * n)
        lo = up + struct.unpack_from("<q", d, up + UPPER_TABLE + 32 * n)[0]
        n = (((x & 127) >> 3) << 8) | (((y & 127) >> 3) << 4) | ((z & 127) >> 3)
        if not self._bit(lo + 32 + 512, n):
            return self._val(lo + LOWER_TABLE + 32 * n)
        leaf = lo + struct.unpack_from("<q", d, lo + LOWER_TABLE + 32 * n)[0]
        n = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7)
        return self._val(leaf + LEAF_TABLE + 32 * n)


@pytest.mark.parametrize("shape,occ", [((1, 1, 1), 0.0), ((5, 9, 13), 0.3), ((24, 16, 8), 0.05),
                                       ((20, 140, 9), 0.02), ((17, 17, 17), 0.0)])
def test_round_trip_and_active_set(tmp_path, shape, occ):
    rec, dens = records(shape, 7, occ)
    p = tmp_path / "g.nvdb"
    r2.nvdb_write(p, rec)
    info, got = r2.nvdb_read(p)
    nz, ny, nx = shape
    act = active_mask(dens)
    assert (info["nx"], info["ny"], info["nz"]) == (nx, ny, nz)
    assert tuple(info["ibb_min"]) == (0, 0, 0)
    assert info["active_voxels"] == int(act.sum())
    assert np.allclose(info["map_mat"], np.eye(3).ravel()) and not np.any(info["map_vec"])
    want = np.where(act[..., None], rec, BG)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape,occ", [((5, 9, 13), 0.3), ((24, 16, 8), 0.05), ((20, 140, 9), 0.02)])
def test_read_leaves_matches_dense_read(tmp_path, shape, occ):
    """The sparse read: one leaf per leaf node of the file, its active voxels and
    their records (x fastest within the leaf), background elsewhere."""
    rec, dens = records(shape, 11, occ)
    p = tmp_path / "g.nvdb"
    r2.nvdb_write(p, rec)
    info, dense = r2.nvdb_read(p)
    li, org, lrec, msk = r2.nvdb_read_leaves(p)
    nz, ny, nx = shape
    act = active_mask(dens)
    assert tuple(li["origin_shift"]) == (0, 0, 0) and len(org) == info["leaf_count"]
    assert (org % 8 == 0).all() and len({tuple(o) for o in org}) == len(org)
    seen = np.zeros(shape, bool)
    for (ox, oy, oz), lr, m in zip(org, lrec, msk):
        bits = np.unpackbits(m, bitorder="little").astype(bool)
        for v in range(512):
            x, y, z = ox + (v & 7), oy + ((v >> 3) & 7), oz + (v >> 6)
            inside = x < nx and y < ny and z < nz
            assert bits[v] == (inside and act[z, y, x])
            if inside:
                assert np.array_equal(lr[v], dense[z, y, x])
                seen[z, y, x] |= bits[v]
    assert np.array_equal(seen, act)


def test_node_counts_follow_from_the_active_set(tmp_path):
    rec, dens = records((40, 150, 70), 3, 0.01)
    p = tmp_path / "g.nvdb"
    r2.nvdb_write(p, rec)
    info, _ = r2.nvdb_read(p, with_records=False)
    z, y, x = np.nonzero(active_mask(dens))
    leaves = {(a >> 3, b >> 3, c >> 3) for a, b, c in zip(x, y, z)}
    lowers = {(a >> 7, b >> 7, c >> 7) for a, b, c in zip(x, y, z)}
    assert info["leaf_count"] == len(leaves)
    assert info["lower_count"] == len(lowers)
    assert info["upper_count"] == 1


def test_top_down_accessor_agrees(tmp_path):
    """Every voxel (and a ring of voxels outside the box) through the tree."""
    rec, dens = records((12, 135, 20), 11, 0.1)  # spans two lower nodes in y
    p = tmp_path / "g.nvdb"
    r2.nvdb_write(p, rec)
    data = p.read_bytes()
    acc = Accessor(data)
    act = active_mask(dens)
    nz, ny, nx = dens.shape
    for z in range(-1, nz + 1):
        for y in range(-1, ny + 1, 3):
            for x in range(-1, nx + 1):
                inside = 0 <= x < nx and 0 <= y < ny and 0 <= z < nz
                want = rec[z, y, x] if inside and act[z, y, x] else BG
                assert np.array_equal(acc.get(x, y, z), want), (x, y, z)


def test_file_header(tmp_path):
    rec, dens = records((9, 9, 9), 5, 0.2)
    p = tmp_path / "g.nvdb"
    r2.nvdb_write(p, rec)
    data = p.read_bytes()
    magic, version, grids, codec = struct.unpack_from("<QIHH", data, 0)
    assert data[:8] == b"NanoVDB0"
    assert (version >> 21, (version >> 10) & 0x7FF, version & 0x3FF) == (32, 3, 3)
    assert (grids, codec) == (1, 0)
    grid_size, file_size, name_key, voxels, gtype = struct.unpack_from("<QQQQI", data, 16)
    name_size = struct.unpack_from("<I", data, 16 + 136)[0]
    assert len(data) == 16 + 176 + name_size + grid_size and file_size == grid_size
    assert (name_key, name_size, data[16 + 176]) == (0, 1, 0)
    assert voxels == int(active_mask(dens).sum())
    assert gtype == 18  # Vec4d
    g = 16 + 176 + name_size
    assert data[g:g + 8] == b"NanoVDB0"
    assert struct.unpack_from("<Q", data, g + 32)[0] == grid_size
    wmin = struct.unpack_from("<3d", data, g + 560)
    wmax = struct.unpack_from("<3d", data, g + 584)
    assert wmin == (0.0, 0.0, 0.0) and wmax == (9.0, 9.0, 9.0)


def test_map_world_box_and_voxel_size(tmp_path):
    """The index -> model map (setupGrid's indexToModel) is stored as the NanoVDB map;
    the world box is the index box (max + 1) through the map, the voxel size the
    column lengths, and the stored inverse inverts the stored matrix."""
    rec, _ = records((6, 10, 14), 2, 0.3)
    c, s_ = np.cos(0.3), np.sin(0.3)
    A = (np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]]) @ np.diag([0.02, 0.03, 0.05])).astype(np.float32)
    t = np.array([-0.4, 0.25, 1.5], np.float32)
    p = tmp_path / "g.nvdb"
    r2.nvdb_write(p, rec, A, t)
    info, got = r2.nvdb_read(p)
    assert np.array_equal(info["map_mat"], A.ravel()) and np.array_equal(info["map_vec"], t)
    data = p.read_bytes()
    g = 16 + 176 + 1
    Ad = np.array(struct.unpack_from("<9d", data, g + 296 + 88)).reshape(3, 3)
    Aid = np.array(struct.unpack_from("<9d", data, g + 296 + 160)).reshape(3, 3)
    td = np.array(struct.unpack_from("<3d", data, g + 296 + 232))
    assert np.array_equal(Ad, A.astype(np.float64)) and np.array_equal(td, t.astype(np.float64))
    assert np.abs(Ad @ Aid - np.eye(3)).max() < 1e-12
    corners = np.array([[x, y, z] for x in (0, 14) for y in (0, 10) for z in (0, 6)], np.float64)
    w = corners @ Ad.T + td
    wmin = np.array(struct.unpack_from("<3d", data, g + 560))
    wmax = np.array(struct.unpack_from("<3d", data, g + 584))
    assert np.allclose(wmin, w.min(0), rtol=0, atol=1e-12) and np.allclose(wmax, w.max(0), rtol=0, atol=1e-12)
    vs = np.array(struct.unpack_from("<3d", data, g + 608))
    assert np.allclose(vs, [0.02, 0.03, 0.05], rtol=1e-6)
    meta_w = np.array(struct.unpack_from("<6d", data, 16 + 40))
    assert np.array_equal(meta_w, np.concatenate([wmin, wmax]))
    with pytest.raises(r2.R2Error) as e:
        r2.nvdb_write(p, rec, np.zeros((3, 3), np.float32), t)
    assert e.value.code == -1  # singular map


def test_errors(tmp_path):
    rec, _ = records((8, 8, 8), 1, 0.5)
    p = tmp_path / "g.nvdb"
    r2.nvdb_write(p, rec)
    data = p.read_bytes()
    (tmp_path / "t.nvdb").write_bytes(data[: len(data) // 2])
    with pytest.raises(r2.R2Error) as e:
        r2.nvdb_read(tmp_path / "t.nvdb")
    assert e.value.code == -5  # R2_ERR_IO
    (tmp_path / "m.nvdb").write_bytes(b"X" + data[1:])
    with pytest.raises(r2.R2Error) as e:
        r2.nvdb_read(tmp_path / "m.nvdb")
    assert e.value.code == -5
    with pytest.raises(r2.R2Error) as e:
        r2.nvdb_read(tmp_path / "missing.nvdb")
    assert e.value.code == -5
    info = capi.NvdbInfo()
    buf = np.zeros(100, np.uint8)
    rc = capi.lib().r2_nvdb_read(str(p).encode(), info, buf.ctypes.data, buf.nbytes, None, 0)
    assert rc == -1  # R2_ERR_INVALID: buffer smaller than 8^3 records
    rc = capi.lib().r2_nvdb_write(str(p).encode(), 0, 1, 1, buf.ctypes.data, None, None, None, 0)
    assert rc == -1


def test_brickgrid_reader_decodes_the_archive(tmp_path):
    """r2_brickgrid_read on a file written by the test-side writer of cereal's
    portable binary rules: ranges, atlas blocks put back in brick layout through the
    indirection, mips and header fields; malformed files are R2_ERR_IO."""
    from brickgrid_io import write_brickgrid

    rng = np.random.default_rng(9)
    nb = (8, 16, 8)
    n = nb[0] * nb[1] * nb[2]
    lo = rng.integers(0, 30000, n).astype(np.uint32)
    hi = lo + rng.integers(0, 2, n).astype(np.uint32) * rng.integers(1, 3000, n).astype(np.uint32)
    ranges = (hi << 16) | lo
    dense = rng.integers(0, 256, (8 * nb[2], 8 * nb[1], 8 * nb[0])).astype(np.uint8)
    empty = ((ranges & 0xFFFF) == (ranges >> 16)).reshape(nb[2], nb[1], nb[0])
    want = dense.copy()
    for bz, by, bx in zip(*np.nonzero(empty)):
        want[8 * bz:8 * bz + 8, 8 * by:8 * by + 8, 8 * bx:8 * bx + 8] = 0
    mips = [rng.integers(0, 2 ** 32, (nb[0] >> (l + 1)) * (nb[1] >> (l + 1)) * (nb[2] >> (l + 1)),
                         dtype=np.uint64).astype(np.uint32) for l in range(3)]
    p = tmp_path / "g.brickgrid"
    write_brickgrid(p, nb, ranges, dense, mips, 0.25, 7.5)
    info, r, u8, m = r2.brickgrid_read(p)
    assert info["n_bricks"] == nb and info["n_nonempty_bricks"] == int((~empty).sum())
    assert info["min_density"] == 0.25 and info["max_density"] == 7.5
    assert np.array_equal(r, ranges) and np.array_equal(u8, want.reshape(-1))
    for a, b in zip(m, mips):
        assert np.array_equal(a, b)
    data = p.read_bytes()
    for bad in (data[:-3], data + b"\0", b"\2" + data[1:]):
        (tmp_path / "bad.brickgrid").write_bytes(bad)
        with pytest.raises(r2.R2Error) as e:
            r2.brickgrid_read(tmp_path / "bad.brickgrid")
        assert e.value.code == -5
    with pytest.raises(r2.R2Error):
        r2.brickgrid_read(tmp_path / "missing.brickgrid")
