"""Reader of the .brickgrid file for the tests: cereal 1.3.2's portable binary
archive of CHostBrickGrid (host_grid_brick.hpp:47-72, serialization.cpp:37-47) as
this build reads cereal's rules (a leading byte 1 for little-endian data,
little-endian fields, std::vector as a uint64 length and its elements), and a
check of a file against the brick grid the device built."""
import struct

import numpy as np


def read_brickgrid(path):
    b = open(path, "rb").read()
    pos = [0]

    def take(fmt):
        v = struct.unpack_from("<" + fmt, b, pos[0])
        pos[0] += struct.calcsize("<" + fmt)
        return v

    def buf3d(dtype):
        stride = take("3I")
        (n,) = take("Q")
        arr = np.frombuffer(b, dtype, n, pos[0]).copy()
        pos[0] += n * np.dtype(dtype).itemsize
        return stride, arr

    (le,) = take("B")
    g = dict(little_endian=le, transform=np.array(take("16f"), np.float32).reshape(4, 4), n_bricks=take("3I"))
    g["min"], g["max"] = take("2f")
    (g["count"],) = take("Q")
    g["indirection"], g["range"], g["atlas"] = buf3d(np.uint32), buf3d(np.uint32), buf3d(np.uint8)
    (nm,) = take("Q")
    g["mips"] = [buf3d(np.uint32) for _ in range(nm)]
    g["trailing"] = len(b) - pos[0]
    return g


def check_against_device(grid, path, records):
    g = read_brickgrid(path)
    inf = grid.info()
    rng_dev, u8_dev, mips_dev = grid.download_bricks()
    nbx, nby, nbz = (int(v) for v in inf["n_bricks"])
    assert g["little_endian"] == 1 and g["trailing"] == 0
    assert np.array_equal(g["transform"], np.eye(4, dtype=np.float32))
    assert g["n_bricks"] == (nbx, nby, nbz)
    nz, ny, nx = records.shape[:3]
    d = records[..., :4].copy().view(np.float32)[..., 0]
    idx = np.indices((nz, ny, nx))
    corner = np.ones_like(d, bool)
    for a, n in zip(idx, (nz, ny, nx)):
        corner &= (a == 0) | (a == n - 1)
    active = (d > 0) | corner
    assert np.float32(g["min"]) == np.where(active, d, 0).min()
    assert np.float32(g["max"]) == d[active].max()
    empty = (rng_dev & 0xFFFF) == (rng_dev >> 16)
    assert g["count"] == int((~empty).sum()) == int(inf["n_nonempty_bricks"])
    stride, ind = g["indirection"]
    assert stride == (nbx, nby, nbz) and np.array_equal(ind[empty], np.zeros(int(empty.sum()), np.uint32))
    assert g["range"][0] == (nbx, nby, nbz) and np.array_equal(g["range"][1], rng_dev)
    zs = 8 * int(np.round(np.ceil(g["count"] / (nbx * nby))))
    astride, atlas = g["atlas"]
    assert astride == (8 * nbx, 8 * nby, zs)
    atlas = atlas.reshape(zs, 8 * nby, 8 * nbx)
    dense = u8_dev.reshape(8 * nbz, 8 * nby, 8 * nbx)
    slots = []
    for bi in np.flatnonzero(~empty):
        p = int(ind[bi])
        px, py, pz = (p >> 22) & 1023, (p >> 12) & 1023, (p >> 2) & 1023
        slots.append(px + nbx * (py + nby * pz))
        bz, by, bx = bi // (nbx * nby), (bi // nbx) % nby, bi % nbx
        assert np.array_equal(atlas[8 * pz:8 * pz + 8, 8 * py:8 * py + 8, 8 * px:8 * px + 8],
                              dense[8 * bz:8 * bz + 8, 8 * by:8 * by + 8, 8 * bx:8 * bx + 8])
    assert slots == list(range(len(slots)))  # handed out in the reference's z, y, x brick order
    off = 0
    assert len(g["mips"]) == 3
    for lvl, (ms, md) in enumerate(g["mips"]):
        assert ms == (nbx >> (lvl + 1), nby >> (lvl + 1), nbz >> (lvl + 1))
        assert np.array_equal(md, mips_dev[lvl])
    return g


def write_brickgrid(path, nb, rng, atlas_dense, mips, vmin=0.0, vmax=1.0):
    """A .brickgrid file by the same rules (the test side of r2_brickgrid_read):
    non-empty bricks get atlas slots in z, y, x brick order."""
    nbx, nby, nbz = nb
    rng = np.asarray(rng, np.uint32).reshape(-1)
    empty = (rng & 0xFFFF) == (rng >> 16)
    count = int((~empty).sum())
    zs = 8 * int(np.ceil(count / (nbx * nby)))
    dense = np.asarray(atlas_dense, np.uint8).reshape(8 * nbz, 8 * nby, 8 * nbx)
    atlas = np.zeros((zs, 8 * nby, 8 * nbx), np.uint8)
    ind = np.zeros(rng.size, np.uint32)
    slot = 0
    for bi in np.flatnonzero(~empty):
        px, py, pz = slot % nbx, (slot // nbx) % nby, slot // (nbx * nby)
        slot += 1
        ind[bi] = (px << 22) | (py << 12) | (pz << 2)
        bz, by, bx = bi // (nbx * nby), (bi // nbx) % nby, bi % nbx
        atlas[8 * pz:8 * pz + 8, 8 * py:8 * py + 8, 8 * px:8 * px + 8] = \
            dense[8 * bz:8 * bz + 8, 8 * by:8 * by + 8, 8 * bx:8 * bx + 8]
    out = [struct.pack("<B", 1), np.eye(4, dtype=np.float32).tobytes(), struct.pack("<3I", *nb),
           struct.pack("<2f", vmin, vmax), struct.pack("<Q", count)]

    def buf(stride, a):
        out.append(struct.pack("<3I", *stride) + struct.pack("<Q", a.size) + a.tobytes())

    buf(nb, ind)
    buf(nb, rng)
    buf((8 * nbx, 8 * nby, zs), atlas.reshape(-1))
    out.append(struct.pack("<Q", 3))
    for lvl in range(3):
        buf((nbx >> (lvl + 1), nby >> (lvl + 1), nbz >> (lvl + 1)), np.asarray(mips[lvl], np.uint32))
    with open(path, "wb") as f:
        f.write(b"".join(out))
