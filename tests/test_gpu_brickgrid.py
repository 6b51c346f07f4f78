"""The .brickgrid file (r2_grid_write_brickgrid, and beside every filtered LOD of
r2_pipeline_run), read back by a parser of cereal's portable binary archive
written here from cereal 1.3.2's rules: a leading byte 1 (little-endian data),
little-endian fields, std::vector as a uint64 length plus elements.  The content
must equal the brick grid the device built (which test_gpu_render checks bit for
bit against the oracle's CHostBrickGrid): ranges, mips, and per non-empty brick
its atlas block behind the encoded pointer.  The reference ships no .brickgrid
file, so the byte layout itself is checked only against this reading of cereal
and of host_grid_brick.hpp:47-72 (parity unpinned).
"""
import pytest

from brickgrid_io import check_against_device
from renderer2_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [64, 13])
def test_brickgrid_file_matches_device(ctx, tmp_path, n):
    w = scenes.make_workload("c2", n=n)
    grid = ctx.upload_grid(w["grids"][0])
    path = tmp_path / "g.brickgrid"
    grid.write_brickgrid(path)
    g = check_against_device(grid, path, w["grids"][0]["records"])
    assert g["count"] > 0


@pytest.mark.parametrize("n", [64, 13])
def test_brickgrid_reader_returns_the_device_grid(ctx, tmp_path, n):
    """r2_brickgrid_read (rt::read_grid) of a written file gives back the device's
    brick grid: ranges, atlas densities in brick layout, mips, extrema."""
    import numpy as np
    import renderer2_amd as r2

    w = scenes.make_workload("c2", n=n)
    grid = ctx.upload_grid(w["grids"][0])
    path = tmp_path / "g.brickgrid"
    grid.write_brickgrid(path)
    info, rng, u8, mips = r2.brickgrid_read(path)
    drng, du8, dmips = grid.download_bricks()
    gi = grid.info()
    assert info["n_bricks"] == tuple(int(v) for v in gi["n_bricks"])
    assert info["n_nonempty_bricks"] == gi["n_nonempty_bricks"]
    assert np.float32(info["max_density"]) == np.float32(gi["max_density"])
    assert np.float32(info["min_density"]) == np.float32(gi["min_density"])
    assert np.array_equal(rng, drng) and np.array_equal(u8, du8)
    for a, b in zip(mips, dmips):
        assert np.array_equal(a, b)
