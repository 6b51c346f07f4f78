"""C ABI checks that need no GPU: the library loads, exports every function that
include/r2.h declares with the ctypes signature table in step, and the host-only
entry points (version, synthetic inputs, argument validation) behave."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from renderer2_amd import _capi as capi
from renderer2_amd import scenes

HEADER = Path(__file__).resolve().parent.parent / "include" / "r2.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w*]+\s*\*?\s*(r2_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    L = capi.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_signature_table_covers_the_header():
    table = {n for n, _, _ in capi.SIGNATURES}
    assert set(declared_functions()) == table


def test_version():
    assert capi.lib().r2_version() == 3


def test_null_arguments_are_rejected():
    L = capi.lib()
    assert L.r2_destroy(None) == 0
    assert L.r2_grid_get_info(None, None) == -1
    assert L.r2_filter_setup(None, C.c_float(1.0), None) == -1
    assert L.r2_server_submit(None, None) == -1
    assert L.r2_server_stop(None) == 0
    assert L.r2_synth_volume(1, 0, 1, 4.0, 3, None) < 0
    nv, nt = C.c_int32(0), C.c_int32(0)
    assert L.r2_synth_mesh(0, 0, 0.01, 1, C.byref(nv), C.byref(nt), None, None, None, None) < 0


def test_synth_volume_deterministic_and_ranged():
    a = scenes.synth_volume(1, 24, seed=3, scale=4.0, octaves=3)
    b = scenes.synth_volume(1, 24, seed=3, scale=4.0, octaves=3)
    c = scenes.synth_volume(1, 24, seed=4, scale=4.0, octaves=3)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    d = a[..., :4].copy().view(np.float32)[..., 0]
    assert d.min() >= 0.0 and d.max() <= 4.0 and d.std() > 0.1
    # isotropic kind: S = I -> sigma = 65535, r = 32767; normal +y; ior half(1.5)
    u16 = a.view(np.uint16).reshape(-1, 16)
    assert np.all(u16[:, 2:5] == 65535) and np.all(u16[:, 5:8] == 32767)
    assert np.all(u16[:, 14].view(np.float16) == np.float16(1.5))


def test_synth_volume_fibres_are_unit_directions():
    a = scenes.synth_volume(2, 16, seed=1, scale=8.0, octaves=4)
    u16 = a.view(np.uint16).reshape(-1, 16).astype(np.float64)
    sig = u16[:, 2:5] / 65535.0
    r = 2.0 * u16[:, 5:8] / 65535.0 - 1.0
    S = np.zeros((len(sig), 3, 3))
    for i in range(3):
        S[:, i, i] = sig[:, i] ** 2
    for k, (i, j) in enumerate(((0, 1), (0, 2), (1, 2))):
        S[:, i, j] = S[:, j, i] = r[:, k] * sig[:, i] * sig[:, j]
    w = np.linalg.eigvalsh(S)
    # S = R diag(1, 1, 0.05) R^T up to the 16-bit codec
    np.testing.assert_allclose(w[:, 0], 0.05, atol=0.01)
    np.testing.assert_allclose(w[:, 1:], 1.0, atol=0.01)


def test_synth_env_kinds():
    white = scenes.synth_env("white", 16, 8)
    assert np.all(white == 1.0)
    sky = scenes.synth_env("sky", 512, 256)
    assert sky.min() > 0 and sky.max() / np.median(sky) > 1e3  # sun over sky


def test_synth_mesh_counts_and_normals():
    m = scenes.synth_mesh(6, 10, 0.05, seed=2)
    assert m["indices"].shape == (12 * 36 + 20, 3)
    assert m["positions"].shape == (6 * 49 + 40, 3)
    assert np.all(m["material"][: 12 * 36] == 0) and np.all(m["material"][12 * 36:] == 1)
    n = np.linalg.norm(m["normals"], axis=1)
    np.testing.assert_allclose(n, 1.0, atol=1e-5)
    # sphere vertices lie at radius 1 +- 0.05, normals point outwards
    p = m["positions"][: 6 * 49]
    rr = np.linalg.norm(p, axis=1)
    assert rr.min() > 0.94 and rr.max() < 1.06
    assert np.all(np.sum(p * m["normals"][: 6 * 49], axis=1) > 0)
    m2 = scenes.synth_mesh(6, 10, 0.05, seed=2)
    assert np.array_equal(m["positions"], m2["positions"])


def test_create_without_gpu_fails_cleanly():
    from renderer2_amd import _hip

    n = C.c_int(0)
    if _hip.hip().hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = capi.lib().r2_create(C.byref(h), 0)
    assert rc < 0 and not h.value


def test_fast_math_only_in_render_sources():
    """The bit-exact kernels (filter, grid build, post-process) must keep separately
    and correctly rounded arithmetic; only the render kernels take RENDER_FLAGS."""
    from renderer2_amd import build as b
    assert b.RENDER_SOURCES == {"r2_render.hip", "r2_render_packed.hip"}
    assert "-ffp-contract=off" in b.COMMON_FLAGS
    assert not any("fast" in f or "denormals" in f or "correctly-rounded" in f for f in b.COMMON_FLAGS)
    assert {"r2_filter.hip", "r2_grid.hip", "r2_post.hip"} <= set(b.SOURCES)


C_PROGRAM = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "r2.h"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  char path[1024], err[256];
  if (r2_version() != R2_VERSION) return 3;
  /* a synthetic 16^3 volume of 32-byte records, written and read back as .nvdb */
  const int n = 16;
  unsigned char* rec = malloc((size_t)n * n * n * 32);
  unsigned char* back = malloc((size_t)n * n * n * 32);
  if (!rec || !back) return 4;
  if (r2_synth_volume(1, n, 7u, 1.0f, 2, rec) != R2_OK) return 5;
  snprintf(path, sizeof path, "%s/v.nvdb", argv[1]);
  if (r2_nvdb_write(path, n, n, n, rec, NULL, NULL, err, (int32_t)sizeof err) != R2_OK) return 6;
  r2_nvdb_info info;
  memset(&info, 0, sizeof info);
  if (r2_nvdb_read(path, &info, back, (int64_t)n * n * n * 32, err, (int32_t)sizeof err) != R2_OK) return 7;
  if (info.nx != n || info.ny != n || info.nz != n || info.active_voxels <= 0) return 8;
  /* a 4x3 frame as the reference's two image files */
  unsigned char rgb[4 * 3 * 3];
  for (int i = 0; i < 36; ++i) rgb[i] = (unsigned char)(i * 7);
  snprintf(path, sizeof path, "%s/output.png", argv[1]);
  if (r2_image_write(path, 4, 3, rgb, R2_IMAGE_PNG, 0, err, (int32_t)sizeof err) != R2_OK) return 9;
  snprintf(path, sizeof path, "%s/output.jpg", argv[1]);
  if (r2_image_write(path, 4, 3, rgb, R2_IMAGE_JPG, 100, err, (int32_t)sizeof err) != R2_OK) return 10;
  if (r2_image_write(path, 4, 3, rgb, R2_IMAGE_JPG, 0, err, (int32_t)sizeof err) != R2_ERR_INVALID) return 11;
  printf("%lld\n", (long long)info.active_voxels);
  free(rec);
  free(back);
  return 0;
}
"""


def test_header_is_c_and_abi_callable_from_c(tmp_path):
    """include/r2.h compiles as strict C11 and the host-only entry points work from a
    C program linked against the library (the binding a C host would use)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "abi.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "abi"
    libdir = capi.LIB_PATH.parent
    cmd = [gcc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", str(HEADER.parent), str(src),
           "-o", str(exe), "-L", str(libdir), "-lrenderer2", "-L/opt/rocm/lib", f"-Wl,-rpath,{libdir}",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, (out.returncode, out.stderr[-2000:])
    assert int(out.stdout.strip()) > 0
    assert (tmp_path / "output.png").stat().st_size > 0 and (tmp_path / "output.jpg").stat().st_size > 0
