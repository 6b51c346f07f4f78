"""Pins of four oracle functions that no other test reaches independently:
the environment texel fetch (env_tex), the environment's lookup pdf
(dist2d_pdf inside env_le), the render-side getRotation of instance
transforms, and the frame post-process.  Each check is fixed by the
mathematics (a closed form, a known matrix, a float64 restatement), chosen so
that a plausible mistake (a missing half-texel shift, H instead of H - 2, a
transposed axis, a dropped rounding, a wrong divisor) fails it.  CPU only."""
import math

import numpy as np
import pytest

import oracle as O
from renderer2_amd import scenes

FLT_MIN = float(np.finfo(np.float32).tiny)


def _env_scene(env):
    rec = scenes.synth_volume(1, 8, 1, 1.0, 1)
    vs = np.float32(1.0) / np.float32(8)
    grid = dict(records=rec, shape=(8, 8, 8), map_mat=(vs, 0, 0, 0, vs, 0, 0, 0, vs), map_vec=(-0.5,) * 3,
                sigma_a=(0.2,) * 3, sigma_s=(0.8,) * 3, voxel_size_filtering=float(vs))
    return O.Scene([grid], [dict(grid=0)], np.ascontiguousarray(env, np.float32), "f64")


def _dir(x, y):
    """Direction whose lat-long texture coordinates are (x, y): phi = 2 pi x from +z
    towards +x, theta = pi y from +y (environmentmap.hpp:44-55)."""
    ph, th = 2 * math.pi * x, math.pi * y
    return np.array([math.sin(ph) * math.sin(th), math.cos(th), math.cos(ph) * math.sin(th)])


# ------------------------------------------------------------- env_tex ---


def test_env_tex_reproduces_linear_texture():
    """Bilinear filtering reproduces a texture that is linear in the texel indices
    exactly, at texel-centre coordinates x W - 1/2, y H - 1/2 (texture.cu:166-189:
    cudaFilterModeLinear, normalised coordinates)."""
    H, W = 12, 20
    j, i = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    coef = np.array([[3, 2, 5], [1, 4, 7], [2, 1, 3]], np.float64)  # per channel: a, b (per column), c (per row)
    env = np.stack([coef[c, 0] + coef[c, 1] * i + coef[c, 2] * j for c in range(3)], -1)
    sc = _env_scene(env)
    rng = np.random.default_rng(2)
    for _ in range(300):
        x = rng.uniform(0.5 / W, (W - 0.5) / W)
        y = rng.uniform(0.5 / H, (H - 0.5) / H)
        got = sc.env_le(_dir(x, y))[:3]
        tx, ty = x * W - 0.5, y * H - 0.5
        want = coef[:, 0] + coef[:, 1] * tx + coef[:, 2] * ty
        assert np.allclose(got, want, rtol=1e-9, atol=1e-9), (x, y, got, want)


def test_env_tex_texel_centres_and_wrap():
    """At a texel's centre the fetch returns that texel; a quarter texel outside the
    first or last column or row blends with the opposite edge (cudaAddressModeWrap
    on both axes)."""
    H, W = 8, 16
    rng = np.random.default_rng(3)
    env = rng.uniform(0.1, 5.0, (H, W, 3))
    sc = _env_scene(env)
    T = env.astype(np.float32).astype(np.float64)
    for jj in range(1, H - 1):
        for ii in range(W):
            got = sc.env_le(_dir((ii + 0.5) / W, (jj + 0.5) / H))[:3]
            assert np.allclose(got, T[jj, ii], rtol=1e-9), (jj, ii)
    yc = (3 + 0.5) / H  # centre of row 3
    assert np.allclose(sc.env_le(_dir(0.25 / W, yc))[:3], 0.25 * T[3, W - 1] + 0.75 * T[3, 0], rtol=1e-9)
    assert np.allclose(sc.env_le(_dir((W - 0.25) / W, yc))[:3], 0.75 * T[3, W - 1] + 0.25 * T[3, 0], rtol=1e-9)
    xc = (5 + 0.5) / W  # centre of column 5
    assert np.allclose(sc.env_le(_dir(xc, 0.25 / H))[:3], 0.25 * T[H - 1, 5] + 0.75 * T[0, 5], rtol=1e-9)
    assert np.allclose(sc.env_le(_dir(xc, (H - 0.25) / H))[:3], 0.75 * T[H - 1, 5] + 0.25 * T[0, 5], rtol=1e-9)


# --------------------------------------------------------- dist2d_pdf ---


@pytest.mark.parametrize("H,W,j0,i0", [(10, 16, 4, 7), (8, 12, 2, 3), (16, 9, 9, 1)])
def test_env_le_pdf_single_texel_tent(H, W, j0, i0):
    """CDistribution2D::pdf (distribution_2d.hpp:73-84) is the unnormalised bilinear
    interpolation of the row functions on a (W - 2) x (H - 2) lattice.  For one
    bright texel (j0, i0) of value v its row function is F = 3 v sin(pi (j0 + 1/2) / H)
    (environmentmap.cu:16-29) and every other entry is 0, so env_le's pdf is the
    tent product F t(x (W-2) - i0) t(y (H-2) - j0) / (2 pi^2 sin(pi y)), with
    t(u) = max(0, 1 - |u|): zero outside y in [(j0-1)/(H-2), (j0+1)/(H-2)]."""
    v = 2.5
    env = np.zeros((H, W, 3), np.float32)
    env[j0, i0] = v
    sc = _env_scene(env)
    F = 3 * v * math.sin(math.pi * (j0 + 0.5) / H)
    rng = np.random.default_rng(H * 100 + W)
    tent = lambda u: max(0.0, 1.0 - abs(u))  # noqa: E731
    hits = 0
    for _ in range(400):
        sx = rng.uniform(i0 - 1.3, i0 + 1.3)
        sy = rng.uniform(j0 - 1.3, j0 + 1.3)
        x, y = sx / (W - 2), sy / (H - 2)
        if not (0 < x < 1 and 0 < y < 1):
            continue
        pdf = sc.env_le(_dir(x, y))[3]
        want = F * tent(sx - i0) * tent(sy - j0) / (2 * math.pi ** 2 * math.sin(math.pi * y))
        assert abs(pdf - want) <= 1e-9 * max(1.0, abs(want)), (sx, sy, pdf, want)
        hits += want > 0
    assert hits > 100
    # four hand-computed points of the tent product
    for (dx, dy, w) in [(0.0, 0.0, 1.0), (0.5, 0.0, 0.5), (0.25, -0.5, 0.375), (-0.75, 0.25, 0.1875)]:
        x, y = (i0 + dx) / (W - 2), (j0 + dy) / (H - 2)
        pdf = sc.env_le(_dir(x, y))[3]
        assert math.isclose(pdf, F * w / (2 * math.pi ** 2 * math.sin(math.pi * y)), rel_tol=1e-9)


# ------------------------------------------------------- getRotation ----


def _inst_scene(orientation, pos=(0, 0, 0), scaling=(1, 1, 1)):
    rec = scenes.synth_volume(1, 8, 1, 1.0, 1)
    vs = np.float32(1.0) / np.float32(8)
    grid = dict(records=rec, shape=(8, 8, 8), map_mat=(vs, 0, 0, 0, vs, 0, 0, 0, vs), map_vec=(-0.5,) * 3,
                sigma_a=(0.2,) * 3, sigma_s=(0.8,) * 3, voxel_size_filtering=float(vs))
    inst = dict(grid=0, pos=pos, orientation=orientation, scaling=scaling)
    return O.Scene([grid], [inst], np.ones((4, 8, 3), np.float32), "f64")


# getRotation (functions.hpp:128-142): the rotation by acos(n . y) about
# normalize(n x y), i.e. the one taking n to +y, with every entry rounded to an
# integer; +y gives the identity and -y gives diag(-1, -1, -1).  Rows of R.
ROTATIONS = [
    ((0, 1, 0), [[1, 0, 0], [0, 1, 0], [0, 0, 1]]),
    ((0, -1, 0), [[-1, 0, 0], [0, -1, 0], [0, 0, -1]]),
    ((1, 0, 0), [[0, -1, 0], [1, 0, 0], [0, 0, 1]]),
    ((-1, 0, 0), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]]),
    ((0, 0, 1), [[1, 0, 0], [0, 0, 1], [0, -1, 0]]),
    ((0, 0, -1), [[1, 0, 0], [0, 0, -1], [0, 1, 0]]),
    # 45 degrees about +z: cos = sin = 0.7071 round to 1, so the rounded matrix is
    # not a rotation (the reference's rounding, kept)
    ((math.sqrt(0.5), math.sqrt(0.5), 0), [[1, -1, 0], [1, 1, 0], [0, 0, 1]]),
]


@pytest.mark.parametrize("n,R", ROTATIONS)
def test_get_rotation_known_matrices(n, R):
    info = _inst_scene(n).instance_info(0)
    m2w = info["model_to_world"]
    assert np.array_equal(m2w[:, :3], np.array(R, np.float64)), m2w
    assert np.array_equal(m2w[:, 3], np.zeros(3))
    if abs(abs(n[1]) - 1) > 0 and abs(n[0] * n[1]) == 0:  # axis-aligned, not +-y: R n = +y exactly
        assert np.array_equal(np.array(R) @ np.array(n, np.float64), [0, 1, 0])


def test_model_to_world_is_translate_rotate_scale():
    """getModelToWorldTransform (sceneobject.cu:375-377) = T(pos) R(orientation) S(scaling):
    the columns of R are scaled, the translation is pos, and world_to_model inverts it."""
    pos, sc = (1.0, -2.0, 3.0), (2.0, 3.0, 4.0)
    info = _inst_scene((1, 0, 0), pos=pos, scaling=sc).instance_info(0)
    R = np.array(ROTATIONS[2][1], np.float64)
    want = np.concatenate([R * np.array(sc)[None, :], np.array(pos)[:, None]], 1)
    assert np.allclose(info["model_to_world"], want, rtol=0, atol=1e-12)
    M = np.vstack([info["model_to_world"], [0, 0, 0, 1]])
    Mi = np.vstack([info["world_to_model"], [0, 0, 0, 1]])
    assert np.allclose(Mi @ M, np.eye(4), atol=1e-12)


# --------------------------------------------------------- post-process --


def _post_f64(hdr, gamma):
    """raytracer.cu:53-127 in float64: k = exp(sum of log(r + g + b + FLT_MIN) / (3 W H))
    - W H FLT_MIN, then c / (c + k), then c^(1 / gamma); bytes = round(255 c)."""
    h, w = hdr.shape[:2]
    x = hdr.astype(np.float64)
    lum = x.sum(-1) + FLT_MIN
    k = math.exp(float(np.sum(np.log(lum)) / (3.0 * w * h))) - w * h * FLT_MIN
    t = (x / (x + k)) ** (1.0 / gamma)
    return t, k


@pytest.mark.parametrize("h,w,gamma", [(20, 48, 2.0), (33, 17, 2.2), (64, 64, 1.0)])
def test_postprocess_matches_float64_closed_form(h, w, gamma):
    rng = np.random.default_rng(h * w)
    hdr = (rng.lognormal(0.0, 1.5, (h, w, 3))).astype(np.float32)
    hdr[rng.uniform(size=(h, w)) < 0.05] = 0.0  # black pixels: the FLT_MIN term dominates their log
    ldr, by = O.postprocess(hdr, gamma)
    t, _ = _post_f64(hdr, gamma)
    assert np.allclose(ldr, t, rtol=0, atol=5e-5)
    v = t * 255.0
    away = np.abs(v - np.floor(v) - 0.5) > 5e-2  # not within float32 rounding of a .5 boundary
    assert away.mean() > 0.8
    assert np.array_equal(by[away], np.floor(v[away] + 0.5).astype(np.uint8))


def test_postprocess_divisor_is_three_w_h():
    """A uniform grey frame c: the key is k = exp(log(3 c) / 3) - W H FLT_MIN, so the
    tonemapped value is c / (c + (3c)^(1/3)) -- the divisor counts channels too."""
    c = 0.8
    hdr = np.full((10, 30, 3), c, np.float32)
    ldr, _ = O.postprocess(hdr, 1.0)
    k = (3 * np.float64(np.float32(c))) ** (1 / 3) - 300 * FLT_MIN
    assert np.allclose(ldr, np.float32(c) / (np.float32(c) + k), rtol=2e-6)
