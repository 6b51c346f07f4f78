"""Pins of the oracle's textured-material filter (CPU only).

The reference filters textured meshes through OptiX: `__anyhit__mesh` drops hits
where the material's alpha mask is not > 0, `__closesthit__ch` interpolates the
texture coordinate and applies the normal map in the triangle's tangent frame,
and CMeshFilter::run sums the texture-or-constant colours
(shaders.optix.cu:48-152, material.hpp:111-156, mesh_filter.hpp:51-122).  No
reference test or fixture covers this, so each part is pinned by what the
mathematics fixes:
  * the texture fetch: texel centres return the texel, a constant texture
    returns its value bit for bit (as tex2D's fixed-point weights, which sum to
    one, do; the normal map's zero test needs it), a texel field linear in the
    texel index is reproduced, and addressing wraps;
  * the alpha mask: a 4-texel-wide mask of (1, 1, 0, 0) is transparent exactly
    on the band u in [0.625, 0.875], so rays through it reach the quad behind
    (float64 Moller-Trumbore brute force);
  * textured colours: the filtered colour of a plane with a texture linear in
    the texture coordinate equals the mean of that linear function over the
    hit points, recomputed in float64 from the probe's hit distances;
  * the normal map: a constant normal texel c maps the plane normal to
    nm.x T + nm.y (-B) + nm.z N with nm = 2 c - 1 and (T, B) = (dP/du, dP/dv),
    or keeps N where nm = 0 (closed forms);
  * without texture coordinates a material texture is read at tc = (0, 0): the
    mean of the four corner texels.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_filter import MAT, _expected_plane, bbox_mesh, quad_mesh


def textured(m, uvs, textures, mat_textures):
    m = dict(m)
    m["uvs"] = np.asarray(uvs, np.float32)
    m["textures"] = [np.asarray(t, np.float32) for t in textures]
    m["mat_textures"] = np.asarray(mat_textures, np.int32).reshape(-1, 4)
    return m


def quad_uvs(n_quads):
    """(u, v) at the four corners of every quad of quad_mesh: (0,0) (1,0) (1,1) (0,1),
    u along the quad's first axis a, v along b."""
    return np.tile(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), (n_quads, 1))


# ------------------------------------------------------------------ fetch --
def test_fetch_texel_centres_linear_field_and_wrap():
    H, W = 5, 7
    j, i = np.mgrid[0:H, 0:W].astype(np.float32)
    tex = np.stack([0.25 + 0.5 * i, -1.0 + 0.125 * j, 3.0 * i - 2.0 * j, np.full_like(i, 0.5)], -1).astype(np.float32)
    # texel centres return the texel
    for jj in range(H):
        for ii in range(W):
            got = O.tex_fetch(tex, (ii + 0.5) / W, (jj + 0.5) / H)
            np.testing.assert_array_equal(got, tex[jj, ii])
    # a field linear in the texel index is reproduced between the centres
    rng = np.random.default_rng(3)
    for _ in range(200):
        tx, ty = rng.uniform(0, W - 1), rng.uniform(0, H - 1)
        got = O.tex_fetch(tex, (tx + 0.5) / W, (ty + 0.5) / H).astype(np.float64)
        want = np.array([0.25 + 0.5 * tx, -1.0 + 0.125 * ty, 3.0 * tx - 2.0 * ty, 0.5])
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=2e-5)
    # wrap: u and u + 1 (and v - 2) fetch the same; u = 0 mixes the first and last column
    for _ in range(50):
        # multiples of 2^-10, so that x + 1 and y - 2 are exact in float32
        x, y = rng.integers(0, 1024) / 1024.0, rng.integers(0, 1024) / 1024.0
        a = O.tex_fetch(tex, x, y)
        np.testing.assert_array_equal(a, O.tex_fetch(tex, x + 1.0, y - 2.0))
    edge = O.tex_fetch(tex, 0.0, 0.5 / H)
    np.testing.assert_allclose(edge, 0.5 * (tex[0, 0] + tex[0, W - 1]), rtol=1e-6)
    # a constant texture returns its value exactly, wherever it is read
    for c in (0.5, 0.1, 1.0 / 3.0, 0.8):
        const = np.full((3, 5, 4), c, np.float32)
        for _ in range(100):
            np.testing.assert_array_equal(O.tex_fetch(const, rng.uniform(-2, 2), rng.uniform(-2, 2)),
                                          np.float32(c))


# ------------------------------------------------------------------ alpha --
def _mt(o, d, p0, p1, p2):
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = e1 @ pv
    if det == 0:
        return np.inf, 0.0, 0.0
    tv = o - p0
    u = (tv @ pv) / det
    qv = np.cross(tv, e1)
    v = (d @ qv) / det
    t = (e2 @ qv) / det
    if u < 0 or v < 0 or u + v > 1 or t <= 0:
        return np.inf, 0.0, 0.0
    return t, u, v


def test_alpha_mask_band_lets_rays_through():
    # front quad at z = 0.1 with the (1, 1, 0, 0) mask, opaque quad at z = -0.1 behind it
    quads = bbox_mesh([((0.0, 0.0, 0.1), (0, 0, 1), 1.6), ((0.0, 0.0, -0.1), (0, 0, 1), 1.6)])
    m = quad_mesh(quads, material=[0, 0, 0, 1])
    mask = np.zeros((1, 4, 4), np.float32)
    mask[0, :2] = 1.0
    mask = np.repeat(mask, 2, axis=0)  # 2 rows of texels, identical
    m = textured(m, quad_uvs(len(quads)), [mask], [[-1, -1, -1, 0], [-1, -1, -1, -1]])
    f = O.MeshFilter(m)
    P = m["positions"].astype(np.float64)
    I = m["indices"]
    U = m["uvs"].astype(np.float64)
    rng = np.random.default_rng(5)
    n_band = n_front = 0
    for _ in range(400):
        o = np.array([rng.uniform(-0.75, 0.75), rng.uniform(-0.75, 0.75), 0.5])
        d = np.array([rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05), -1.0])
        d /= np.linalg.norm(d)
        tri, out = f.closest_hit(o, d, 5.0)
        # brute force: the front quad's triangles are 4 and 5, the back quad's 6 and 7
        best = None
        for t in (4, 5, 6, 7):
            th, b1, b2 = _mt(o, d, P[I[t, 0]], P[I[t, 1]], P[I[t, 2]])
            if not np.isfinite(th):
                continue
            uv = (1 - b1 - b2) * U[I[t, 0]] + b1 * U[I[t, 1]] + b2 * U[I[t, 2]]
            if t in (4, 5):
                if abs(uv[0] - 0.625) < 1e-3 or abs(uv[0] - 0.875) < 1e-3:
                    best = "skip"
                    break
                if 0.625 <= uv[0] <= 0.875:  # transparent band of the mask
                    n_band += 1
                    continue
                n_front += 1
            if best is None or th < best[1]:
                best = (t, th)
        if best == "skip":
            continue
        assert best is not None
        assert tri // 2 == best[0] // 2, (tri, best)
        assert abs(out[0] - best[1]) <= 1e-5
    assert n_band > 50 and n_front > 200


# ---------------------------------------------------------------- colours --
def test_textured_diffuse_is_the_mean_over_hits():
    # one plane through the grid; its diffuse texture is linear in the texel index,
    # and the quad's texture coordinates stay inside the texel centres' range
    W = H = 8
    j, i = np.mgrid[0:H, 0:W].astype(np.float32)
    tex = np.stack([0.1 + 0.1 * i, 0.2 + 0.05 * j, 0.9 - 0.05 * i - 0.04 * j, np.ones_like(i)], -1)
    quads = bbox_mesh([((0.03, -0.02, 0.01), (0, 0, 1), 1.8)])
    m = quad_mesh(quads)
    lo, hi = 0.5 / W, 1.0 - 0.5 / W
    uvs = quad_uvs(len(quads)) * (hi - lo) + lo
    m = textured(m, uvs, [tex], [[0, -1, -1, -1], [-1, -1, -1, -1]])
    f = O.MeshFilter(m)
    g = f.setup(0.25)
    rec = f.run(g, samples=32, iterations=2, seed=4)
    d = O.decode_records(rec)
    hit = np.nonzero(d["density"] > 0)[0]
    assert hit.size >= 20
    P = m["positions"].astype(np.float64)
    # the plane is quad 2: vertex 8 has (u, v) = (0, 0), vertex 9 (1, 0), vertex 11 (0, 1);
    # tc of a point on it is affine in position (the quad is a parallelogram)
    A = P[8]
    ea, eb = P[9] - A, P[11] - A
    G = np.array([[ea @ ea, ea @ eb], [ea @ eb, eb @ eb]])
    checked = 0
    for vid in hit[:40]:
        pr = f.probe(g, int(vid), samples=32, iterations=2, seed=4)
        sel = pr["tri"] >= 0
        pts = pr["o"][sel].astype(np.float64) + pr["t"][sel, None].astype(np.float64) * pr["d"][sel].astype(np.float64)
        ab = np.linalg.solve(G, np.stack([(pts - A) @ ea, (pts - A) @ eb]))
        uv = lo + (hi - lo) * ab.T  # (u, v) along a and b
        tx, ty = uv[:, 0] * W - 0.5, uv[:, 1] * H - 0.5
        col = np.stack([0.1 + 0.1 * tx, 0.2 + 0.05 * ty, 0.9 - 0.05 * tx - 0.04 * ty], -1).mean(axis=0)
        want = np.floor(255.0 * col)
        assert np.all(np.abs(d["diffuse"][vid].astype(np.float64) - want) <= 1), (vid, d["diffuse"][vid], want)
        # the specular colour stays the material constant
        np.testing.assert_array_equal(d["specular"][vid], np.floor(255.0 * MAT["mat_specular"][0]).astype(np.uint8))
        checked += 1
    assert checked >= 20


def test_texture_without_texture_coordinates_reads_tc_zero():
    W, H = 4, 3
    rng = np.random.default_rng(2)
    tex = rng.uniform(0.1, 0.9, (H, W, 4)).astype(np.float32)
    m = quad_mesh(bbox_mesh([((0.0, 0.0, 0.02), (0, 0, 1), 1.7)]))
    m["textures"] = [tex]
    m["mat_textures"] = np.array([[-1, 0, -1, -1], [-1, -1, -1, -1]], np.int32)
    f = O.MeshFilter(m)
    rec = f.run(f.setup(0.3), samples=16, iterations=2, seed=1)
    d = O.decode_records(rec)
    hit = d["density"] > 0
    assert hit.sum() >= 10
    corners = 0.25 * (tex[0, 0] + tex[0, W - 1] + tex[H - 1, 0] + tex[H - 1, W - 1])
    want = np.floor(255.0 * corners[:3].astype(np.float64))
    assert np.all(np.abs(d["specular"][hit].astype(np.float64) - want) <= 1)
    np.testing.assert_array_equal(d["diffuse"][hit], np.broadcast_to(
        np.floor(255.0 * MAT["mat_diffuse"][0].astype(np.float32)).astype(np.uint8), d["diffuse"][hit].shape))


# ------------------------------------------------------------- normal map --
# quad_mesh's plane with normal z has a = (0, 1, 0) and b = (-1, 0, 0): with
# (u, v) along (a, b), T = dP/du ~ a and B = dP/dv ~ b, so the frame (T, -B, N)
# is (y, x, z) and the normal nm.x y + nm.y x + nm.z z
@pytest.mark.parametrize("texel,normal", [
    ((0.5, 1.0, 0.5), (1.0, 0.0, 0.0)),
    ((1.0, 0.5, 0.5), (0.0, 1.0, 0.0)),
    ((0.5, 0.5, 1.0), (0.0, 0.0, 1.0)),
    ((0.5, 0.5, 0.5), (0.0, 0.0, 1.0)),   # nm = 0: the geometric normal
    ((0.8, 0.5, 0.9), (0.0, 0.6, 0.8)),
])
def test_normal_map_closed_form(texel, normal):
    tex = np.tile(np.array([*texel, 1.0], np.float32), (2, 2, 1))
    quads = bbox_mesh([((0.02, -0.01, 0.015), (0, 0, 1), 1.7)])
    m = textured(quad_mesh(quads), quad_uvs(len(quads)), [tex], [[-1, -1, 0, -1], [-1, -1, -1, -1]])
    f = O.MeshFilter(m)
    rec = f.run(f.setup(0.25), samples=32, iterations=2, seed=6)
    d = O.decode_records(rec)
    hit = d["density"] > 0
    assert hit.sum() >= 20
    sig, r, nn = _expected_plane(normal, 0)
    assert np.all(np.abs(d["sigma"][hit].astype(np.float64) - np.floor(sig)) <= 1)
    assert np.all(np.abs(d["r"][hit].astype(np.float64) - np.floor(r)) <= 1)
    assert np.all(np.abs(d["normal"][hit].astype(np.float64) - np.floor(nn)) <= 1)
