"""FLIP metric (renderer2_amd/flip.py): invariants of the published metric."""
import numpy as np

from renderer2_amd import flip as F


def _img(seed=0, h=48, w=64):
    rng = np.random.default_rng(seed)
    base = np.linspace(0, 1, w)[None, :, None] * np.linspace(0.2, 1, h)[:, None, None] * np.array([1.0, 0.6, 0.3])
    return np.clip(base + 0.05 * rng.standard_normal((h, w, 3)), 0, 1)


def test_identical_images_score_zero():
    a = _img()
    m, e = F.flip(a, a)
    assert m == 0.0 and e.shape == a.shape[:2]


def test_range_and_monotone_in_error():
    a = _img()
    rng = np.random.default_rng(1)
    noise = rng.standard_normal(a.shape)
    vals = [F.flip(a, np.clip(a + s * noise, 0, 1))[0] for s in (0.01, 0.03, 0.1, 0.3)]
    assert all(0.0 <= v <= 1.0 for v in vals)
    assert all(x < y for x, y in zip(vals, vals[1:]))


def test_black_white_is_large_and_symmetric_in_colour_error():
    b, w = np.zeros((32, 32, 3)), np.ones((32, 32, 3))
    m, _ = F.flip(b, w)
    assert m > 0.9
    a = _img(2)
    c = a.copy()
    c[..., 0] = np.clip(c[..., 0] + 0.2, 0, 1)
    assert F.flip(a, c)[0] > F.flip(a, np.clip(a + 0.02, 0, 1))[0]


def test_srgb_roundtrip_and_hdr_helper():
    x = np.linspace(0, 1, 101)
    np.testing.assert_allclose(F.linear_to_srgb(F.srgb_to_linear(x)), x, atol=1e-12)
    h = np.full((16, 16, 3), 0.25)
    assert F.flip_hdr(h, h) == 0.0
    assert F.flip_hdr(h, 1.5 * h) > 0.0
