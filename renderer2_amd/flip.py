"""LDR FLIP image difference (Andersson, Nilsson, Akenine-Moller, Oskarsson,
Astrom, Fairchild: "FLIP: A Difference Evaluator for Alternating Images",
HPG 2020), written from the paper for reporting rendered frames against the
FP64 oracle.  The reference repository measures FLIP with a Python script that
needs OpenCV/OpenEXR (absent here), so values are pinned by the metric's own
invariants (0 for identical images, range [0, 1], monotone in added error).

Pipeline: sRGB -> linear RGB -> YCxCz; per-channel contrast-sensitivity
filtering; back to linear RGB (clamped), CIELAB, Hunt adjustment, HyAB distance,
raised to qc = 0.7 and redistributed (pc = 0.4, pt = 0.95) against the
green-blue distance; edge and point features of the achromatic channel give the
feature difference (qf = 0.5); the error is dE_c ^ (1 - dE_f).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

QC, PC, PT, QF = 0.7, 0.4, 0.95, 0.5
WHITE = np.array([0.950428545, 1.0, 1.088900371])
RGB2XYZ = np.array([[0.4124564, 0.3575761, 0.1804375],
                    [0.2126729, 0.7151522, 0.0721750],
                    [0.0193339, 0.1191920, 0.9503041]])
XYZ2RGB = np.linalg.inv(RGB2XYZ)


def srgb_to_linear(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def linear_to_srgb(c):
    c = np.clip(np.asarray(c, np.float64), 0.0, 1.0)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * c ** (1 / 2.4) - 0.055)


def hdr_to_ldr(hdr, exposure: float = 0.0):
    """Fixed exposure, clamp, sRGB encode."""
    return linear_to_srgb(np.asarray(hdr, np.float64) * 2.0 ** exposure)


def _xyz_to_ycxcz(xyz):
    n = xyz / WHITE
    return np.stack([116.0 * n[..., 1] - 16.0, 500.0 * (n[..., 0] - n[..., 1]), 200.0 * (n[..., 1] - n[..., 2])], -1)


def _ycxcz_to_xyz(y):
    yn = (y[..., 0] + 16.0) / 116.0
    return np.stack([y[..., 1] / 500.0 + yn, yn, yn - y[..., 2] / 200.0], -1) * WHITE


def _lab(rgb_lin):
    n = (rgb_lin @ RGB2XYZ.T) / WHITE
    d = 6.0 / 29.0
    f = np.where(n > d ** 3, np.cbrt(n), n / (3 * d * d) + 4.0 / 29.0)
    return np.stack([116.0 * f[..., 1] - 16.0, 500.0 * (f[..., 0] - f[..., 1]), 200.0 * (f[..., 1] - f[..., 2])], -1)


def _hunt(lab):
    return np.stack([lab[..., 0], 0.01 * lab[..., 0] * lab[..., 1], 0.01 * lab[..., 0] * lab[..., 2]], -1)


def _hyab(a, b):
    d = a - b
    return np.abs(d[..., 0]) + np.sqrt(d[..., 1] ** 2 + d[..., 2] ** 2)


def _csf_kernels(ppd):
    params = [(1.0, 0.0047, 0.0, 1e-5), (1.0, 0.0053, 0.0, 1e-5), (34.1, 0.04, 13.5, 0.025)]
    bmax = max(max(p[1], p[3]) for p in params)
    r = int(np.ceil(3.0 * np.sqrt(bmax / (2.0 * np.pi ** 2)) * ppd))
    x, y = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1))
    z = (x / ppd) ** 2 + (y / ppd) ** 2
    ks = []
    for a1, b1, a2, b2 in params:
        g = a1 * np.sqrt(np.pi / b1) * np.exp(-np.pi ** 2 * z / b1) + a2 * np.sqrt(np.pi / b2) * np.exp(-np.pi ** 2 * z / b2)
        ks.append(g / g.sum())
    return ks


def _feature_kernels(ppd):
    sd = 0.5 * 0.082 * ppd
    r = int(np.ceil(3.0 * sd))
    x, y = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1))
    g = np.exp(-(x ** 2 + y ** 2) / (2.0 * sd * sd))
    out = []
    for k in (-x * g, (x ** 2 / (sd * sd) - 1.0) * g):
        k = k.astype(np.float64)
        pos, neg = k[k > 0].sum(), -k[k < 0].sum()
        k[k > 0] /= pos
        k[k < 0] /= neg
        out.append(k)
    return out


def _features(ach, ppd):
    edge, point = _feature_kernels(ppd)
    res = []
    for k in (edge, point):
        fx = ndimage.convolve(ach, k, mode="nearest")
        fy = ndimage.convolve(ach, k.T, mode="nearest")
        res.append(np.sqrt(fx * fx + fy * fy))
    return res


def flip(reference_srgb, test_srgb, ppd: float = 67.0):
    """Mean LDR-FLIP and the per-pixel map for two sRGB images in [0, 1], shape (H, W, 3)."""
    ref = np.clip(np.asarray(reference_srgb, np.float64), 0.0, 1.0)
    tst = np.clip(np.asarray(test_srgb, np.float64), 0.0, 1.0)
    ref_y = _xyz_to_ycxcz(srgb_to_linear(ref) @ RGB2XYZ.T)
    tst_y = _xyz_to_ycxcz(srgb_to_linear(tst) @ RGB2XYZ.T)
    ks = _csf_kernels(ppd)
    filt = []
    for img in (ref_y, tst_y):
        f = np.stack([ndimage.convolve(img[..., c], ks[c], mode="nearest") for c in range(3)], -1)
        lin = np.clip(_ycxcz_to_xyz(f) @ XYZ2RGB.T, 0.0, 1.0)
        filt.append(_hunt(_lab(lin)))
    cmax = _hyab(_hunt(_lab(np.array([0.0, 1.0, 0.0]))), _hunt(_lab(np.array([0.0, 0.0, 1.0])))) ** QC
    dc = _hyab(filt[0], filt[1]) ** QC
    pcc = PC * cmax
    dc = np.where(dc < pcc, PT / pcc * dc, PT + (dc - pcc) / (cmax - pcc) * (1.0 - PT))
    fr = _features((ref_y[..., 0] + 16.0) / 116.0, ppd)
    ft = _features((tst_y[..., 0] + 16.0) / 116.0, ppd)
    df = (np.maximum(np.abs(fr[0] - ft[0]), np.abs(fr[1] - ft[1])) / np.sqrt(2.0)) ** QF
    err = np.clip(dc, 0.0, 1.0) ** (1.0 - df)
    return float(err.mean()), err


def flip_hdr(reference_hdr, test_hdr, exposure: float = 0.0, ppd: float = 67.0) -> float:
    """Mean LDR-FLIP after a fixed exposure and sRGB encoding of linear HDR frames."""
    return flip(hdr_to_ldr(reference_hdr, exposure), hdr_to_ldr(test_hdr, exposure), ppd)[0]
