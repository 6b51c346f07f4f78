"""Image files of a post-processed frame (r2_image_write), host only (no GPU).

The reference writes every frame as output.jpg (stbi_write_jpg, quality 100) and
output.png (stbi_write_png, compression level 0), both vertically flipped
(visualisation/src/image_writer.cpp:8-21, raytracer/pipeline.cu:84-85).  stb is not
in this image, so the byte streams are not compared with stb's.  What these tests
pin, with PIL (libjpeg / zlib) as the independent decoder:

  * PNG: the decoded pixels equal the frame flipped top to bottom, bit for bit;
  * JPEG: at quality 100 the decoded pixels are no further from the frame than
    libjpeg's own quality-100 4:4:4 encoding of the same pixels is (the colour
    conversion and DCT rounding of both sides: a max error of 3 or 4 of 255 on
    most images), and at lower qualities within a PSNR that only a correct
    DCT, zig-zag order and Huffman coding reach;
  * the quantisation tables equal the ones libjpeg's own encoder writes at
    the same quality (the IJG scaling of the T.81 Annex K tables), and the
    chroma sampling is 4:4:4 above quality 90 and 4:2:0 at 90 and below;
  * extreme coefficients: black and white blocks side by side (DC differences
    of category 11) and a pixel checkerboard (the largest AC terms);
  * error codes for bad sizes, formats, qualities and paths.
"""
import ctypes as C

import numpy as np
import pytest
from PIL import Image, JpegImagePlugin

import renderer2_amd as r2
from renderer2_amd import capi


def _decode(path):
    return np.asarray(Image.open(path).convert("RGB")).astype(np.int64)


def libjpeg_q100_error(rgb: np.ndarray) -> np.ndarray:
    """|decoded - original| of libjpeg's own quality-100, 4:4:4 JPEG of `rgb` (rows
    top to bottom): the error a correct quality-100 encoder is held to."""
    import io

    buf = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(rgb, np.uint8)).save(buf, format="JPEG", quality=100, subsampling=0)
    buf.seek(0)
    return np.abs(np.asarray(Image.open(buf).convert("RGB")).astype(np.int64) - rgb.astype(np.int64))


def _psnr(a, b):
    mse = float(((a - b) ** 2).mean())
    return 99.0 if mse == 0 else 10 * np.log10(255.0 ** 2 / mse)


def _smooth(h, w):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    r = 128 + 100 * np.sin(xx / 9.0) * np.cos(yy / 13.0)
    g = 255 * yy / max(h - 1, 1)
    b = 60 + 120 * (xx + yy) / (w + h)
    return np.clip(np.stack([r, g, b], -1), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("h,w", [(1, 1), (1, 100), (37, 53), (64, 64)])
def test_png_exact(tmp_path, h, w):
    img = np.random.default_rng(h * 1000 + w).integers(0, 256, (h, w, 3), dtype=np.uint8)
    p = tmp_path / "output.png"
    r2.image_write(p, img, "png")
    dec = np.asarray(Image.open(p))
    assert dec.shape == (h, w, 3)
    assert np.array_equal(dec, np.flipud(img))


@pytest.mark.parametrize("h,w", [(1, 1), (8, 8), (9, 17), (37, 53), (100, 3)])
def test_jpg_quality_100(tmp_path, h, w):
    rng = np.random.default_rng(h * 7 + w)
    for img in (rng.integers(0, 256, (h, w, 3), dtype=np.uint8), _smooth(h, w)):
        p = tmp_path / "output.jpg"
        r2.image_write(p, img, "jpg", 100)
        im = Image.open(p)
        assert im.size == (w, h) and im.mode == "RGB"
        assert JpegImagePlugin.get_sampling(im) == 0  # 4:4:4
        assert all(v == 1 for t in im.quantization.values() for v in t)
        err = np.abs(_decode(p) - np.flipud(img).astype(np.int64))
        lib = libjpeg_q100_error(np.flipud(img))
        assert err.max() <= max(lib.max(), 3) and err.mean() <= lib.mean() + 0.1, (err.max(), err.mean(), lib.max())


@pytest.mark.parametrize("q", [1, 10, 25, 50, 75, 90, 91, 95, 100])
def test_jpg_tables_match_libjpeg(tmp_path, q):
    img = _smooth(24, 40)
    ours, theirs = tmp_path / "a.jpg", tmp_path / "b.jpg"
    r2.image_write(ours, img, "jpg", q)
    Image.fromarray(np.flipud(img)).save(theirs, quality=q, subsampling=0 if q > 90 else 2)
    a, b = Image.open(ours), Image.open(theirs)
    assert JpegImagePlugin.get_sampling(a) == (0 if q > 90 else 2)
    assert sorted(a.quantization) == sorted(b.quantization) == [0, 1]
    for k in (0, 1):
        assert list(a.quantization[k]) == list(b.quantization[k])


@pytest.mark.parametrize("q", [50, 75, 90, 95])
def test_jpg_lossy_quality(tmp_path, q):
    """Same tables and sampling as libjpeg's encoder, so the PSNR must be close to
    libjpeg's on the same image; a wrong zig-zag, DCT basis or chroma placement
    loses tens of dB."""
    img = _smooth(61, 83)
    ours, theirs = tmp_path / "a.jpg", tmp_path / "b.jpg"
    r2.image_write(ours, img, "jpg", q)
    Image.fromarray(np.flipud(img)).save(theirs, quality=q, subsampling=0 if q > 90 else 2)
    ref = np.flipud(img).astype(np.int64)
    pa, pb = _psnr(_decode(ours), ref), _psnr(_decode(theirs), ref)
    assert pa >= pb - 1.0, (pa, pb)
    assert pa >= 35.0


def test_jpg_extreme_coefficients(tmp_path):
    img = np.zeros((16, 32, 3), np.uint8)
    img[:, 8:16] = 255                        # black | white | ... blocks: DC steps of 2040
    img[:, 24:32] = 255
    yy, xx = np.mgrid[0:16, 16:24]
    img[:, 16:24] = (((xx + yy) & 1) * 255)[..., None]  # pixel checkerboard: largest AC terms
    p = tmp_path / "x.jpg"
    for q in (100, 90):
        r2.image_write(p, img, "jpg", q)
        err = np.abs(_decode(p) - np.flipud(img).astype(np.int64))
        assert err.max() <= (1 if q == 100 else 4), err.max()  # libjpeg's own encoder: 0 and 2


def test_jpg_markers(tmp_path):
    p = tmp_path / "m.jpg"
    r2.image_write(p, _smooth(20, 30), "jpg", 100)
    b = p.read_bytes()
    assert b[:2] == b"\xff\xd8" and b[-2:] == b"\xff\xd9"
    assert b[2:4] == b"\xff\xe0" and b[6:11] == b"JFIF\x00"
    i = b.index(b"\xff\xc0")                   # baseline SOF0: 8-bit, h, w, 3 components
    assert b[i + 4] == 8 and int.from_bytes(b[i + 5:i + 7], "big") == 20
    assert int.from_bytes(b[i + 7:i + 9], "big") == 30 and b[i + 9] == 3


INVALID, IO = -1, -5  # R2_ERR_INVALID, R2_ERR_IO


def test_image_errors(tmp_path):
    lib = capi.lib()
    img = np.zeros((4, 4, 3), np.uint8)
    err = C.create_string_buffer(256)
    path = str(tmp_path / "e.jpg").encode()
    for fmt, q in ((2, 100), (capi.R2_IMAGE_JPG, 0), (capi.R2_IMAGE_JPG, 101)):
        assert lib.r2_image_write(path, 4, 4, img.ctypes.data, fmt, q, err, 256) == INVALID
    assert lib.r2_image_write(path, 0, 4, img.ctypes.data, capi.R2_IMAGE_PNG, 0, err, 256) == INVALID
    assert lib.r2_image_write(path, 4, 4, None, capi.R2_IMAGE_PNG, 0, err, 256) == INVALID
    wide = np.zeros((1, 65536, 3), np.uint8)
    assert lib.r2_image_write(path, 65536, 1, wide.ctypes.data, capi.R2_IMAGE_JPG, 100, err, 256) == INVALID
    bad = str(tmp_path / "no" / "such" / "dir" / "e.png").encode()
    assert lib.r2_image_write(bad, 4, 4, img.ctypes.data, capi.R2_IMAGE_PNG, 0, err, 256) == IO
    assert err.value
    with pytest.raises(ValueError):
        r2.image_write(tmp_path / "e.bmp", img, "bmp")
