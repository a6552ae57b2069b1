"""Image I/O of the driver layer: the BMP mask loader (the reference's load_mask_from_image,
lbm_smemaa_main.py:53-68, without OpenCV) and the PPM dump of lbm_visualize output."""
import struct

import numpy as np
import pytest

from lbm import geometry as G


def _write_bmp(path, bgr, top_down=False, bpp=24):
    """Minimal uncompressed BMP writer (test fixture)."""
    h, w, _ = bgr.shape
    nb = bpp // 8
    stride = (w * nb + 3) & ~3
    rows = bgr if top_down else bgr[::-1]
    pix = bytearray()
    for r in rows:
        line = r.astype(np.uint8)
        if nb == 4:
            line = np.concatenate([line, np.full((w, 1), 255, np.uint8)], 1)
        b = line.tobytes()
        pix += b + b"\0" * (stride - len(b))
    hdr = struct.pack("<2sIHHI", b"BM", 54 + len(pix), 0, 0, 54)
    dib = struct.pack("<IiiHHIIiiII", 40, w, -h if top_down else h, 1, bpp, 0, len(pix), 2835, 2835, 0, 0)
    open(path, "wb").write(hdr + dib + bytes(pix))


def _disc_image(w, h, cx, cy, r):
    img = np.full((h, w, 3), 255, np.uint8)
    yy, xx = np.mgrid[0:h, 0:w]
    img[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] = 0
    return img


@pytest.mark.parametrize("top_down,bpp", [(False, 24), (True, 24), (False, 32)])
def test_read_bmp_round_trip(tmp_path, top_down, bpp):
    img = _disc_image(37, 21, 12, 9, 5)
    img[0, 0] = (10, 20, 30)  # B, G, R at the top-left pixel
    p = tmp_path / "a.bmp"
    _write_bmp(p, img, top_down, bpp)
    back = G.read_bmp(p)
    assert back.shape == (21, 37, 3)
    assert np.array_equal(back, img)


def test_mask_from_image_same_size(tmp_path):
    W, H = 64, 32
    img = _disc_image(W, H, 20, 16, 6)
    p = tmp_path / "m.bmp"
    _write_bmp(p, img)
    m = G.mask_from_image(p, W, H)
    want = (img[:, :, 1] == 0).astype(np.uint8)
    want[0, :] = want[H - 1, :] = 1
    want[1:H - 1, 0] = want[1:H - 1, W - 1] = 0
    assert np.array_equal(m, want)


def test_mask_from_image_resized(tmp_path):
    """A disc drawn at half resolution and resized 2x keeps its centre and roughly its area (bilinear
    geometry; OpenCV's exact fixed-point weights are not reproduced)."""
    img = _disc_image(64, 32, 20, 16, 6)
    p = tmp_path / "m.bmp"
    _write_bmp(p, img)
    m = G.mask_from_image(p, 128, 64)
    inner = m[1:-1, 1:-1].astype(bool)
    ys, xs = np.nonzero(inner)
    assert abs(xs.mean() + 1 - 40.5) < 1.5 and abs(ys.mean() + 1 - 32.5) < 1.5
    area = inner.sum()
    assert 0.6 * np.pi * 12 ** 2 < area < 1.2 * np.pi * 12 ** 2


def test_resize_identity_and_constant():
    img = np.random.default_rng(0).integers(0, 256, (9, 7, 3), dtype=np.uint8)
    assert np.array_equal(G.resize_linear(img, 7, 9), img)
    const = np.full((5, 5, 3), 77, np.uint8)
    assert np.all(G.resize_linear(const, 13, 11) == 77)


def test_write_ppm(tmp_path):
    rgba = np.zeros((3, 4, 4), np.uint8)
    rgba[..., 0] = np.arange(12).reshape(3, 4)
    rgba[..., 3] = 255
    p = tmp_path / "v.ppm"
    G.write_ppm(p, rgba)
    data = open(p, "rb").read()
    assert data.startswith(b"P6\n4 3\n255\n")
    body = np.frombuffer(data[len(b"P6\n4 3\n255\n"):], np.uint8).reshape(3, 4, 3)
    assert np.array_equal(body, rgba[..., :3])
