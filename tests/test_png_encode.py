"""PNG output (tiny-skia Pixmap::encode_png, crates/resvg/src/main.rs:104-117): the
oracle's scanlines, decoded by an independent decoder (PIL), give back the
demultiplied pixmap; on the reference's own fixtures they give back the reference's
PNGs pixel for pixel; the per-row filter choice follows the minimum-sum rule."""
import io
import zlib

import numpy as np
import pytest
from PIL import Image

from fixture_scenes import demultiply_f64, load_png
from oracle import bindings as orc
from resvg_amd import png
from scene_fixtures import all_scene_fixtures, fixture_ts


def random_premultiplied(rng, h, w):
    a = rng.integers(0, 256, (h, w, 1))
    a[rng.random((h, w, 1)) < 0.2] = 255
    a[rng.random((h, w, 1)) < 0.1] = 0
    c = rng.integers(0, 256, (h, w, 3)) * a // 255
    return np.concatenate([c, a], axis=2).astype(np.uint8)


def decode(data: bytes) -> np.ndarray:
    im = Image.open(io.BytesIO(data))
    assert im.mode == "RGBA"
    return np.asarray(im).copy()


def oracle_png(px: np.ndarray) -> bytes:
    h, w = px.shape[:2]
    return png.from_rows(w, h, orc.png_filter_rows(px).tobytes())


@pytest.mark.parametrize("h,w", [(1, 1), (1, 9), (7, 1), (23, 37), (64, 300)])
def test_rows_decode_to_the_demultiplied_pixmap(h, w):
    px = random_premultiplied(np.random.default_rng(h * 1000 + w), h, w)
    assert np.array_equal(decode(oracle_png(px)), demultiply_f64(px))


def _filters(cur: np.ndarray, prev: np.ndarray) -> np.ndarray:
    """The five filtered versions (None, Sub, Up, Average, Paeth) of one row of
    bytes, PNG specification section 9.2."""
    x = cur.astype(np.int32)
    a = np.concatenate([np.zeros(4, np.int32), x[:-4]])
    b = prev.astype(np.int32)
    c = np.concatenate([np.zeros(4, np.int32), b[:-4]])
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    paeth = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
    return np.stack([x, x - a, x - b, x - (a + b) // 2, x - paeth]) & 0xFF


def test_each_row_takes_the_filter_of_least_signed_sum():
    rng = np.random.default_rng(1)
    yy, xx = np.mgrid[0:40, 0:50]
    images = [random_premultiplied(rng, 12, 31)]
    # smooth images, on which each predictor wins somewhere
    for f in (xx * 3, yy * 5, (xx + yy) * 2, (xx * yy) % 256, (xx * 7 + (yy % 3) * 40)):
        rgb = np.stack([f % 256, (f // 2) % 256, (255 - f) % 256], axis=-1)
        images.append(np.concatenate([rgb, np.full(f.shape + (1,), 255)], axis=-1).astype(np.uint8))
    # each value the mean of its left and upper neighbours plus a little noise, with
    # random first row and column: Average predicts it, Sub, Up and Paeth do not
    v = np.zeros((30, 40, 3), np.int64)
    v[0], v[:, 0] = rng.integers(0, 256, (40, 3)), rng.integers(0, 256, (30, 3))
    for y in range(1, 30):
        for x in range(1, 40):
            v[y, x] = np.clip((v[y, x - 1] + v[y - 1, x]) // 2 + rng.integers(-1, 2, 3), 0, 255)
    images.append(np.concatenate([v, np.full((30, 40, 1), 255)], axis=-1).astype(np.uint8))
    chosen = set()
    for px in images:
        h, w = px.shape[:2]
        rows = orc.png_filter_rows(px).reshape(h, 4 * w + 1)
        straight = demultiply_f64(px).reshape(h, 4 * w)
        prev = np.zeros(4 * w, np.uint8)
        for y in range(h):
            f = _filters(straight[y], prev)
            sums = np.abs(f.astype(np.uint8).view(np.int8).astype(np.int64)).sum(axis=1)
            t = int(np.argmin(sums))  # the first of equal sums
            assert rows[y, 0] == t
            assert np.array_equal(rows[y, 1:], f[t].astype(np.uint8))
            chosen.add(t)
            prev = straight[y]
    assert chosen == {0, 1, 2, 3, 4}


EXACT_FIXTURES = ["feGaussianBlur/small-stdDeviation", "feGaussianBlur/simple-case", "painting/opacity/50percent",
                  "masking/mask/simple-case", "feTurbulence/type=fractalNoise", "feTurbulence/seed=20",
                  "feComposite/operator=arithmetic", "painting/mix-blend-mode/hue"]


@pytest.mark.parametrize("name", EXACT_FIXTURES)
def test_reference_pngs_are_the_encoded_renders(name):
    """The fixture PNGs are resvg's renders written by encode_png (then recompressed
    by oxipng, which keeps the pixels): the oracle's render encoded the same way
    decodes to the same pixels (the 3-pixel frame band and pixels transparent in both
    left out, as the reference's comparison does)."""
    sc = all_scene_fixtures()[name]
    exp = load_png(name)
    h, w = exp.shape[:2]
    px = np.zeros((h, w, 4), np.uint8)
    orc.render(sc.to_abi().abi, fixture_ts(name), px)
    got = decode(oracle_png(px)).astype(np.int32)
    exp = exp.astype(np.int32)
    differ = (got != exp).any(axis=-1) & ~((got[..., 3] == 0) & (exp[..., 3] == 0))
    semi = (px[..., 3] > 0) & (px[..., 3] < 255)
    assert int(differ[3:-3, 3:-3].sum()) == 0
    assert int(semi[3:-3, 3:-3].sum()) > 10000  # the demultiply is exercised


def test_container_layout():
    data = png.from_rows(2, 1, bytes([0, 1, 2, 3, 4, 5, 6, 7, 8]))
    assert data[:8] == png.SIGNATURE and data[12:16] == b"IHDR" and data.endswith(png.chunk(b"IEND", b""))
    with pytest.raises(ValueError):
        png.from_rows(2, 2, bytes(9))


@pytest.mark.parametrize("n", [0, 1, (1 << 20) - 1, 1 << 20, (1 << 20) + 1, 5 * (1 << 20) + 12345])
def test_parallel_deflate_is_one_zlib_stream(n):
    rng = np.random.default_rng(n)
    data = (np.cumsum(rng.integers(-3, 4, n)) % 256).astype(np.uint8).tobytes()
    z = png.deflate(data, 6, workers=4)
    assert zlib.decompress(z) == data
    d = zlib.decompressobj()
    assert d.decompress(z) == data and d.eof and d.unused_data == b""


def test_large_png_decodes():
    rng = np.random.default_rng(3)
    px = random_premultiplied(rng, 700, 600)  # 1.68 MB of rows: several deflate pieces
    assert np.array_equal(decode(oracle_png(px)), demultiply_f64(px))
