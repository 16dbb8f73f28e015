"""rvk_png_filter_rows (tiny-skia Pixmap::encode_png's demultiply and scanline
filtering) against the oracle, through the C-ABI: byte for byte on random pixmaps of
widths below, at and above one block's 256 pixels, on a rendered document, and the
whole PNG file (Context.encode_png) equal to the oracle's and decoding to the
demultiplied pixmap."""
import io

import numpy as np
import pytest
from PIL import Image

from fixture_scenes import demultiply_f64
from oracle import bindings as orc
from resvg_amd import abi, png
from resvg_amd.workloads import blur_offset_composite_doc

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _premul(h, w, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, 1))
    a[rng.random((h, w, 1)) < 0.2] = 255
    a[rng.random((h, w, 1)) < 0.1] = 0
    c = rng.integers(0, 256, (h, w, 3)) * a // 255
    return np.concatenate([c, a], axis=2).astype(np.uint8)


def _rows(ctx, px):
    import torch
    h, w = px.shape[:2]
    out = torch.full((h * (4 * w + 1),), 77, dtype=torch.uint8, device="cuda")
    ctx.png_filter_rows(_dev(px), out)
    ctx.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("h,w", [(1, 1), (5, 3), (3, 255), (4, 256), (6, 257), (7, 1000), (9, 4096)])
def test_rows_match_the_oracle(ctx, h, w):
    px = _premul(h, w, h * 7919 + w)
    print(f"png rows {h}x{w}", flush=True)
    assert np.array_equal(_rows(ctx, px), orc.png_filter_rows(px))


def test_rendered_document_and_file(ctx):
    import torch
    sc, (w, h) = blur_offset_composite_doc(seed=2, size=512)
    px = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    ctx.render(sc, abi.Transform.identity(), px)
    ctx.synchronize()
    host = px.cpu().numpy()
    rows = orc.png_filter_rows(host)
    assert np.array_equal(_rows(ctx, host), rows)
    data = ctx.encode_png(px)
    assert data == png.from_rows(w, h, rows.tobytes())
    assert np.array_equal(np.asarray(Image.open(io.BytesIO(data))), demultiply_f64(host))
    # rows of more than one filter type (the adaptive choice is exercised)
    assert len(set(rows.reshape(h, 4 * w + 1)[:, 0].tolist())) >= 2
