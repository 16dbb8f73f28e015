"""The reference's rasterizer-independent filter fixtures, restated as filter graphs.

Each fixture SVG (crates/resvg/tests/tests/filters/<name>.svg) is a 200x200 viewBox
rendered at 300 px (scale 1.5, crates/resvg/tests/integration/main.rs:77-89) with a
filtered, pixel-aligned solid rect.  ``render_filtered_rect`` replays render_group
(crates/resvg/src/render.rs:49-143) for that single group with a pluggable backend
(oracle on numpy, or the HIP library on torch tensors), and ``compare_png`` applies
the reference's fixture rule (main.rs:151-226): demultiply in f64, skip pixels that
are transparent in both, flag any channel differing by more than 1.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from resvg_amd import abi
from resvg_amd.filters import SOURCE_GRAPHIC, FilterGraph

PNG_DIR = Path(__file__).resolve().parent / "golden" / "resvg_png"
SCALE = 1.5
SIZE = 300

SEAGREEN = (46, 139, 87)
LIGHTBLUE = (173, 216, 230)
BLUE = (0, 0, 255)
GREEN = (0, 128, 0)
RED = (255, 0, 0)
BLACK = (0, 0, 0)
AABBCC = (170, 187, 204)


@dataclass
class Fixture:
    name: str
    rect: tuple  # user-space rect (x, y, w, h)
    color: tuple
    graph: FilterGraph
    # user-space rects that the fixture SVG strokes (1 unit, black) over the result
    guides: tuple = ()


def region_for(rect):
    """usvg's default filter region: -10%/120% of the element bbox (objectBoundingBox)."""
    x, y, w, h = rect
    return (x - 0.1 * w, y - 0.1 * h, 1.2 * w, 1.2 * h)


def _f32(v):
    return float(np.float32(v))


def _region(rect):
    return tuple(_f32(v) for v in region_for(rect))


def fixtures() -> list[Fixture]:
    out = []
    R1 = (20.0, 20.0, 160.0, 160.0)

    def fx(name, rect, color, build):
        g = FilterGraph(_region(rect))
        build(g)
        out.append(Fixture(name, rect, color, g))

    fx("feGaussianBlur/simple-case", R1, SEAGREEN, lambda g: g.gaussian_blur(4.0))
    fx("feGaussianBlur/small-stdDeviation", R1, SEAGREEN, lambda g: g.gaussian_blur(1.0))
    fx("feGaussianBlur/stdDeviation=5-0", (40.0, 40.0, 120.0, 120.0), SEAGREEN, lambda g: g.gaussian_blur(5.0, 0.0))
    fx("feOffset/simple-case", R1, SEAGREEN, lambda g: g.offset(20.0, 40.0))
    fx("feOffset/negative-offset", R1, SEAGREEN, lambda g: g.offset(-20.0, -40.0))

    def comp(op, flood, k=(0, 0, 0, 0), flood_rect=None):
        def b(g):
            kw = {}
            if flood_rect is not None:
                kw["rect"] = flood_rect
            f = g.flood(flood, **kw)
            g.composite(op, k=k, in1=f, in2=SOURCE_GRAPHIC)
        return b

    fx("feComposite/operator=arithmetic", R1, SEAGREEN, comp(abi.COMPOSITE_ARITHMETIC, BLUE, (0.1, 0.2, 0.3, 0.4)))
    fx("feComposite/operator=over", R1, SEAGREEN, comp(abi.COMPOSITE_OVER, LIGHTBLUE, flood_rect=(50.0, 50.0, 100.0, 100.0)))
    fx("feComposite/operator=in", R1, RED, comp(abi.COMPOSITE_IN, SEAGREEN))
    fx("feComposite/operator=xor", R1, RED, comp(abi.COMPOSITE_XOR, SEAGREEN))

    def blend(mode, flood):
        def b(g):
            f = g.flood(flood)
            g.blend(mode, in1=f, in2=SOURCE_GRAPHIC)
        return b

    fx("feBlend/mode=normal", R1, RED, blend("source_over", SEAGREEN))
    fx("feBlend/mode=multiply", R1, SEAGREEN, blend("multiply", LIGHTBLUE))
    fx("feBlend/mode=screen", R1, SEAGREEN, blend("screen", GREEN))
    fx("feBlend/mode=darken", R1, SEAGREEN, blend("darken", LIGHTBLUE))
    fx("feBlend/mode=lighten", R1, SEAGREEN, blend("lighten", BLUE))
    fx("feBlend/mode=hue", R1, SEAGREEN, blend("hue", BLUE))
    fx("feBlend/mode=color-burn", R1, SEAGREEN, blend("color_burn", BLUE))
    fx("feFlood/seagreen", R1, BLACK, lambda g: g.flood(SEAGREEN))

    def ct(fb):
        return lambda g: g.component_transfer([None, None, fb, None])

    fx("feComponentTransfer/type=gamma-on-blue", R1, AABBCC, ct(("gamma", 2.0, 1.0, 0.5)))
    fx("feComponentTransfer/type=table-on-blue", R1, AABBCC, ct(("table", [1, 1, 0, 0])))
    fx("feComponentTransfer/type=linear-on-blue", R1, AABBCC, ct(("linear", 0.5, 0.25)))
    fx("feComponentTransfer/type=discrete-on-blue", R1, AABBCC, ct(("discrete", [1, 1, 0, 0])))

    R2 = (50.0, 50.0, 100.0, 100.0)

    def turb(bfx, bfy, octaves=1, seed=0, stitch=False, fractal=False):
        return lambda g: g.turbulence(bfx, bfy, octaves, seed, stitch, fractal)

    fx("feTurbulence/type=fractalNoise", R2, BLACK, turb(0.1, 0.1, fractal=True))
    fx("feTurbulence/no-attributes", R2, BLACK, turb(0.0, 0.0))
    fx("feTurbulence/seed=20", R2, BLACK, turb(0.05, 0.05, seed=20))
    fx("feTurbulence/stitchTiles=stitch", R2, BLACK, turb(0.03, 0.03, stitch=True))
    fx("feTurbulence/numOctaves=5", R2, BLACK, turb(0.05, 0.05, octaves=5))
    fx("feTurbulence/baseFrequency=0.05-0.01", R2, BLACK, turb(0.05, 0.01))
    _tile_fixtures(out)
    return out


def _tile_fixtures(out):
    """filters/feTile/*: feFlood on a small subregion, feOffset by (5, 5), feTile."""
    def graph(region, flood_rect, color, tile_rect=None):
        g = FilterGraph(tuple(_f32(v) for v in region))
        f = g.flood(color, rect=flood_rect, cs=abi.CI_SRGB)
        o = g.offset(5.0, 5.0, in1=f, cs=abi.CI_SRGB)
        g.tile(in1=o, rect=tile_rect, cs=abi.CI_SRGB)
        return g

    r = (40.0, 40.0, 120.0, 120.0)
    out.append(Fixture("feTile/simple-case", r, BLACK, graph(region_for(r), (28.0, 28.0, 10.0, 10.0), SEAGREEN),
                       guides=((28.0, 28.0, 144.0, 144.0), (28.0, 28.0, 10.0, 10.0), r, (40.0, 40.0, 10.0, 10.0))))
    # flood outside the filter region: the tile is empty, so the filter fails and the
    # layer is cleared
    out.append(Fixture("feTile/empty-region", r, BLACK, graph(region_for(r), (0.0, 0.0, 10.0, 10.0), RED)))
    # an empty group with a userSpaceOnUse filter region: nothing drawn under it
    reg = (50.0, 50.0, 100.0, 100.0)
    out.append(Fixture("feTile/with-region", (50.0, 50.0, 0.0, 0.0), BLACK,
                       graph(reg, (50.0, 50.0, 10.0, 10.0), SEAGREEN), guides=(reg,)))


def guide_mask(guides, size: int = SIZE) -> np.ndarray:
    """Pixels touched by the 1-unit strokes of ``guides`` (with a 1 px margin)."""
    m = np.zeros((size, size), dtype=bool)
    for x, y, w, h in guides:
        l, t, r, b = (v * SCALE for v in (x, y, x + w, y + h))
        lo = lambda e: max(int(math.floor(e - 0.75)) - 1, 0)  # noqa: E731
        hi = lambda e: min(int(math.ceil(e + 0.75)) + 1, size)  # noqa: E731
        m[lo(t):hi(b), lo(l):hi(l)] = m[lo(t):hi(b), lo(r):hi(r)] = True
        m[lo(t):hi(t), lo(l):hi(r)] = m[lo(b):hi(b), lo(l):hi(r)] = True
    return m


def layer_geometry(fx: Fixture):
    """ibbox and layer transform of the filtered group (render.rs:49-143)."""
    rx, ry, rw, rh = fx.graph.rect
    s = np.float32(SCALE)
    l, t = np.float32(rx) * s, np.float32(ry) * s
    r, b = (np.float32(rx) + np.float32(rw)) * s, (np.float32(ry) + np.float32(rh)) * s
    ix, iy = math.floor(l), math.floor(t)
    iw, ih = max(1, math.ceil(float(r - l))), max(1, math.ceil(float(b - t)))
    ts = abi.Transform(float(s), 0.0, 0.0, float(s), float(np.float32(-ix)), float(np.float32(-iy)))
    return (ix, iy, iw, ih), ts


def rect_pixels(fx: Fixture, ibbox):
    """Pixel span of the (pixel-aligned) rect inside the layer."""
    x, y, w, h = fx.rect
    s = np.float32(SCALE)
    x0 = int(np.float32(x) * s) - ibbox[0]
    y0 = int(np.float32(y) * s) - ibbox[1]
    x1 = int((np.float32(x) + np.float32(w)) * s) - ibbox[0]
    y1 = int((np.float32(y) + np.float32(h)) * s) - ibbox[1]
    return x0, y0, x1, y1


def render_filtered_rect_oracle(fx: Fixture) -> np.ndarray:
    from oracle import bindings as orc
    ibbox, ts = layer_geometry(fx)
    layer = np.zeros((ibbox[3], ibbox[2], 4), dtype=np.uint8)
    x0, y0, x1, y1 = rect_pixels(fx, ibbox)
    layer[y0:y1, x0:x1] = (*fx.color, 255)
    orc.filter_apply(fx.graph, ts, layer)
    canvas = np.zeros((SIZE, SIZE, 4), dtype=np.uint8)
    orc.draw_pixmap(canvas, ibbox[0], ibbox[1], layer)
    return canvas


def render_filtered_rect_gpu(ctx, fx: Fixture) -> np.ndarray:
    import torch
    ibbox, ts = layer_geometry(fx)
    layer = torch.zeros((ibbox[3], ibbox[2], 4), dtype=torch.uint8, device="cuda")
    x0, y0, x1, y1 = rect_pixels(fx, ibbox)
    layer[y0:y1, x0:x1] = torch.tensor([*fx.color, 255], dtype=torch.uint8, device="cuda")
    ctx.filter_apply(fx.graph, ts, layer)
    canvas = torch.zeros((SIZE, SIZE, 4), dtype=torch.uint8, device="cuda")
    ctx.draw_pixmap(canvas, ibbox[0], ibbox[1], layer)
    torch.cuda.synchronize()
    return canvas.cpu().numpy()


def load_png(name: str) -> np.ndarray:
    from PIL import Image
    return np.asarray(Image.open(PNG_DIR / (name.replace("/", "__") + ".png")).convert("RGBA")).copy()


def demultiply_f64(px: np.ndarray) -> np.ndarray:
    """The test harness's demultiply (main.rs:204-211): (c / (a/255) + 0.5) as u8 in f64."""
    a = px[..., 3:4].astype(np.float64) / 255.0
    with np.errstate(divide="ignore", invalid="ignore"):
        c = px[..., :3].astype(np.float64) / a + 0.5
    c = np.where(np.isnan(c), 0.0, np.clip(c, 0.0, 255.0))
    out = px.copy()
    out[..., :3] = np.floor(c).astype(np.uint8)
    return out


def compare_png(actual_premul: np.ndarray, expected: np.ndarray, band: int = 3, guides=()) -> int:
    """Number of differing pixels under the reference's rule, ignoring the
    stroked frame band (device pixels 0..band-1 and SIZE-band..SIZE-1) and the
    strokes of ``guides``."""
    act = demultiply_f64(actual_premul).astype(np.int32)
    exp = expected.astype(np.int32)
    both_clear = (act[..., 3] == 0) & (exp[..., 3] == 0)
    diff = (np.abs(act - exp) > 1).any(axis=-1) & ~both_clear
    mask = np.ones(diff.shape, dtype=bool)
    mask[:band, :] = mask[-band:, :] = mask[:, :band] = mask[:, -band:] = False
    mask &= ~guide_mask(guides, diff.shape[0])
    return int((diff & mask).sum())
