"""GPU parity of every filter primitive and of the filter-graph executor.

Each case runs the HIP kernel through the C-ABI (resvg_amd.Context) and the oracle
on the same seeded input; integer paths must be bit-exact, and so are the float
paths here because every kernel keeps the reference's operation order (the only
per-pixel libm call on the device is powf in lighting, tested within 1 LSB).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import bindings as orc
from resvg_amd import abi
from resvg_amd.filters import SOURCE_ALPHA, SOURCE_GRAPHIC, FilterGraph, distant_light, point_light, spot_light

from fixture_scenes import compare_png, fixtures, load_png, render_filtered_rect_gpu, render_filtered_rect_oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rand_premul(h, w, seed):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    a = px[..., 3:4].astype(np.uint16)
    px[..., :3] = (px[..., :3].astype(np.uint16) * a // 255).astype(np.uint8)
    return px


def smooth_premul(h, w, seed):
    """Blobby images: realistic inputs for lighting/convolve with flat areas."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    a = np.zeros((h, w))
    for _ in range(6):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(3, max(4, min(w, h) / 2))
        a += np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
    a = np.clip(a * 255, 0, 255).astype(np.uint8)
    col = rng.integers(0, 256, size=3)
    px = np.zeros((h, w, 4), np.uint8)
    px[..., 3] = a
    px[..., :3] = (col[None, None, :] * a[..., None].astype(np.uint16) // 255).astype(np.uint8)
    return px


def dev(a):
    return torch.from_numpy(a.copy()).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def assert_exact(got, exp, what=""):
    if not np.array_equal(got, exp):
        d = np.abs(got.astype(int) - exp.astype(int))
        idx = np.argwhere(d > 0)[:5]
        raise AssertionError(f"{what}: {int((d > 0).any(-1).sum())} px differ, max {d.max()}, first {idx.tolist()}")


SHAPES = [(1, 1), (3, 50), (37, 19), (64, 64), (130, 257)]


@pytest.mark.parametrize("shape", SHAPES)
def test_colorspace_and_alpha(ctx, shape):
    px = rand_premul(*shape, seed=1)
    for name in ("into_linear_rgb", "into_srgb", "demultiply_alpha", "multiply_alpha"):
        t = dev(px)
        getattr(ctx, name)(t)
        e = px.copy()
        getattr(orc, name)(e)
        assert_exact(host(t), e, name)


@pytest.mark.parametrize("shape", SHAPES + [(600, 700)])
@pytest.mark.parametrize("sig", [(2.0, 2.0), (4.0, 4.0), (6.0, 0.5), (0.0, 5.0), (24.0, 24.0), (24.0, 3.0),
                                 (100.0, 60.0)])
def test_box_blur(ctx, shape, sig):
    px = rand_premul(*shape, seed=2)
    t = dev(px)
    ctx.box_blur(t, *sig)
    e = px.copy()
    orc.box_blur(e, *sig)
    assert_exact(host(t), e, f"box {sig}")


@pytest.mark.parametrize("shape", SHAPES + [(300, 410)])
@pytest.mark.parametrize("sig", [(0.5, 0.5), (1.5, 1.0), (1.9, 0.0), (0.0, 0.7), (40.0, 3.0), (0.0, 9.0)])
def test_iir_blur(ctx, shape, sig):
    """Sigmas below 2 (the filter graph's IIR range) and, through the C-ABI entry,
    larger ones (dnu up to 0.95: the longest warm-up and the most fixpoint rounds)."""
    px = rand_premul(*shape, seed=3)
    t = dev(px)
    ctx.iir_blur(t, *sig)
    e = px.copy()
    orc.iir_blur(e, *sig)
    assert_exact(host(t), e, f"iir {sig}")


def sparse_premul(h, w, seed, density=0.003):
    """Isolated pixels on a transparent background: every chain has long runs of
    zeros next to content, where values decay towards f64 underflow."""
    rng = np.random.default_rng(seed)
    px = np.zeros((h, w, 4), np.uint8)
    a = np.where(rng.random((h, w)) < density, rng.integers(1, 256, size=(h, w)), 0)
    px[..., 3] = a
    px[..., :3] = (rng.integers(0, 256, size=(h, w, 3)) * a[..., None] // 255).astype(np.uint8)
    return px


@pytest.mark.parametrize("shape,sig", [((1, 4096), s) for s in [(1.99, 1.99), (0.5, 1.5), (0.05, 0.0), (0.0, 1.2),
                                                                  (0.0, 0.0)]] +
                         [(sh, s) for sh in [(4096, 2), (3, 4097), (4097, 5), (700, 1500), (1024, 1024)]
                          for s in [(1.99, 1.99), (0.5, 1.5), (0.05, 0.0), (0.0, 1.2), (0.0, 0.0)]] +
                         [((1800, 2600), (1.5, 0.5)), ((2600, 1800), (0.5, 1.5))])
@pytest.mark.parametrize("kind", ["sparse", "blobs"])
def test_iir_blur_long_chains_and_transparent_runs(ctx, shape, sig, kind):
    """Chains of every on-chip chunk length (8 to 64 values a lane, up to 4096 values)
    and past the on-chip limit (the streaming form), on images whose chains mix
    content with long transparent runs."""
    px = sparse_premul(*shape, seed=5) if kind == "sparse" else smooth_premul(*shape, seed=5)
    t = dev(px)
    ctx.iir_blur(t, *sig)
    e = px.copy()
    orc.iir_blur(e, *sig)
    assert_exact(host(t), e, f"iir {sig} {kind}")


@pytest.mark.parametrize("shape", SHAPES + [(500, 333)])
@pytest.mark.parametrize("r", [(1.0, 1.0), (2.5, 0.5), (4.0, 4.0), (3.5, 7.5), (8.0, 8.0), (8.0, 3.0),
                               (12.0, 16.0), (20.0, 18.0), (16.5, 2.0)])
@pytest.mark.parametrize("op", [abi.MORPH_ERODE, abi.MORPH_DILATE])
def test_morphology(ctx, shape, r, op):
    px = rand_premul(*shape, seed=4)
    t = dev(px)
    ctx.morphology(t, op, *r)
    e = px.copy()
    orc.morphology(e, op, *r)
    assert_exact(host(t), e, f"morph {op} {r}")


def _conv_prim(cols, rows, kernel, edge, preserve, divisor=None, bias=0.0):
    g = FilterGraph((0, 0, 1, 1))
    g.convolve_matrix(cols, rows, kernel, divisor=divisor, bias=bias, edge_mode=edge, preserve_alpha=preserve)
    return g.primitives[0], g.values


@pytest.mark.parametrize("shape", [(1, 1), (5, 7), (64, 33)])
@pytest.mark.parametrize("order", [(3, 3), (5, 5), (3, 1), (9, 9)])
@pytest.mark.parametrize("edge", [abi.EDGE_NONE, abi.EDGE_DUPLICATE, abi.EDGE_WRAP])
@pytest.mark.parametrize("preserve", [False, True])
def test_convolve(ctx, shape, order, edge, preserve):
    rng = np.random.default_rng(hash((shape, order, edge, preserve)) & 0xffff)
    k = rng.integers(-4, 5, size=order[0] * order[1]).astype(float)
    prim, vals = _conv_prim(order[0], order[1], k, edge, preserve, bias=0.1 if preserve else 0.0)
    px = smooth_premul(*shape, seed=5) if shape[0] > 8 else rand_premul(*shape, seed=5)
    t = dev(px)
    ctx.convolve_matrix(t, prim, vals)
    e = px.copy()
    orc.convolve_matrix(e, prim, vals)
    assert_exact(host(t), e, "convolve")


@pytest.mark.parametrize("kind,values", [
    (abi.COLOR_MATRIX_MATRIX, None), (abi.COLOR_MATRIX_SATURATE, [0.3]), (abi.COLOR_MATRIX_SATURATE, [2.5]),
    (abi.COLOR_MATRIX_HUE_ROTATE, [37.0]), (abi.COLOR_MATRIX_LUMINANCE_TO_ALPHA, [])])
def test_color_matrix(ctx, kind, values):
    rng = np.random.default_rng(6)
    if values is None:
        values = rng.uniform(-1, 1, size=20).astype(np.float32).tolist()
    g = FilterGraph((0, 0, 1, 1))
    g.color_matrix(kind, values)
    px = rand_premul(45, 61, 6)
    t = dev(px)
    ctx.color_matrix(t, g.primitives[0])
    e = px.copy()
    orc.color_matrix(e, g.primitives[0])
    assert_exact(host(t), e, "color matrix")


@pytest.mark.parametrize("funcs", [
    [("table", [0.0, 1.0]), ("discrete", [0.2, 0.9, 0.4]), ("linear", 0.5, 0.25), ("gamma", 2.0, 0.7, 0.1)],
    [None, ("table", [1, 1, 0, 0]), ("gamma", 1.3, 2.2, -0.05), ("discrete", [])],
    [("table", [0.9, -0.3, 1.2, 0.5, 0.0]), None, None, ("linear", -1.0, 1.0)],
])
def test_component_transfer(ctx, funcs):
    g = FilterGraph((0, 0, 1, 1))
    g.component_transfer(funcs)
    px = rand_premul(33, 47, 7)
    t = dev(px)
    ctx.component_transfer(t, g.primitives[0], g.values)
    e = px.copy()
    orc.component_transfer(e, g.primitives[0], g.values)
    assert_exact(host(t), e, "component transfer")


@pytest.mark.parametrize("k", [(0.1, 0.2, 0.3, 0.4), (0.0, 1.0, 0.0, 0.0), (1.5, -0.5, 0.7, -0.1), (0, 0, 0, 0)])
def test_arithmetic(ctx, k):
    s1, s2, d0 = rand_premul(29, 31, 8), rand_premul(29, 31, 9), rand_premul(29, 31, 10)
    d = dev(d0)
    ctx.composite_arithmetic(dev(s1), dev(s2), d, *k)
    e = d0.copy()
    orc.composite_arithmetic(s1, s2, e, *k)
    assert_exact(host(d), e, "arithmetic")


@pytest.mark.parametrize("args", [
    (0.0, 0.0, 1.5, 1.5, 0.1, 0.1, 1, 0, False, True), (60.0, 60.0, 1.5, 1.5, 0.05, 0.01, 4, 20, False, False),
    (10.5, -3.0, 2.0, 0.5, 0.03, 0.03, 2, -7, True, False), (0.0, 0.0, 1.0, 1.0, 0.08, 0.02, 4, 3, True, True)])
def test_turbulence(ctx, args):
    out = torch.zeros((70, 90, 4), dtype=torch.uint8, device="cuda")
    ctx.turbulence(out, *args)
    e = np.zeros((70, 90, 4), np.uint8)
    orc.turbulence(e, *args)
    assert_exact(host(out), e, "turbulence")


@pytest.mark.parametrize("specular", [False, True])
@pytest.mark.parametrize("light", [distant_light(45.0, 30.0), point_light(20.0, 15.0, 30.0),
                                   spot_light(10.0, 10.0, 40.0, 40.0, 30.0, 0.0, 3.0, 30.0),
                                   spot_light(50.0, 5.0, 20.0, 0.0, 40.0, 0.0, 1.0, None)])
def test_lighting(ctx, specular, light):
    src = smooth_premul(48, 64, 11)
    d = torch.zeros((48, 64, 4), dtype=torch.uint8, device="cuda")
    ctx.lighting(dev(src), d, specular, 3.0, 1.2, 8.0 if specular else 1.0, (250, 200, 120), light)
    e = np.zeros_like(src)
    orc.lighting(src, e, specular, 3.0, 1.2, 8.0 if specular else 1.0, (250, 200, 120), light)
    got = host(d)
    # per-pixel powf is the device libm's: allow 1 LSB, and report the histogram
    diff = np.abs(got.astype(int) - e.astype(int))
    assert diff.max() <= 1, f"max diff {diff.max()}"
    assert (diff > 0).mean() < 0.01


def test_displacement_map(ctx):
    src, m = rand_premul(40, 50, 12), rand_premul(40, 50, 13)
    d = torch.zeros((40, 50, 4), dtype=torch.uint8, device="cuda")
    ctx.displacement_map(dev(src), dev(m), d, 1.5, 1.5, 12.0, abi.CHANNEL_R, abi.CHANNEL_A)
    e = np.zeros_like(src)
    orc.displacement_map(src, m, e, 1.5, 1.5, 12.0, abi.CHANNEL_R, abi.CHANNEL_A)
    assert_exact(host(d), e, "displacement")


@pytest.mark.parametrize("mode", list(range(len(abi.BLEND_MODES))), ids=abi.BLEND_MODES)
@pytest.mark.parametrize("opacity", [1.0, 0.37])
def test_draw_pixmap_modes(ctx, mode, opacity):
    dst0, src = rand_premul(40, 52, 14), rand_premul(25, 30, 15)
    d = dev(dst0)
    ctx.draw_pixmap(d, 7, -4, dev(src), opacity, mode)
    e = dst0.copy()
    orc.draw_pixmap(e, 7, -4, src, opacity, mode)
    assert_exact(host(d), e, abi.BLEND_MODES[mode])


FIX = fixtures()


@pytest.mark.parametrize("fx", FIX, ids=[f.name for f in FIX])
def test_fixture_graph_gpu_vs_oracle_and_png(ctx, fx):
    got = render_filtered_rect_gpu(ctx, fx)
    assert_exact(got, render_filtered_rect_oracle(fx), fx.name)
    assert compare_png(got, load_png(fx.name), guides=fx.guides) == 0


def _random_graph(rng, region):
    g = FilterGraph(region)
    last = SOURCE_GRAPHIC
    kinds = ["blur_box", "blur_iir", "offset", "morph", "conv", "cmatrix", "ctransfer", "turb", "light", "blend",
             "composite", "flood", "merge", "shadow", "alpha"]
    for _ in range(int(rng.integers(1, 6))):
        k = kinds[int(rng.integers(len(kinds)))]
        cs = int(rng.integers(2))
        kw = {"cs": cs}
        if rng.random() < 0.2:
            x, y, w, h = region
            kw["rect"] = (x + w * 0.1, y + h * 0.2, w * 0.7, h * 0.5)
        if k == "blur_box":
            last = g.gaussian_blur(float(rng.uniform(2, 8)), float(rng.uniform(0, 6)), in1=last, **kw)
        elif k == "blur_iir":
            last = g.gaussian_blur(float(rng.uniform(0.2, 1.9)), float(rng.uniform(0.2, 1.9)), in1=last, **kw)
        elif k == "offset":
            last = g.offset(float(rng.uniform(-9, 9)), float(rng.uniform(-9, 9)), in1=last, **kw)
        elif k == "morph":
            last = g.morphology(int(rng.integers(2)), float(rng.uniform(0.5, 6)), float(rng.uniform(0.5, 6)),
                                in1=last, **kw)
        elif k == "conv":
            last = g.convolve_matrix(3, 3, rng.integers(-4, 5, 9).astype(float), in1=last, **kw)
        elif k == "cmatrix":
            last = g.color_matrix(abi.COLOR_MATRIX_MATRIX, rng.uniform(-1, 1, 20).astype(np.float32).tolist(),
                                  in1=last, **kw)
        elif k == "ctransfer":
            last = g.component_transfer([("gamma", 1.5, 0.8, 0.05), None, ("table", [0.1, 0.9, 0.3]), None],
                                        in1=last, **kw)
        elif k == "turb":
            last = g.turbulence(0.05, 0.07, 2, int(rng.integers(100)), False, bool(rng.integers(2)), **kw)
        elif k == "light":
            last = g.lighting(bool(rng.integers(2)), distant_light(30.0, 40.0), 2.0, 1.0, 1.0, in1=last, **kw)
        elif k == "blend":
            last = g.blend(int(rng.integers(len(abi.BLEND_MODES))), in1=last, in2=SOURCE_GRAPHIC, **kw)
        elif k == "composite":
            op = int(rng.integers(6))
            last = g.composite(op, k=tuple(rng.uniform(-1, 1, 4).tolist()), in1=last, in2=SOURCE_ALPHA, **kw)
        elif k == "flood":
            last = g.flood(tuple(rng.integers(0, 256, 3).tolist()), float(rng.uniform(0, 1)), **kw)
        elif k == "merge":
            last = g.merge([last, SOURCE_GRAPHIC], **kw)
        elif k == "shadow":
            last = g.drop_shadow(3.0, 4.0, 2.5, 2.5, (10, 20, 200), 0.7, in1=last, **kw)
        else:
            last = g.gaussian_blur(0.0, 0.0, in1=SOURCE_ALPHA, **kw)
    return g


@pytest.mark.parametrize("seed", range(12))
def test_random_filter_graphs(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    h, w = int(rng.integers(20, 90)), int(rng.integers(20, 90))
    layer = smooth_premul(h, w, seed)
    region = (float(rng.uniform(-5, 5)), float(rng.uniform(-5, 5)), float(w * rng.uniform(0.6, 1.2)),
              float(h * rng.uniform(0.6, 1.2)))
    g = _random_graph(rng, region)
    ts = abi.Transform(float(rng.uniform(0.7, 1.6)), 0.0, 0.0, float(rng.uniform(0.7, 1.6)),
                       float(rng.uniform(-3, 3)), float(rng.uniform(-3, 3)))
    t = dev(layer)
    err_g = ctx.filter_apply(g, ts, t)
    e = layer.copy()
    err_o = orc.filter_apply(g, ts, e)
    assert err_g == err_o
    got = host(t)
    diff = np.abs(got.astype(int) - e.astype(int))
    has_light = any(p.kind in (abi.FE_DIFFUSE_LIGHTING, abi.FE_SPECULAR_LIGHTING) for p in g.primitives)
    if has_light:
        assert diff.max() <= 1
    else:
        assert_exact(got, e, f"graph seed {seed}")


@pytest.mark.parametrize("case", ["box_blur", "morphology_small", "morphology_large"])
def test_filters_on_a_pixmap_not_16_byte_aligned(ctx, case):
    # the fused box blur and morphology kernels take 16-byte loads and stores only
    # where the pixmap and its rows are 16-byte aligned; a pixmap 4 bytes past such a
    # boundary (rows 4-pixel aligned) takes the single-pixel paths with the same result
    h, w = 70, 132
    px = rand_premul(h, w, seed=9)
    aligned = dev(px)
    buf = torch.zeros(h * w * 4 + 16, dtype=torch.uint8, device="cuda")
    shifted = buf[4:4 + h * w * 4].view(h, w, 4)
    assert shifted.data_ptr() % 16 == 4 and aligned.data_ptr() % 16 == 0
    shifted.copy_(aligned)
    exp = px.copy()
    if case == "box_blur":
        for t in (aligned, shifted):
            ctx.box_blur(t, 3.0, 3.0)
        orc.box_blur(exp, 3.0, 3.0)
    else:
        r = 3.0 if case == "morphology_small" else 8.0
        for t in (aligned, shifted):
            ctx.morphology(t, abi.MORPH_ERODE, r, r)
        orc.morphology(exp, abi.MORPH_ERODE, r, r)
    assert_exact(host(shifted), host(aligned), f"{case}: shifted vs aligned")
    assert_exact(host(aligned), exp, f"{case}: vs the oracle")


@pytest.mark.parametrize("seed", range(10))
def test_tile_graphs(ctx, seed):
    """feTile of a subregioned input (inside, across or outside the layer; from a
    blur, a flood moved by feOffset, or the whole SourceGraphic), repeated over the
    filter region or a subregion of its own, then read by one more primitive."""
    rng = np.random.default_rng(700 + seed)
    h, w = int(rng.integers(24, 96)), int(rng.integers(24, 96))
    layer = smooth_premul(h, w, seed)
    region = (float(rng.uniform(-6, 4)), float(rng.uniform(-6, 4)), float(w * rng.uniform(0.7, 1.3)),
              float(h * rng.uniform(0.7, 1.3)))
    x, y, rw, rh = region
    sub = (x + float(rng.uniform(-0.3, 0.8)) * rw, y + float(rng.uniform(-0.3, 0.8)) * rh,
           float(rng.uniform(2, 0.5 * rw)), float(rng.uniform(2, 0.5 * rh)))
    cs = int(rng.integers(2))
    g = FilterGraph(region)
    kind = seed % 3
    if kind == 0:
        src = g.gaussian_blur(1.5, in1=SOURCE_GRAPHIC, rect=sub, cs=cs)
    elif kind == 1:
        f = g.flood((200, 40, 90), 0.8, rect=sub, cs=cs)
        src = g.offset(float(rng.uniform(-6, 6)), float(rng.uniform(-6, 6)), in1=f, cs=cs)
    else:
        src = SOURCE_GRAPHIC
    tile_rect = None if rng.random() < 0.5 else (x + 0.1 * rw, y + 0.15 * rh, 0.75 * rw, 0.6 * rh)
    t = g.tile(in1=src, rect=tile_rect, cs=cs)
    g.composite(abi.COMPOSITE_OVER, in1=t, in2=SOURCE_GRAPHIC, cs=1 - cs)
    ts = abi.Transform(float(rng.uniform(0.7, 1.6)), 0.0, 0.0, float(rng.uniform(0.7, 1.6)),
                       float(rng.uniform(-3, 3)), float(rng.uniform(-3, 3)))
    tt = dev(layer)
    err_g = ctx.filter_apply(g, ts, tt)
    e = layer.copy()
    err_o = orc.filter_apply(g, ts, e)
    assert err_g == err_o
    assert_exact(host(tt), e, f"tile seed {seed} kind {kind}")
