"""The drop-in boundary beyond rvk_render, against the oracle: rvk_render_node
(resvg::render_node), the per-operation replacements of the tiny-skia calls the
render walk makes (fill_path, stroke_path, Mask::fill_path, Mask::from_pixmap,
Mask::invert, apply_mask, fill_rect with Clear), all through the C-ABI."""
import numpy as np
import pytest

from oracle import bindings as orc
from resvg_amd import abi
from resvg_amd.scene import PaintSpec, Scene, ellipse_path, rect_path

from test_gpu_render import _assert_same, _clip_or_mask_scene, _paint, _random_path, _random_scene

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(ctx, t):
    ctx.synchronize()
    return t.cpu().numpy()


def _rand_premul(h, w, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, 1), dtype=np.uint8)
    return np.concatenate([np.minimum(rng.integers(0, 256, (h, w, 3), dtype=np.uint8), a), a], -1)


# ---------------------------------------------------------------------------
# rvk_render_node
# ---------------------------------------------------------------------------
def _node_cases(sc):
    """Nodes of the document tree (reachable from the root group, so not the
    content of clip paths and masks, which render_node does not address), root
    excluded."""
    nodes, stack = [], [sc.root]
    while stack:
        for kind, idx in sc.groups[stack.pop()]["children"]:
            nodes.append((kind, idx))
            if kind == abi.NODE_GROUP:
                stack.append(idx)
    return sorted(nodes)


@pytest.mark.parametrize("seed", range(3))
def test_render_node_matches_oracle(ctx, seed):
    sc = _random_scene(900 + seed, 160, 10) if seed != 2 else _clip_or_mask_scene(910)
    ts = abi.Transform(1.2, 0.0, 0.0, 1.2, 3.0, -2.0)
    s = sc.to_abi()
    checked = 0
    for kind, idx in _node_cases(sc)[:12]:
        bb = sc.abs_layer_bbox(kind, idx)
        if bb is None:
            continue
        w, h = max(1, int(np.ceil(bb[2] * 1.2))), max(1, int(np.ceil(bb[3] * 1.2)))
        base = _rand_premul(h, w, idx)
        px = _dev(base)
        drawn = ctx.render_node(sc, (kind, idx), ts, px, abs_layer_bbox=bb)
        ref = base.copy()
        assert orc.render_node(s.abi, kind, idx, bb, ts, ref) == 1 and drawn
        _assert_same(_host(ctx, px), ref, f"render_node seed {seed} node {(kind, idx)}")
        checked += 1
    assert checked >= 4


def test_render_node_without_a_bounding_box_draws_nothing(ctx):
    import torch
    sc = Scene(64, 64)
    # a horizontal line without a stroke: usvg gives it no bounding box (render_node's None)
    sc.path(sc.root, [abi.VERB_MOVE, abi.VERB_LINE], [(5.0, 20.0), (50.0, 20.0)], fill=PaintSpec.solid((255, 0, 0)))
    v, p = rect_path(4, 4, 30, 30)
    sc.path(sc.root, v, p, fill=PaintSpec.solid((0, 0, 255)))
    assert sc.abs_layer_bbox(abi.NODE_PATH, 0) is None
    px = torch.zeros((40, 40, 4), dtype=torch.uint8, device="cuda")
    assert ctx.render_node(sc, (abi.NODE_PATH, 0), abi.Transform.identity(), px) is False
    assert not _host(ctx, px).any()
    # the rect's box starts at (4, 4): drawn at the pixmap origin
    assert ctx.render_node(sc, (abi.NODE_PATH, 1), abi.Transform.identity(), px) is True
    got = _host(ctx, px)
    assert (got[:30, :30, 2] == 255).all() and not got[30:, :, :].any() and not got[:, 30:, :].any()


# ---------------------------------------------------------------------------
# fill_path / stroke_path / Clear fills
# ---------------------------------------------------------------------------
def _one_path_scene(size, verbs, pts, **kw):
    sc = Scene(size, size)
    sc.path(sc.root, verbs, pts, **kw)
    return sc


@pytest.mark.parametrize("seed", range(8))
def test_fill_path_matches_oracle(ctx, seed):
    rng = np.random.default_rng(1000 + seed)
    size = 140
    base = _rand_premul(size, size, seed)
    px = _dev(base)
    ref = base.copy()
    for i in range(6):
        verbs, pts = _random_path(rng, size, int(rng.integers(1, 10)), margin=0.2 if i == 0 else 0.0)
        paint = _paint(rng, size)
        rule = abi.FILL_EVENODD if rng.random() < 0.5 else abi.FILL_NONZERO
        aa = rng.random() < 0.8
        ts = abi.Transform(*map(float, np.float32([1.0, 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal(),
                                                   1.0, rng.uniform(-5, 5), rng.uniform(-5, 5)])))
        ctx.fill_path(px, verbs, pts, paint, rule, ts, anti_alias=aa)
        sc = _one_path_scene(size, verbs, pts, fill=paint, fill_rule=rule, anti_alias=aa)
        orc.fill_path(sc.to_abi().abi, 0, ts, ref)
    _assert_same(_host(ctx, px), ref, f"fill_path seed {seed}")


@pytest.mark.parametrize("seed", range(4))
def test_clear_fill_matches_oracle(ctx, seed):
    # clip path children: BlendMode::Clear lerps every channel toward zero by coverage
    rng = np.random.default_rng(1100 + seed)
    size = 120
    base = _rand_premul(size, size, 50 + seed)
    px = _dev(base)
    ref = base.copy()
    for _ in range(4):
        verbs, pts = _random_path(rng, size, int(rng.integers(2, 8)))
        rule = abi.FILL_EVENODD if rng.random() < 0.5 else abi.FILL_NONZERO
        ctx.fill_path(px, verbs, pts, PaintSpec.solid((0, 0, 0)), rule, blend_mode=abi.BLEND["clear"])
        sc = _one_path_scene(size, verbs, pts, fill=PaintSpec.solid((0, 0, 0)), fill_rule=rule)
        orc.clear_path(sc.to_abi().abi, 0, abi.Transform.identity(), ref)
    _assert_same(_host(ctx, px), ref, f"clear fill seed {seed}")


@pytest.mark.parametrize("seed", range(8))
def test_stroke_path_matches_oracle(ctx, seed):
    rng = np.random.default_rng(1200 + seed)
    size = 150
    base = _rand_premul(size, size, 80 + seed)
    px = _dev(base)
    ref = base.copy()
    for i in range(5):
        verbs, pts = _random_path(rng, size, int(rng.integers(1, 8)), closed=rng.random() < 0.5)
        paint = _paint(rng, size)
        width = float(np.float32(rng.uniform(0.3, 1.0) if i == 0 else rng.uniform(1.0, 14.0)))  # one hairline
        cap = [abi.CAP_BUTT, abi.CAP_ROUND, abi.CAP_SQUARE][int(rng.integers(3))]
        join = [abi.JOIN_MITER, abi.JOIN_MITER_CLIP, abi.JOIN_ROUND, abi.JOIN_BEVEL][int(rng.integers(4))]
        miter = float(np.float32(rng.uniform(1.0, 8.0)))
        dashes = [float(np.float32(v)) for v in rng.uniform(1, 15, 2 * int(rng.integers(1, 3)))] \
            if rng.random() < 0.4 else None
        off = float(np.float32(rng.uniform(0, 10))) if dashes else 0.0
        aa = rng.random() < 0.85
        ctx.stroke_path(px, verbs, pts, paint, width, cap, join, miter, dashes, off, anti_alias=aa)
        sc = _one_path_scene(size, verbs, pts, stroke=paint, stroke_width=width, stroke_cap=cap, stroke_join=join,
                             miter_limit=miter, dasharray=dashes, dashoffset=off, anti_alias=aa)
        orc.stroke_path(sc.to_abi().abi, 0, abi.Transform.identity(), ref)
    _assert_same(_host(ctx, px), ref, f"stroke_path seed {seed}")


def test_fill_path_rejects_other_blend_modes_and_skips_flat_paths(ctx):
    import torch
    import resvg_amd
    px = torch.zeros((32, 32, 4), dtype=torch.uint8, device="cuda")
    v, p = rect_path(2, 2, 20, 20)
    with pytest.raises(resvg_amd.RvkError) as e:
        ctx.fill_path(px, v, p, PaintSpec.solid((1, 2, 3)), blend_mode=abi.BLEND["multiply"])
    assert e.value.status == abi.ERR_UNSUPPORTED
    # a zero-height path cannot be filled (path.rs:36-38); a verb list asking for more
    # points than given is rejected before anything is drawn
    ctx.fill_path(px, [abi.VERB_MOVE, abi.VERB_LINE, abi.VERB_CLOSE], [(1, 5), (30, 5)], PaintSpec.solid((9, 9, 9)))
    with pytest.raises(resvg_amd.RvkError) as e:
        ctx.fill_path(px, [abi.VERB_MOVE, abi.VERB_CUBIC], [(1, 5), (30, 5)], PaintSpec.solid((9, 9, 9)))
    assert e.value.status == abi.ERR_INVALID_ARGUMENT
    assert not _host(ctx, px).any()


# ---------------------------------------------------------------------------
# masks
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (37, 29), (64, 64)])
def test_mask_from_pixmap_invert_apply(ctx, shape):
    import torch
    h, w = shape
    src = _rand_premul(h, w, h * 131 + w)
    dst = _rand_premul(h, w, h * 7 + w + 1)
    for lum in (False, True):
        m = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        ctx.mask_from_pixmap(_dev(src), m, abi.MASK_TYPE_LUMINANCE if lum else abi.MASK_TYPE_ALPHA)
        exp = orc.mask_from_pixmap(src, lum)
        got = _host(ctx, m)
        assert np.array_equal(got, exp), ("from_pixmap", lum)
        ctx.mask_invert(m)
        assert np.array_equal(_host(ctx, m), 255 - exp)
        px = _dev(dst)
        ctx.apply_mask(px, m)
        ref = dst.copy()
        orc.apply_mask(ref, np.ascontiguousarray(255 - exp))
        _assert_same(_host(ctx, px), ref, f"apply_mask {shape} luminance={lum}")


@pytest.mark.parametrize("seed", range(4))
def test_mask_fill_path_is_the_path_coverage(ctx, seed):
    import torch
    rng = np.random.default_rng(1300 + seed)
    size = 96
    verbs, pts = (_random_path(rng, size, 6) if seed < 3 else ellipse_path(40.3, 50.1, 30.7, 20.2))
    rule = abi.FILL_EVENODD if seed % 2 else abi.FILL_NONZERO
    ts = abi.Transform(1.0, 0.0, 0.2, 0.9, 2.5, -1.5)
    m = torch.full((size, size), 77, dtype=torch.uint8, device="cuda")
    ctx.mask_fill_path(m, verbs, pts, rule, True, ts)
    sc = _one_path_scene(size, verbs, pts, fill=PaintSpec.solid((255, 255, 255)), fill_rule=rule)
    exp = orc.path_coverage(sc.to_abi().abi, 0, ts, False, size, size)
    assert np.array_equal(_host(ctx, m), exp)


def test_clear_rect(ctx):
    base = _rand_premul(20, 30, 3)
    for x, y, rw, rh in ((5, 4, 10, 7), (-3, -2, 8, 50), (25, 15, 100, 100), (40, 0, 5, 5)):
        px = _dev(base)
        ctx.clear_rect(px, x, y, rw, rh)
        exp = base.copy()
        exp[max(0, y):max(0, y + rh), max(0, x):max(0, x + rw)] = 0
        assert np.array_equal(_host(ctx, px), exp), (x, y, rw, rh)


def test_context_follows_torch_current_stream(ctx):
    # work issued under `with torch.cuda.stream(s)` is ordered on s
    import torch
    s = torch.cuda.Stream()
    sc = _random_scene(950, 96, 6)
    ts = abi.Transform.identity()
    with torch.cuda.stream(s):
        px = torch.zeros((96, 96, 4), dtype=torch.uint8, device="cuda")
        ctx.render(sc, ts, px)
        out = px.clone()  # on s, after the render
    s.synchronize()
    ref = np.zeros((96, 96, 4), np.uint8)
    orc.render(sc.to_abi().abi, ts, ref)
    _assert_same(out.cpu().numpy(), ref, "render on a side stream")
