"""Pins the oracle's restatement of tiny-skia's EdgeClipper (scan/edge_clipper.rs,
scan/line_clipper.rs; Skia SkEdgeClipper / SkLineClipper) by what the geometry fixes:
textbook line/rect intersections, clipped pieces that stay inside the clip and on the
original curves, monotonic curve pieces, and winding numbers of interior points that
clipping must not change.  The rendered effect is pinned end to end by the
filters/feImage/link-to-an-element-with-transform fixture (a skewed rect crossing the
feImage pixmap), whose reference PNG the unclipped edges miss by 9 pixels."""
import numpy as np
import pytest

from oracle import bindings as orc

W, H = 40, 30


def _row(*pts):
    r = np.zeros(9, np.float32)
    r[0] = len(pts)
    for i, (x, y) in enumerate(pts):
        r[1 + 2 * i], r[2 + 2 * i] = x, y
    return r


def _pts(row):
    n = int(row[0])
    return [(float(row[1 + 2 * i]), float(row[2 + 2 * i])) for i in range(n)]


def test_line_crossing_two_sides_is_cut_at_the_intersections():
    # y = x from (-10, -10) to (30, 30) against [0, 40] x [0, 30]: enters at (0, 0), leaves at (30, 30)
    out = orc.edge_clip(_row((-10.0, -10.0), (50.0, 50.0)), W, H)
    assert [_pts(r) for r in out] == [[(0.0, 0.0), (30.0, 30.0)]]


def test_parts_left_and_right_become_vertical_lines_on_the_clip_edge():
    # from (-20, 5) to (60, 25): slope 1/4; crosses x = 0 at y = 10 and x = 40 at y = 20
    out = [_pts(r) for r in orc.edge_clip(_row((-20.0, 5.0), (60.0, 25.0)), W, H)]
    assert out == [[(0.0, 5.0), (0.0, 10.0)], [(0.0, 10.0), (40.0, 20.0)], [(40.0, 20.0), (40.0, 25.0)]]
    # the same line walked backwards keeps its direction piece by piece
    back = [_pts(r) for r in orc.edge_clip(_row((60.0, 25.0), (-20.0, 5.0)), W, H)]
    assert back == [[(40.0, 25.0), (40.0, 20.0)], [(40.0, 20.0), (0.0, 10.0)], [(0.0, 10.0), (0.0, 5.0)]]


def test_lines_wholly_outside():
    assert len(orc.edge_clip(_row((5.0, -8.0), (30.0, -1.0)), W, H)) == 0   # above
    assert len(orc.edge_clip(_row((5.0, 31.0), (30.0, 60.0)), W, H)) == 0   # below
    assert [_pts(r) for r in orc.edge_clip(_row((-5.0, 2.0), (-3.0, 15.0)), W, H)] == [[(0.0, 2.0), (0.0, 15.0)]]
    # to the right: kept as a vertical line (tiny-skia never culls to the right)
    assert [_pts(r) for r in orc.edge_clip(_row((45.0, 20.0), (41.0, 3.0)), W, H)] == [[(40.0, 20.0), (40.0, 3.0)]]


def _sample(row, n=400):
    p = np.array(_pts(row), np.float64)
    t = np.linspace(0.0, 1.0, n)[:, None]
    if len(p) == 2:
        return p[0] * (1 - t) + p[1] * t
    if len(p) == 3:
        return p[0] * (1 - t) ** 2 + 2 * p[1] * t * (1 - t) + p[2] * t * t
    return (p[0] * (1 - t) ** 3 + 3 * p[1] * t * (1 - t) ** 2 + 3 * p[2] * t * t * (1 - t) + p[3] * t ** 3)


def _polyline(rows):
    segs = []
    for r in rows:
        s = _sample(r, 160)
        segs.append(np.stack([s[:-1], s[1:]], axis=1))
    return np.concatenate(segs) if segs else np.zeros((0, 2, 2))


def _winding(lines, px, py):
    """Signed crossings of the horizontal ray to the left of (px, py), per point."""
    a, b = lines[:, 0], lines[:, 1]
    y0, y1 = a[:, 1][None, :], b[:, 1][None, :]
    up = (y0 <= py[:, None]) & (y1 > py[:, None])
    down = (y1 <= py[:, None]) & (y0 > py[:, None])
    t = (py[:, None] - y0) / np.where(y1 == y0, 1.0, y1 - y0)
    x = a[:, 0][None, :] + t * (b[:, 0] - a[:, 0])[None, :]
    left = x < px[:, None]
    return (up & left).sum(1) - (down & left).sum(1)


def _dist_to(lines, px, py):
    a, b = lines[:, 0][None], lines[:, 1][None]
    p = np.stack([px, py], -1)[:, None]
    d = b - a
    t = np.clip(((p - a) * d).sum(-1) / np.maximum((d * d).sum(-1), 1e-12), 0, 1)
    q = a + t[..., None] * d
    return np.sqrt(((p - q) ** 2).sum(-1)).min(1)


def _random_contour(rng, n):
    rows = []
    start = cur = rng.uniform([-25, -20], [W + 25, H + 20])
    for i in range(n):
        k = int(rng.integers(2, 5))
        nxt = start if i == n - 1 else rng.uniform([-25, -20], [W + 25, H + 20])
        ctrl = [tuple(rng.uniform([-30, -25], [W + 30, H + 25])) for _ in range(k - 2)]
        rows.append(_row(tuple(cur), *ctrl, tuple(nxt)))
        cur = nxt
    return np.array(rows)


@pytest.mark.parametrize("seed", range(12))
def test_clipping_keeps_interior_winding_numbers(seed):
    rng = np.random.default_rng(seed)
    rows = _random_contour(rng, int(rng.integers(3, 9)))
    out = orc.edge_clip(rows, W, H)
    pts = np.concatenate([np.array(_pts(r)) for r in out]) if len(out) else np.zeros((0, 2))
    assert (pts[:, 0] >= 0).all() and (pts[:, 0] <= W).all()
    assert (pts[:, 1] >= 0).all() and (pts[:, 1] <= H).all()
    for r in out:  # every curve piece is monotonic in x and in y
        if r[0] > 2:
            s = _sample(r)
            for axis in (0, 1):
                d = np.diff(s[:, axis])
                assert (d >= -1e-3).all() or (d <= 1e-3).all()
    gx, gy = np.meshgrid(np.linspace(0.5, W - 0.5, 48), np.linspace(0.5, H - 0.5, 36))
    px, py = gx.ravel(), gy.ravel()
    orig, clip = _polyline(rows), _polyline(out)
    far = (_dist_to(orig, px, py) > 0.05) & (_dist_to(clip, px, py) > 0.05)
    assert far.sum() > 800
    assert np.array_equal(_winding(orig, px[far], py[far]), _winding(clip, px[far], py[far]))


def test_cubic_chop_points_lie_on_the_curve():
    # a cubic crossing the top edge and the right edge: the pieces' end points sit on
    # the original curve (checked against numpy's roots of the double polynomial)
    src = _row((10.0, -15.0), (25.0, -5.0), (45.0, 12.0), (52.0, 28.0))
    p = np.array(_pts(src), np.float64)
    out = orc.edge_clip(src, W, H)
    curve = [r for r in out if r[0] == 4]
    assert len(curve) == 1
    c = np.array(_pts(curve[0]))
    assert c[0][1] == 0.0 and c[-1][0] == W

    def at(axis, v):
        a = -p[0, axis] + 3 * p[1, axis] - 3 * p[2, axis] + p[3, axis]
        b = 3 * p[0, axis] - 6 * p[1, axis] + 3 * p[2, axis]
        cc = -3 * p[0, axis] + 3 * p[1, axis]
        r = [t.real for t in np.roots([a, b, cc, p[0, axis] - v]) if abs(t.imag) < 1e-9 and 0 <= t.real <= 1]
        t = r[0]
        return (1 - t) ** 3 * p[0] + 3 * t * (1 - t) ** 2 * p[1] + 3 * t * t * (1 - t) * p[2] + t ** 3 * p[3]

    assert abs(at(1, 0.0)[0] - c[0][0]) < 1e-3
    assert abs(at(0, float(W))[1] - c[-1][1]) < 1e-3
    # the part beyond the right edge is a vertical line from there down to the end's y
    assert [_pts(r) for r in out if r[0] == 2] == [[(float(W), float(c[-1][1])), (float(W), 28.0)]]
