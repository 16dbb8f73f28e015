"""FP64 CPU oracle of the differentiable f32 form (test infrastructure only: used by
tests/, never by the product).

The model restated in float64 with numpy, following the same definitions as the
device kernels (resvg_amd/csrc/diff.hip): SourceOver of per-path sources
(rgb·tint·A, A), A = coverage·opacity·stop alpha, the stop colour interpolated along a
pad-spread linear gradient (resvg path.rs:118-177, render.rs:133); a separable
Gaussian blur of radius min(ceil(3σ), 48) with zero padding; a 4x5 colour matrix on
the demultiplied colour, clamped and premultiplied (color_matrix.rs:11-110); and the
composite over the paths' canvas (filter/mod.rs:731-793).  Its gradient is taken by
central differences in float64, which is what the device's analytic adjoints are
checked against.
"""
from __future__ import annotations

import math

import numpy as np

PATH_PARAMS = 24
MAX_RADIUS = 48


def _paths(cov, ends, p, w, h):
    n = cov.shape[0]
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    px, py = xs + 0.5, ys + 0.5
    s = np.zeros((h, w, 4))
    for i in range(n):
        c = cov[i].astype(np.float64) / 255.0
        q = p[i * PATH_PARAMS:(i + 1) * PATH_PARAMS]
        tint, op, col, off = q[0:3], q[3], q[4:20].reshape(4, 4), q[20:24]
        x0, y0, x1, y1 = (float(v) for v in ends[i])
        dx, dy = x1 - x0, y1 - y0
        dd = dx * dx + dy * dy
        t = ((px - x0) * dx + (py - y0) * dy) / dd if dd > 0 else np.zeros_like(px)
        t = np.clip(t, 0.0, 1.0)
        k = np.where(t < off[1], 0, np.where(t < off[2], 1, 2))
        lo, hi = off[k], off[k + 1]
        span = hi - lo
        u = np.where(span > 0, (t - lo) / np.where(span > 0, span, 1.0), 0.0)
        u = np.where(t <= off[0], 0.0, np.where(t >= off[3], 1.0, u))
        k = np.where(t <= off[0], 0, np.where(t >= off[3], 2, k))
        ca, cb = col[k], col[k + 1]
        rgba = ca + u[..., None] * (cb - ca)
        A = c * op * rgba[..., 3]
        src = np.concatenate([rgba[..., :3] * tint * A[..., None], A[..., None]], -1)
        s = src + (1.0 - A[..., None]) * s
    return s


def blur_weights(sigma: float):
    sigma = min(max(sigma, 0.05), 16.0)
    r = min(int(math.ceil(3.0 * sigma)), MAX_RADIUS)
    k = np.arange(-r, r + 1, dtype=np.float64)
    g = np.exp(-k * k / (2.0 * sigma * sigma))
    return r, g / g.sum()


def _blur(img, sigma):
    r, wts = blur_weights(sigma)
    h, w, _ = img.shape
    out = np.zeros_like(img)
    pad = np.zeros((h, w + 2 * r, 4))
    pad[:, r:r + w] = img
    for j, wk in enumerate(wts):
        out += wk * pad[:, j:j + w]
    pad = np.zeros((h + 2 * r, w, 4))
    pad[r:r + h] = out
    out2 = np.zeros_like(img)
    for j, wk in enumerate(wts):
        out2 += wk * pad[j:j + h]
    return out2


ALPHA_FLOOR = 1.0 / 255.0


def _color_matrix(b, m):
    # demultiplied by max(a, 1/255): exact for the alphas of a u8 pixmap (non-zero ones
    # are >= 1/255), a continuous ramp to zero below
    a = b[..., 3]
    den = np.maximum(a, ALPHA_FLOOR)
    v = np.stack([b[..., 0] / den, b[..., 1] / den, b[..., 2] / den, a], -1)
    o = v @ m[:, :4].T + m[:, 4]
    q = np.clip(o, 0.0, 1.0)
    return np.concatenate([q[..., :3] * q[..., 3:4], q[..., 3:4]], -1)


def forward(cov, ends, params, width: int, height: int):
    """The output image (height x width x 4, float64)."""
    p = np.asarray(params, np.float64)
    n = cov.shape[0]
    s = _paths(cov, ends, p, width, height)
    b = _blur(s, float(p[n * PATH_PARAMS]))
    m = p[n * PATH_PARAMS + 1:n * PATH_PARAMS + 21].reshape(4, 5)
    f = _color_matrix(b, m)
    return f + (1.0 - f[..., 3:4]) * s


def grad_fd(cov, ends, params, width, height, grad_out, indices=None, h=1e-6):
    """Central differences of sum(grad_out · forward) for the given parameter indices."""
    p = np.asarray(params, np.float64)
    go = np.asarray(grad_out, np.float64)
    idx = range(len(p)) if indices is None else indices
    g = np.zeros(len(p))
    for i in idx:
        step = h * max(1.0, abs(p[i]))
        a, b = p.copy(), p.copy()
        a[i] += step
        b[i] -= step
        g[i] = (np.sum(go * forward(cov, ends, a, width, height)) -
                np.sum(go * forward(cov, ends, b, width, height))) / (2.0 * step)
    return g


# ---------------------------------------------------------------------------
# Analytic coverage: closed chains of k cubics (user-space points P0 .. P3k after a
# path's affine sx, ky, kx, sy, tx, ty), flattened into FLAT segments per cubic at
# uniform t and closed from P3k to P0; per pixel, the integral of the winding number
# over the pixel (signed area accumulation along rows), coverage min(1, |area|).
# ---------------------------------------------------------------------------
FLAT = 16


def geom_floats(k):
    return 6 + 2 * (3 * k + 1)


def flatten(g, k):
    a = np.asarray(g[:6], np.float64)
    pts = np.asarray(g[6:6 + 2 * (3 * k + 1)], np.float64).reshape(-1, 2)
    out = []
    for j in range(k * FLAT + 1):
        c = min(j // FLAT, k - 1)
        t = (j - c * FLAT) / FLAT
        u = 1.0 - t
        b = (u * u * u, 3 * u * u * t, 3 * u * t * t, t * t * t)
        x = sum(b[m] * pts[3 * c + m, 0] for m in range(4))
        y = sum(b[m] * pts[3 * c + m, 1] for m in range(4))
        out.append((a[0] * x + a[2] * y + a[4], a[1] * x + a[3] * y + a[5]))
    return out


def winding_area(verts, w, h):
    """Per pixel, the integral of the polygon's winding number (rows x columns).

    Along each pixel row, an edge's stretch is split where it crosses x = 0 and x = w:
    a piece left of the canvas is moved onto x = 0 (it changes the winding of every
    pixel of the row alike), a piece right of it is dropped, so the result is the
    exact integral over every canvas pixel whatever lies outside."""
    acc = np.zeros((h, w + 2))
    nv = len(verts)
    for j in range(nv):
        (ax, ay), (bx, by) = verts[j], verts[(j + 1) % nv]
        if ay == by:
            continue
        d = 1.0
        if ay > by:
            ax, ay, bx, by, d = bx, by, ax, ay, -1.0
        dxdy = (bx - ax) / (by - ay)
        for y in range(max(0, math.floor(ay)), min(h, math.ceil(by))):
            ya, yb = max(ay, y), min(by, y + 1)
            dy = yb - ya
            if dy <= 0:
                continue
            xa = ax + dxdy * (ya - ay)
            xb = ax + dxdy * (yb - ay)
            ts = sorted(t for t in ((0.0 - xa) / (xb - xa), (w - xa) / (xb - xa)) if 0.0 < t < 1.0) \
                if xb != xa else []
            cuts = [0.0] + ts + [1.0]
            for t0, t1 in zip(cuts[:-1], cuts[1:]):
                pa = min(max(xa + t0 * (xb - xa), 0.0), w)
                pb = min(max(xa + t1 * (xb - xa), 0.0), w)
                _row_piece(acc[y], pa, pb, dy * d * (t1 - t0))
    return np.cumsum(acc[:, :w], axis=1)


def _row_piece(row, xa, xb, dd):
    """One straight piece of an edge within a pixel row, x in [0, w], its share dd of
    the row's height (signed): adds to `row` the differences whose prefix sum is the
    piece's contribution to each pixel's winding area."""
    x0, x1 = min(xa, xb), max(xa, xb)
    x0f = math.floor(x0)
    x0i = int(x0f)
    x1c = math.ceil(x1)
    if x1c <= x0f + 1:
        xm = 0.5 * (xa + xb) - x0f
        row[x0i] += dd * (1 - xm)
        row[x0i + 1] += dd * xm
    else:
        # the trapezoid between x0 and x1, split by the pixel columns
        s = 1.0 / (x1 - x0)
        for xi in range(x0i, int(x1c) + 1):
            # area, within [xi, xi + 1), of the part of the row right of the edge
            # = integral over the row of clamp(xi + 1 - x(y), 0, 1) minus the same for
            # xi - 1 (the accumulation's difference form)
            row[xi] += dd * (_ramp(xi + 1, x0, x1, s) - _ramp(xi, x0, x1, s))


def _ramp(c, x0, x1, s):
    """Mean over the row of clamp(c - x, 0, 1) for x uniform in [x0, x1] (s = 1 / (x1 - x0))."""
    return (_antider(c, x1) - _antider(c, x0)) * s


def _antider(c, x):
    # G(x) = ∫_{-inf..x} clamp(c - t, 0, 1) dt up to a constant: slope 1 for t < c - 1,
    # a parabola on [c - 1, c], flat after c
    if x <= c - 1:
        return x
    if x >= c:
        return c - 0.5
    u = x - (c - 1)
    return (c - 1) + u - 0.5 * u * u


def analytic_coverage(geom, k, w, h):
    g = np.asarray(geom, np.float64).reshape(-1, geom_floats(k))
    return np.stack([np.minimum(1.0, np.abs(winding_area(flatten(gi, k), w, h))) for gi in g])


def forward_geom(geom, k, ends, params, width, height):
    """The analytic mode's output: coverage from the geometry, the rest as forward()."""
    cov = analytic_coverage(geom, k, width, height)
    return _forward_cov(cov, ends, params, width, height)


def _forward_cov(cov_f, ends, params, width, height):
    p = np.asarray(params, np.float64)
    n = cov_f.shape[0]
    s = _paths(cov_f * 255.0, ends, p, width, height)
    b = _blur(s, float(p[n * PATH_PARAMS]))
    m = p[n * PATH_PARAMS + 1:n * PATH_PARAMS + 21].reshape(4, 5)
    f = _color_matrix(b, m)
    return f + (1.0 - f[..., 3:4]) * s


def grad_fd_geom(geom, k, ends, params, width, height, grad_out, indices, h=1e-6):
    g0 = np.asarray(geom, np.float64).reshape(-1)
    go = np.asarray(grad_out, np.float64)
    g = np.zeros(len(g0))
    for i in indices:
        step = h * max(1.0, abs(g0[i]))
        a, b = g0.copy(), g0.copy()
        a[i] += step
        b[i] -= step
        g[i] = (np.sum(go * forward_geom(a, k, ends, params, width, height)) -
                np.sum(go * forward_geom(b, k, ends, params, width, height))) / (2.0 * step)
    return g
