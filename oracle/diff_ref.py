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


def _color_matrix(b, m):
    a = b[..., 3]
    safe = np.where(a > 0, a, 1.0)
    v = np.stack([np.where(a > 0, b[..., 0] / safe, 0.0), np.where(a > 0, b[..., 1] / safe, 0.0),
                  np.where(a > 0, b[..., 2] / safe, 0.0), a], -1)
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
