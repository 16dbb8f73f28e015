"""Independent numpy formulations used to pin the oracle (and, through it, the GPU).

These are deliberately different algorithms from the oracle's restatements: window
sums from cumulative sums instead of running sums, min/max over shifted copies
instead of per-pixel loops, row-vectorised recurrences instead of scalar loops.
Float arithmetic is done in numpy float32/float64, which is IEEE without fused
multiply-add, the same arithmetic as the reference's Rust.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def as_u8(v: np.ndarray) -> np.ndarray:
    """Rust `f32 as u8`: saturate, NaN -> 0, truncate."""
    v = np.asarray(v)
    out = np.where(np.isnan(v), 0, np.clip(v, 0, 255))
    return np.floor(out).astype(np.uint8)


def box_sizes(sigma: float) -> list[int]:
    """create_box_gauss (box_blur.rs:37-71) in float32."""
    s = F(sigma)
    if not s > 0:
        return [1] * 5
    n = F(5.0)
    w_ideal = np.sqrt(F(12.0) * s * s / n) + F(1.0)
    wl = int(np.floor(w_ideal))
    if wl % 2 == 0:
        wl -= 1
    wu = wl + 2
    wf = F(wl)
    m_ideal = (F(12.0) * s * s - n * wf * wf - F(4.0) * n * wf - F(3.0) * n) / (F(-4.0) * wf - F(4.0))
    mr = float(m_ideal)
    m = int(np.floor(abs(mr) + 0.5)) * (1 if mr >= 0 else -1)
    m = max(m, 0)
    return [wl if i < m else wu for i in range(5)]


def _box_pass(img: np.ndarray, r: int, axis: int) -> np.ndarray:
    """Zero-padded window sum over [i-r, i+r] along axis, * f32(1/(2r+1)), RNE."""
    if r == 0:
        return img.copy()
    a = img.astype(np.int64)
    n = a.shape[axis]
    pad = [(0, 0)] * a.ndim
    pad[axis] = (r + 1, r)
    p = np.pad(a, pad)
    cs = np.cumsum(p, axis=axis)
    hi = np.take(cs, np.arange(2 * r + 1, 2 * r + 1 + n), axis=axis)
    lo = np.take(cs, np.arange(0, n), axis=axis)
    s = (hi - lo).astype(F)
    iarr = F(1.0) / F(2 * r + 1)
    return np.rint(s * iarr).astype(np.uint8)


def box_blur(img: np.ndarray, sx: float, sy: float) -> np.ndarray:
    bh, bv = box_sizes(sx), box_sizes(sy)
    out = img.copy()
    for i in range(5):
        out = _box_pass(out, (bv[i] - 1) // 2, axis=0)
        out = _box_pass(out, (bh[i] - 1) // 2, axis=1)
    return out


def iir_blur(img: np.ndarray, sx: float, sy: float) -> np.ndarray:
    """iir_blur.rs:47-146 with the recurrences vectorised across rows/columns."""
    h, w, _ = img.shape
    out = np.empty_like(img)

    def coef(s):
        lam = (s * s) / (2.0 * 4.0)
        return lam, (1.0 + 2.0 * lam - np.sqrt(1.0 + 4.0 * lam)) / (2.0 * lam)

    lx, dx = coef(sx) if sx > 0 else (1.0, 1.0)
    ly, dy = coef(sy) if sy > 0 else (1.0, 1.0)
    base = np.sqrt(dx * dy) / np.sqrt(lx * ly)
    post = base * base
    post = post * post
    post = post * post
    for c in range(4):
        b = img[..., c].astype(np.float64) / 255.0
        if sx > 0:
            for _ in range(4):
                for x in range(1, w):
                    b[:, x] = b[:, x] + dx * b[:, x - 1]
                for x in range(w - 1, 0, -1):
                    b[:, x - 1] = b[:, x - 1] + dx * b[:, x]
        if sy > 0:
            for _ in range(4):
                for y in range(1, h):
                    b[y] = b[y] + dy * b[y - 1]
                for y in range(h - 1, 0, -1):
                    b[y - 1] = b[y - 1] + dy * b[y]
        b = b * post
        out[..., c] = as_u8(b * 255.0)
    return out


def morphology(img: np.ndarray, erode: bool, rx: float, ry: float) -> np.ndarray:
    h, w, _ = img.shape
    cols = min(int(np.ceil(F(rx))) * 2, w)
    rows = min(int(np.ceil(F(ry))) * 2, h)
    tx, ty = cols // 2, rows // 2
    init = 255 if erode else 0
    acc = np.full(img.shape, init, dtype=np.int32)
    a = img.astype(np.int32)
    for oy in range(rows):
        for ox in range(cols):
            dx, dy = ox - tx, oy - ty
            sh = np.full(img.shape, init, dtype=np.int32)
            ys0, ys1 = max(0, -dy), min(h, h - dy)
            xs0, xs1 = max(0, -dx), min(w, w - dx)
            if ys1 > ys0 and xs1 > xs0:
                sh[ys0:ys1, xs0:xs1] = a[ys0 + dy:ys1 + dy, xs0 + dx:xs1 + dx]
            acc = np.minimum(acc, sh) if erode else np.maximum(acc, sh)
    return acc.astype(np.uint8)


def composite_arithmetic(s1, s2, dst, k1, k2, k3, k4):
    k1, k2, k3, k4 = F(k1), F(k2), F(k3), F(k4)
    i1 = s1.astype(F) / F(255.0)
    i2 = s2.astype(F) / F(255.0)
    r = k1 * i1 * i2 + k2 * i1 + k3 * i2 + k4
    a = np.where(r[..., 3] > 1, F(1), np.where(r[..., 3] >= 0, r[..., 3], F(0)))
    a4 = a[..., None]
    col = np.where(r[..., :3] > a4, a4, np.where(r[..., :3] >= 0, r[..., :3], F(0)))
    out = dst.copy()
    keep = ~((a == 0) | ((a > 0) & (a.view(np.uint32) <= 4)))
    res = np.concatenate([as_u8(col * F(255.0)), as_u8(a4 * F(255.0))], axis=-1)
    out[keep] = res[keep]
    return out


def park_miller(seed: int, steps: int) -> int:
    """The textbook minimal-standard generator x <- 16807 x mod (2^31 - 1)."""
    x = seed
    for _ in range(steps):
        x = (16807 * x) % 2147483647
    return x
