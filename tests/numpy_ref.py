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


# ---------------------------------------------------------------------------
# Helpers with Rust's numeric semantics, vectorised
# ---------------------------------------------------------------------------
def bound(lo, v, hi):
    """filter/mod.rs:242-254 f32_bound: v > hi -> hi, v >= lo -> v, else (incl. NaN) lo."""
    v = np.asarray(v, F)
    return np.where(v > F(hi), F(hi), np.where(v >= F(lo), v, F(lo))).astype(F)


def approx_zero_ulps4(v):
    """strict-num approx_zero_ulps(4) for f32: exactly zero, or positive within 4 ULPs of 0."""
    v = np.asarray(v, F)
    return (v == 0) | ((v > 0) & (v.view(np.uint32) <= 4))


def _libm_f32(name):
    import ctypes
    fn = getattr(ctypes.CDLL("libm.so.6"), name)
    fn.restype = ctypes.c_float
    return fn


def libm_powf(x, y):
    """The platform powf (what Rust's f32::powf calls on Linux), elementwise."""
    import ctypes
    fn = _libm_f32("powf")
    fn.argtypes = [ctypes.c_float, ctypes.c_float]
    x = np.asarray(x, F)
    out = np.empty(x.shape, F)
    for i, xv in np.ndenumerate(x):
        out[i] = fn(float(xv), float(y))
    return out


def libm_f32(name, x):
    import ctypes
    fn = _libm_f32(name)
    fn.argtypes = [ctypes.c_float]
    return F(fn(float(F(x))))


def to_radians(deg):
    """Rust f32::to_radians: deg * (PI / 180) with the constant folded in f32."""
    return F(deg) * (F(np.pi) / F(180.0))


# ---------------------------------------------------------------------------
# feConvolveMatrix, from the SVG 1.1 definition
#   RESULT(X,Y) = (sum_{I<orderY} sum_{J<orderX} SOURCE(X - targetX + J, Y - targetY + I)
#                  * kernelMatrix(orderX - J - 1, orderY - I - 1)) / divisor + bias * ALPHA(X,Y)
# summed I-major in f32 (the order the reference accumulates in), colours clamped
# to [0, alpha] (premultiplied) or [0, 1] * alpha with preserveAlpha.
# ---------------------------------------------------------------------------
EDGE_NONE, EDGE_DUPLICATE, EDGE_WRAP = 0, 1, 2


def convolve_matrix(img, kernel, divisor, bias, target_x, target_y, edge, preserve_alpha):
    """kernel: (orderY, orderX) array, kernel[row][col] as written in kernelMatrix."""
    kern = np.asarray(kernel, F)
    order_y, order_x = kern.shape
    h, w, _ = img.shape
    src = img.astype(F) / F(255.0)
    ys, xs = np.mgrid[0:h, 0:w]
    acc = np.zeros((h, w, 4), F)
    for i in range(order_y):
        for j in range(order_x):
            sx, sy = xs - target_x + j, ys - target_y + i
            if edge == EDGE_NONE:
                valid = (sx >= 0) & (sx < w) & (sy >= 0) & (sy < h)
            else:
                valid = np.ones((h, w), bool)
            if edge == EDGE_DUPLICATE:
                sx, sy = np.clip(sx, 0, w - 1), np.clip(sy, 0, h - 1)
            elif edge == EDGE_WRAP:
                sx, sy = np.mod(sx, w), np.mod(sy, h)
            else:
                sx, sy = np.clip(sx, 0, w - 1), np.clip(sy, 0, h - 1)
            k = kern[order_y - i - 1, order_x - j - 1]
            term = src[sy, sx] * k
            acc = np.where(valid[..., None], acc + term, acc).astype(F)
    if preserve_alpha:
        new_a = src[..., 3]
    else:
        new_a = (acc[..., 3] / F(divisor) + F(bias)).astype(F)
    bna = bound(0.0, new_a, 1.0)
    col = (acc[..., :3] / F(divisor) + F(bias) * new_a[..., None]).astype(F)
    if preserve_alpha:
        col = (bound(0.0, col, 1.0) * bna[..., None]).astype(F)
    else:
        col = np.where(col > bna[..., None], bna[..., None], np.where(col >= 0, col, F(0))).astype(F)
    out = np.empty_like(img)
    out[..., :3] = as_u8(col * F(255.0) + F(0.5))
    out[..., 3] = as_u8(bna * F(255.0) + F(0.5))
    return out


# ---------------------------------------------------------------------------
# feColorMatrix, from the SVG 1.1 definition: [R' G' B' A'] = M (4x5) . [R G B A 1]
# on the demultiplied colour in [0, 1], each row summed left to right in f32.
# saturate / hueRotate use the spec's 3x3 matrices and keep alpha;
# luminanceToAlpha writes A' = 0.2125 R + 0.7154 G + 0.0721 B and zero colour.
# ---------------------------------------------------------------------------
def saturate_matrix(s):
    """The spec's saturate matrix: entry = base + coefficient * s, in f32."""
    s = max(F(s), F(0.0))
    e = [[(0.213, 0.787), (0.715, -0.715), (0.072, -0.072)],
         [(0.213, -0.213), (0.715, 0.285), (0.072, -0.072)],
         [(0.213, -0.213), (0.715, -0.715), (0.072, 0.928)]]
    return np.array([[F(b) + F(c) * s for b, c in row] for row in e], F)


def hue_rotate_matrix(deg):
    rad = to_radians(deg)
    a1, a2 = libm_f32("cosf", rad), libm_f32("sinf", rad)
    # (base, cos coefficient, sin coefficient) per entry, the spec's hueRotate matrix
    e = [[(0.213, 0.787, -0.213), (0.715, -0.715, -0.715), (0.072, -0.072, 0.928)],
         [(0.213, -0.213, 0.143), (0.715, 0.285, 0.140), (0.072, -0.072, -0.283)],
         [(0.213, -0.213, -0.787), (0.715, -0.715, 0.715), (0.072, 0.928, 0.072)]]

    def term(b, c, s):
        return (F(b) + F(c) * a1) + F(s) * a2
    return np.array([[term(*x) for x in row] for row in e], F)


def color_matrix(img, kind, values):
    """img unpremultiplied RGBA8.  kind 0 matrix, 1 saturate, 2 hueRotate, 3 luminanceToAlpha."""
    c = img.astype(F) / F(255.0)
    out = img.copy()
    if kind == 0:
        m = np.asarray(values, F).reshape(4, 5)
        for row in range(4):
            v = c[..., 0] * m[row, 0]
            for col in range(1, 4):
                v = v + c[..., col] * m[row, col]
            v = v + m[row, 4]
            out[..., row] = as_u8(bound(0.0, v, 1.0) * F(255.0))
    elif kind in (1, 2):
        m = saturate_matrix(values[0]) if kind == 1 else hue_rotate_matrix(values[0])
        for row in range(3):
            v = c[..., 0] * m[row, 0] + c[..., 1] * m[row, 1] + c[..., 2] * m[row, 2]
            out[..., row] = as_u8(bound(0.0, v, 1.0) * F(255.0))
    else:
        v = c[..., 0] * F(0.2125) + c[..., 1] * F(0.7154) + c[..., 2] * F(0.0721)
        out[..., :3] = 0
        out[..., 3] = as_u8(bound(0.0, v, 1.0) * F(255.0))
    return out


# ---------------------------------------------------------------------------
# feComponentTransfer, from the SVG 1.1 definitions of table, discrete, linear and
# gamma (C in [0, 1] the demultiplied channel), evaluated for one channel value.
# ---------------------------------------------------------------------------
def transfer(func, cu):
    """func: ("table"|"discrete", values) | ("linear", slope, intercept) | ("gamma", amp, exp, off)."""
    c = F(cu) / F(255.0)
    kind = func[0]
    if kind == "table":
        v = [F(x) for x in func[1]]
        n = len(v) - 1
        k = min(int(np.floor(c * F(n))), n)
        r = v[k] if k == n else v[k] + (c - F(k) / F(n)) * F(n) * (v[k + 1] - v[k])
    elif kind == "discrete":
        v = [F(x) for x in func[1]]
        n = len(v)
        r = v[min(int(np.floor(c * F(n))), n - 1)]
    elif kind == "linear":
        r = F(func[1]) * c + F(func[2])
    else:
        r = F(func[1]) * libm_powf(c, func[2])[()] + F(func[3])
    return int(as_u8(bound(0.0, F(r), 1.0) * F(255.0)))


# ---------------------------------------------------------------------------
# Lighting, from the SVG 1.1 definitions.  The surface normal uses the spec's
# Sobel operators; its nine border cases follow one rule: along each axis take the
# central difference when both neighbours exist, else the one-sided difference
# against the pixel itself, weight the cross-axis rows/columns (1, 2, 1) truncated
# to those that exist, and scale by 2 / (sum of weights * difference span).
# ---------------------------------------------------------------------------
def sobel_normal(alpha):
    """(factor_x, factor_y, nx_sum, ny_sum) per pixel as f32/int arrays, sums of the
    integer Sobel operator on alpha (the surface is -surfaceScale * sums * factor)."""
    a = alpha.astype(np.int64)
    h, w = a.shape
    fx = np.zeros((h, w), F)
    fy = np.zeros((h, w), F)
    nx = np.zeros((h, w), np.int64)
    ny = np.zeros((h, w), np.int64)
    wt = {-1: 1, 0: 2, 1: 1}
    for y in range(h):
        for x in range(w):
            xl, xr = (x - 1 if x > 0 else x), (x + 1 if x < w - 1 else x)
            yt, yb = (y - 1 if y > 0 else y), (y + 1 if y < h - 1 else y)
            rows = [d for d in (-1, 0, 1) if 0 <= y + d < h]
            cols = [d for d in (-1, 0, 1) if 0 <= x + d < w]
            nx[y, x] = sum(wt[d] * (a[y + d, xr] - a[y + d, xl]) for d in rows)
            ny[y, x] = sum(wt[d] * (a[yb, x + d] - a[yt, x + d]) for d in cols)
            fx[y, x] = F(2.0) / F(sum(wt[d] for d in rows) * (xr - xl))
            fy[y, x] = F(2.0) / F(sum(wt[d] for d in cols) * (yb - yt))
    return fx, fy, nx, ny


def _norm3(x, y, z):
    """Vector3::normalized().unwrap_or(v): (v / |v|) unless |v| is within 4 ULPs of 0."""
    ln = np.sqrt(x * x + y * y + z * z).astype(F)
    z0 = approx_zero_ulps4(ln)
    safe = np.where(z0, F(1), ln).astype(F)
    return (np.where(z0, x, x / safe).astype(F), np.where(z0, y, y / safe).astype(F),
            np.where(z0, z, z / safe).astype(F))


def lighting(src, specular, surface_scale, constant, spec_exp, color, light):
    """light: dict kind 'distant' (azimuth, elevation), 'point' (x, y, z) or 'spot'
    (x, y, z, px, py, pz, exp, cone or None).  Returns the RGBA8 output (diffuse:
    alpha 255; specular: alpha = max(r, g, b)).  Images smaller than 3x3 return None."""
    h, w, _ = src.shape
    if w < 3 or h < 3:
        return None
    alpha = src[..., 3]
    fx, fy, nxs, nys = sobel_normal(alpha)
    ss = F(surface_scale)
    s255 = F(ss / F(255.0))
    nrm_x = (-nxs).astype(np.int16).astype(F)
    nrm_y = (-nys).astype(np.int16).astype(F)
    flat = approx_zero_ulps4(nrm_x) & approx_zero_ulps4(nrm_y)
    ys, xs = np.mgrid[0:h, 0:w]
    if light["kind"] == "distant":
        az, el = to_radians(light["azimuth"]), to_radians(light["elevation"])
        ca, sa, ce, se = (libm_f32("cosf", az), libm_f32("sinf", az), libm_f32("cosf", el), libm_f32("sinf", el))
        lx = np.full((h, w), ca * ce, F)
        ly = np.full((h, w), sa * ce, F)
        lz = np.full((h, w), se, F)
    else:
        nz = (alpha.astype(F) / F(255.0) * ss).astype(F)
        lx, ly, lz = _norm3((F(light["x"]) - xs.astype(F)).astype(F), (F(light["y"]) - ys.astype(F)).astype(F),
                            (F(light["z"]) - nz).astype(F))
    col = np.zeros((h, w, 3), np.uint8)
    col[:] = np.asarray(color[:3], np.uint8)
    if light["kind"] == "spot":
        dx, dy, dz = (F(light["px"]) - F(light["x"]), F(light["py"]) - F(light["y"]), F(light["pz"]) - F(light["z"]))
        dx, dy, dz = (np.asarray(v)[()] for v in _norm3(np.asarray(dx, F), np.asarray(dy, F), np.asarray(dz, F)))
        mlds = (-(lx * dx + ly * dy + lz * dz)).astype(F)
        black = mlds <= 0
        if light.get("cone") is not None:
            black |= mlds < libm_f32("cosf", to_radians(light["cone"]))
        f = libm_powf(np.where(black, F(0), mlds), light["exp"])
        lit = as_u8(bound(0.0, np.asarray(color[:3], F)[None, None, :] * f[..., None], 255.0) + F(0.5))
        col = np.where(black[..., None], np.uint8(0), lit)
    # the normal (n.x, n.y, 1) scaled by surfaceScale / 255 and the Sobel factor
    nx = (nrm_x * s255 * fx).astype(F)
    ny = (nrm_y * s255 * fy).astype(F)
    nlen = np.sqrt(nx * nx + ny * ny + F(1.0)).astype(F)
    if not specular:
        k = np.where(flat, lz, (nx * lx + ny * ly + F(1.0) * lz) / nlen).astype(F)
        factor = (F(constant) * k).astype(F)
    else:
        hx, hy, hz = lx + F(0.0), ly + F(0.0), lz + F(1.0)
        hl = np.sqrt(hx * hx + hy * hy + hz * hz).astype(F)
        ndh = np.where(flat, hz / hl, (nx * hx + ny * hy + F(1.0) * hz) / nlen / hl).astype(F)
        one = F(spec_exp) == F(1.0) or abs(int(F(spec_exp).view(np.uint32)) - int(F(1.0).view(np.uint32))) <= 4
        k = ndh if one else libm_powf(ndh, spec_exp)
        factor = np.where(approx_zero_ulps4(hl), F(0), F(constant) * k).astype(F)
    out = np.zeros((h, w, 4), np.uint8)
    out[..., :3] = as_u8(bound(0.0, col.astype(F) * factor[..., None], 255.0) + F(0.5))
    out[..., 3] = out[..., :3].max(axis=-1) if specular else 255
    return out


# ---------------------------------------------------------------------------
# feDisplacementMap, from the SVG 1.1 definition
#   P'(x,y) <- P(x + scale * (XC(x,y) - .5), y + scale * (YC(x,y) - .5))
# with the canvas scale folded in, the displaced position rounded half away from
# zero; pixels displaced outside the image keep the destination's value.
# ---------------------------------------------------------------------------
def rust_round(v):
    v = np.asarray(v, F)
    t = np.trunc(v)
    frac = np.abs(v - t)
    return (t + np.where(frac >= F(0.5), np.sign(v), F(0))).astype(F)


def displacement_map(src, map_, dst, sx, sy, scale, x_channel, y_channel):
    h, w, _ = src.shape
    ys, xs = np.mgrid[0:h, 0:w]
    dx = map_[..., x_channel].astype(F) / F(255.0) - F(0.5)
    dy = map_[..., y_channel].astype(F) / F(255.0) - F(0.5)
    fx = rust_round(xs.astype(F) + dx * F(sx) * F(scale))
    fy = rust_round(ys.astype(F) + dy * F(sy) * F(scale))
    ox = np.clip(np.nan_to_num(fx), -2**31, 2**31 - 1).astype(np.int64)
    oy = np.clip(np.nan_to_num(fy), -2**31, 2**31 - 1).astype(np.int64)
    valid = (ox >= 0) & (ox < w) & (oy >= 0) & (oy < h)
    out = dst.copy()
    out[valid] = src[oy[valid], ox[valid]]
    return out


# ---------------------------------------------------------------------------
# feTurbulence lattice set-up: init() of the SVG 1.1 reference implementation, with
# the minimal-standard generator written as x -> 16807 x mod (2^31 - 1) (the spec
# uses Schrage's decomposition of the same product).
# ---------------------------------------------------------------------------
def turbulence_init(seed: int):
    M, BS = 2147483647, 256
    if seed <= 0:
        seed = (-seed) % (M - 1) + 1
    if seed > M - 1:
        seed = M - 1
    lat = list(range(BS)) + [0] * (BS + 2)
    grad = np.zeros((4, BS + BS + 2, 2), np.float64)
    for k in range(4):
        for i in range(BS):
            for j in range(2):
                seed = (16807 * seed) % M
                grad[k, i, j] = float((seed % (BS + BS)) - BS) / BS
            s = float(np.sqrt(grad[k, i, 0] * grad[k, i, 0] + grad[k, i, 1] * grad[k, i, 1]))
            grad[k, i, 0] /= s
            grad[k, i, 1] /= s
    for i in range(BS - 1, 0, -1):
        kk = lat[i]
        seed = (16807 * seed) % M
        j = seed % BS
        lat[i] = lat[j]
        lat[j] = kk
    for i in range(BS + 2):
        lat[BS + i] = lat[i]
        grad[:, BS + i, :] = grad[:, i, :]
    return np.array(lat, np.int32), grad
