"""Pins the oracle's convolve matrix, colour matrix, component transfer, lighting,
displacement map and turbulence lattice (CPU only).

Each oracle function is checked two ways:

* against an independent numpy formulation written from the SVG 1.1 definition
  (numpy_ref.py: shifted-array sums instead of per-pixel loops, a general rule for
  the Sobel border cases instead of nine hand-written ones, the minimal-standard
  generator as a modular product instead of Schrage's decomposition), on random,
  ragged and tiny inputs;
* against closed forms chosen so that a plausible mistake fails: an identity or
  shift kernel (an unflipped kernel or a wrong target offset moves the image), a
  linear alpha ramp whose Sobel normal is the same at every pixel, border cases
  included (a swapped sign or factor changes the lit value), channel permutations
  (a transposed matrix permutes differently), an exact displacement shift.
"""
import numpy as np
import pytest

from oracle import bindings as orc
from resvg_amd import abi
from resvg_amd.filters import FilterGraph, distant_light, point_light, spot_light

import numpy_ref as npr

F = np.float32


def rand_premul(h, w, seed):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    a = px[..., 3:4].astype(np.uint16)
    px[..., :3] = (px[..., :3].astype(np.uint16) * a // 255).astype(np.uint8)
    return px


def rand_u8(h, w, seed):
    return np.random.default_rng(seed).integers(0, 256, size=(h, w, 4), dtype=np.uint8)


def smooth_alpha(h, w, seed):
    """Blobby alpha (flat areas, slopes and peaks) on random colour."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    a = np.zeros((h, w))
    for _ in range(4):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(1.5, max(2.0, min(w, h) / 2))
        a += np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
    px = rand_u8(h, w, seed + 1)
    px[..., 3] = np.clip(a * 255, 0, 255).astype(np.uint8)
    return px


# ---------------------------------------------------------------------------
# convolve matrix (convolve_matrix.rs:15-111)
# ---------------------------------------------------------------------------
def conv_prim(kernel, divisor=None, bias=0.0, tx=None, ty=None, edge=abi.EDGE_DUPLICATE, preserve=False):
    k = np.asarray(kernel, dtype=np.float64)
    rows, cols = k.shape
    g = FilterGraph((0.0, 0.0, 1.0, 1.0))
    g.convolve_matrix(cols, rows, k.ravel().tolist(), divisor=divisor, bias=bias, target_x=tx, target_y=ty,
                      edge_mode=edge, preserve_alpha=preserve)
    p = g.primitives[0]
    return p, g.values, (p.u.convolve.divisor, p.u.convolve.bias, p.u.convolve.target_x, p.u.convolve.target_y)


def oracle_conv(px, kernel, **kw):
    p, vals, (div, bias, tx, ty) = conv_prim(kernel, **kw)
    out = px.copy()
    orc.convolve_matrix(out, p, vals)
    return out, (div, bias, tx, ty)


CONV_SHAPES = [(1, 1), (2, 7), (5, 3), (13, 17), (4, 4)]
CONV_ORDERS = [(1, 1), (3, 3), (5, 5), (1, 3), (4, 2)]


@pytest.mark.parametrize("shape", CONV_SHAPES)
@pytest.mark.parametrize("order", CONV_ORDERS)
@pytest.mark.parametrize("edge", [abi.EDGE_NONE, abi.EDGE_DUPLICATE, abi.EDGE_WRAP])
@pytest.mark.parametrize("preserve", [False, True])
def test_convolve_equals_spec_sum(shape, order, edge, preserve):
    rows, cols = order
    seed = hash((shape, order, edge, preserve)) & 0xffff
    rng = np.random.default_rng(seed)
    kernel = rng.integers(-4, 5, size=(rows, cols)).astype(np.float64)
    px = rand_premul(*shape, seed) if not preserve else rand_u8(*shape, seed)
    tx, ty = int(rng.integers(0, cols)), int(rng.integers(0, rows))
    for divisor, bias in ((None, 0.0), (float(rng.uniform(0.5, 7.0)), float(rng.uniform(-0.4, 0.4)))):
        out, (div, b, tx_, ty_) = oracle_conv(px, kernel, divisor=divisor, bias=bias, tx=tx, ty=ty, edge=edge,
                                              preserve=preserve)
        exp = npr.convolve_matrix(px, kernel, div, b, tx_, ty_, edge, preserve)
        assert np.array_equal(out, exp), (divisor, bias)


@pytest.mark.parametrize("order,target", [((3, 3), (1, 1)), ((3, 3), (0, 2)), ((5, 4), (3, 1)), ((2, 3), (2, 0))])
def test_convolve_identity_kernel(order, target):
    # a single 1 where the flipped kernel meets the target pixel: the exact identity
    rows, cols = order
    tx, ty = target
    k = np.zeros(order)
    k[rows - 1 - ty, cols - 1 - tx] = 1.0
    px = rand_premul(9, 11, 4)
    for edge in (abi.EDGE_NONE, abi.EDGE_DUPLICATE, abi.EDGE_WRAP):
        out, _ = oracle_conv(px, k, divisor=1.0, tx=tx, ty=ty, edge=edge)
        assert np.array_equal(out, px)


@pytest.mark.parametrize("order,target", [((3, 3), (1, 1)), ((3, 5), (0, 0)), ((4, 2), (1, 3))])
def test_convolve_shift_kernel(order, target):
    # kernelMatrix(0, 0) = 1 picks SOURCE(X - targetX + orderX - 1, Y - targetY + orderY - 1):
    # an integer shift by (orderX - 1 - targetX, orderY - 1 - targetY)
    rows, cols = order
    tx, ty = target
    k = np.zeros(order)
    k[0, 0] = 1.0
    px = rand_premul(10, 12, 8)
    sx, sy = cols - 1 - tx, rows - 1 - ty
    h, w, _ = px.shape
    out, _ = oracle_conv(px, k, divisor=1.0, tx=tx, ty=ty, edge=abi.EDGE_NONE)
    exp = np.zeros_like(px)
    exp[:h - sy, :w - sx] = px[sy:, sx:]
    assert np.array_equal(out, exp)
    out, _ = oracle_conv(px, k, divisor=1.0, tx=tx, ty=ty, edge=abi.EDGE_DUPLICATE)
    yy = np.minimum(np.arange(h) + sy, h - 1)
    xx = np.minimum(np.arange(w) + sx, w - 1)
    assert np.array_equal(out, px[yy][:, xx])
    out, _ = oracle_conv(px, k, divisor=1.0, tx=tx, ty=ty, edge=abi.EDGE_WRAP)
    assert np.array_equal(out, np.roll(px, (-sy, -sx), axis=(0, 1)))


def test_convolve_divisor_and_bias_closed_form():
    # a constant opaque grey under a kernel summing to d, divided by d: unchanged
    # inside; a bias b adds b * alpha = b to every colour channel (alpha 1)
    px = np.full((6, 6, 4), 255, np.uint8)
    px[..., :3] = 100
    k = np.array([[1.0, 2.0, 1.0], [2.0, 4.0, 2.0], [1.0, 2.0, 1.0]])
    out, _ = oracle_conv(px, k, edge=abi.EDGE_DUPLICATE)
    assert np.array_equal(out, px)
    out, _ = oracle_conv(px, k, bias=0.2, edge=abi.EDGE_DUPLICATE, preserve=True)
    assert (out[..., 3] == 255).all()
    assert (np.abs(out[..., :3].astype(int) - int(100 + 0.2 * 255 + 0.5)) <= 1).all()


# ---------------------------------------------------------------------------
# colour matrix (color_matrix.rs:11-110)
# ---------------------------------------------------------------------------
def oracle_cm(px, kind, values):
    g = FilterGraph((0.0, 0.0, 1.0, 1.0))
    g.color_matrix(kind, values)
    out = px.copy()
    orc.color_matrix(out, g.primitives[0])
    return out


@pytest.mark.parametrize("shape", [(1, 1), (3, 7), (16, 16)])
@pytest.mark.parametrize("case", [
    (abi.COLOR_MATRIX_MATRIX, None), (abi.COLOR_MATRIX_SATURATE, [0.0]), (abi.COLOR_MATRIX_SATURATE, [0.3]),
    (abi.COLOR_MATRIX_SATURATE, [1.0]), (abi.COLOR_MATRIX_SATURATE, [2.5]), (abi.COLOR_MATRIX_SATURATE, [-1.0]),
    (abi.COLOR_MATRIX_HUE_ROTATE, [0.0]), (abi.COLOR_MATRIX_HUE_ROTATE, [30.0]),
    (abi.COLOR_MATRIX_HUE_ROTATE, [90.0]), (abi.COLOR_MATRIX_HUE_ROTATE, [200.0]),
    (abi.COLOR_MATRIX_HUE_ROTATE, [-45.0]), (abi.COLOR_MATRIX_LUMINANCE_TO_ALPHA, [])])
def test_color_matrix_equals_spec(shape, case):
    kind, values = case
    for seed in range(3):
        px = rand_u8(*shape, seed)
        vals = values if values is not None else \
            np.random.default_rng(seed).uniform(-1, 1, 20).astype(np.float32).tolist()
        assert np.array_equal(oracle_cm(px, kind, vals), npr.color_matrix(px, kind, vals))


def test_color_matrix_identity_and_permutation():
    px = rand_u8(8, 32, 3)
    px[0, :, :] = np.arange(256).reshape(8, 32)[0:1, :, None]  # every channel value appears
    ident = np.zeros((4, 5))
    ident[range(4), range(4)] = 1.0
    assert np.array_equal(oracle_cm(px, abi.COLOR_MATRIX_MATRIX, ident.ravel().tolist()), px)
    # r' = b, g' = r, b' = g, a' = a: a transposed evaluation would give r' = g, g' = b, b' = r
    perm = np.zeros((4, 5))
    perm[0, 2] = perm[1, 0] = perm[2, 1] = perm[3, 3] = 1.0
    out = oracle_cm(px, abi.COLOR_MATRIX_MATRIX, perm.ravel().tolist())
    assert np.array_equal(out, px[..., [2, 0, 1, 3]])
    # the constant column: r' = 0.5 everywhere, as u8 127 (truncation)
    const = np.zeros((4, 5))
    const[0, 4] = 0.5
    const[3, 3] = 1.0
    out = oracle_cm(px, abi.COLOR_MATRIX_MATRIX, const.ravel().tolist())
    assert (out[..., 0] == 127).all() and (out[..., 1:3] == 0).all() and np.array_equal(out[..., 3], px[..., 3])


def test_color_matrix_saturate_hue_luminance_closed_forms():
    px = rand_u8(8, 8, 5)
    # saturate(0) is the luminance grey: r' = g' = b'
    out = oracle_cm(px, abi.COLOR_MATRIX_SATURATE, [0.0])
    assert np.array_equal(out[..., 0], out[..., 1]) and np.array_equal(out[..., 1], out[..., 2])
    assert np.array_equal(out[..., 3], px[..., 3])
    # saturate(1) and hueRotate(0) are the identity up to the matrix's f32 rounding
    for kind, v in ((abi.COLOR_MATRIX_SATURATE, [1.0]), (abi.COLOR_MATRIX_HUE_ROTATE, [0.0])):
        out = oracle_cm(px, kind, v)
        assert np.abs(out.astype(int) - px.astype(int)).max() <= 1
        assert np.array_equal(out[..., 3], px[..., 3])
    # luminanceToAlpha of a grey g: weights summing to 1 (just under it in f32), so
    # alpha g or g - 1, colour 0
    grey = np.zeros((1, 256, 4), np.uint8)
    grey[0, :, :3] = np.arange(256)[:, None]
    grey[0, :, 3] = 255
    out = oracle_cm(grey, abi.COLOR_MATRIX_LUMINANCE_TO_ALPHA, [])
    d = np.arange(256) - out[0, :, 3].astype(int)
    assert ((d == 0) | (d == 1)).all() and out[0, 0, 3] == 0 and out[0, 128, 3] == 128
    assert not out[..., :3].any()


# ---------------------------------------------------------------------------
# component transfer (component_transfer.rs:10-72): every input value
# ---------------------------------------------------------------------------
TRANSFER_SETS = [
    [("table", [0.0, 1.0]), ("discrete", [0.2, 0.6, 1.0]), ("linear", 0.5, 0.25), ("gamma", 2.0, 3.0, 0.1)],
    [("table", [1.0, 0.0, 0.5, 0.25, 0.9]), ("discrete", [0.9]), ("linear", -1.0, 1.0), ("gamma", 1.0, 0.45, 0.0)],
    [("table", [0.3]), ("discrete", [0.0, 1.0, 0.0, 1.0, 0.5]), ("linear", 2.5, -0.7), ("gamma", 0.7, 1.0, -0.2)],
    [("gamma", 1.5, 0.8, 0.05), ("table", [0.1, 0.9, 0.3]), ("linear", 0.0, 0.5), ("discrete", [0.25, 0.75])],
]


@pytest.mark.parametrize("funcs", TRANSFER_SETS, ids=[f"set{i}" for i in range(len(TRANSFER_SETS))])
def test_component_transfer_all_inputs(funcs):
    g = FilterGraph((0.0, 0.0, 1.0, 1.0))
    g.component_transfer(funcs)
    px = np.zeros((16, 16, 4), np.uint8)
    px[...] = np.arange(256, dtype=np.uint8).reshape(16, 16)[..., None]
    out = px.copy()
    orc.component_transfer(out, g.primitives[0], g.values)
    for ch, f in enumerate(funcs):
        exp = np.array([npr.transfer(f, v) for v in range(256)], np.uint8).reshape(16, 16)
        assert np.array_equal(out[..., ch], exp), (ch, f)


def test_component_transfer_identity_and_empty_tables_keep_input():
    g = FilterGraph((0.0, 0.0, 1.0, 1.0))
    g.component_transfer([None, ("table", []), ("discrete", []), ("identity",)])
    px = rand_u8(9, 5, 2)
    out = px.copy()
    orc.component_transfer(out, g.primitives[0], g.values)
    assert np.array_equal(out, px)


# ---------------------------------------------------------------------------
# lighting (lighting.rs:132-485)
# ---------------------------------------------------------------------------
LIGHTS = {
    "distant": (distant_light(30.0, 40.0), {"kind": "distant", "azimuth": 30.0, "elevation": 40.0}),
    "distant_low": (distant_light(200.0, 5.0), {"kind": "distant", "azimuth": 200.0, "elevation": 5.0}),
    "point": (point_light(2.5, 1.0, 6.0), {"kind": "point", "x": 2.5, "y": 1.0, "z": 6.0}),
    "spot": (spot_light(1.0, 2.0, 9.0, 4.0, 5.0, 0.0, 2.0),
             {"kind": "spot", "x": 1.0, "y": 2.0, "z": 9.0, "px": 4.0, "py": 5.0, "pz": 0.0, "exp": 2.0,
              "cone": None}),
    "spot_cone": (spot_light(3.0, 3.0, 5.0, 3.5, 4.0, 0.0, 1.0, cone=30.0),
                  {"kind": "spot", "x": 3.0, "y": 3.0, "z": 5.0, "px": 3.5, "py": 4.0, "pz": 0.0, "exp": 1.0,
                   "cone": 30.0}),
}


def oracle_lighting(src, specular, ss, kc, exp, color, light):
    dst = np.zeros_like(src)
    orc.lighting(src, dst, specular, ss, kc, exp, color, light)
    return dst


@pytest.mark.parametrize("shape", [(3, 3), (4, 5), (7, 6), (9, 12)])
@pytest.mark.parametrize("light", list(LIGHTS))
@pytest.mark.parametrize("mode", [("diffuse", 1.0), ("specular", 1.0), ("specular", 8.0), ("specular", 0.5)])
def test_lighting_equals_spec(shape, light, mode):
    specular = mode[0] == "specular"
    src = smooth_alpha(*shape, seed=hash((shape, light)) & 0xffff)
    abi_light, spec_light = LIGHTS[light]
    for ss, kc in ((1.0, 1.0), (4.5, 0.7)):
        got = oracle_lighting(src, specular, ss, kc, mode[1], (250, 200, 120), abi_light)
        exp = npr.lighting(src, specular, ss, kc, mode[1], (250, 200, 120), spec_light)
        assert np.array_equal(got, exp), (ss, kc)


def test_lighting_below_3x3_does_nothing():
    for shape in ((2, 5), (5, 2), (1, 1)):
        src = smooth_alpha(*shape, seed=1)
        assert not oracle_lighting(src, False, 1.0, 1.0, 1.0, (255, 255, 255), distant_light(0.0, 45.0)).any()


def test_lighting_flat_surface_is_kd_times_lz():
    # N = (0, 0, 1) everywhere: diffuse = colour * kd * L.z; elevation 90 gives L.z = 1
    src = np.zeros((6, 7, 4), np.uint8)
    src[..., 3] = 77
    out = oracle_lighting(src, False, 3.0, 1.0, 1.0, (250, 200, 120), distant_light(10.0, 90.0))
    assert (out[..., :3] == np.array([250, 200, 120], np.uint8)).all() and (out[..., 3] == 255).all()
    out = oracle_lighting(src, False, 3.0, 0.5, 1.0, (250, 200, 120), distant_light(10.0, 30.0))
    exp = np.floor(np.array([250, 200, 120]) * 0.5 * np.sin(np.radians(30.0)) + 0.5)
    assert (np.abs(out[..., :3].astype(int) - exp.astype(int)) <= 1).all()


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("slope", [12, -12, 30])
@pytest.mark.parametrize("shape", [(3, 3), (5, 8), (8, 5)])
def test_lighting_linear_ramp_has_one_normal_everywhere(axis, slope, shape):
    # alpha = a0 + slope * coordinate: the spec's Sobel normal, border cases included,
    # is (-2 ss slope / 255, 0, 1) (or its y form) at every pixel, so a distant light
    # along +x (+y) lights every pixel the same: colour * kd * N.L / |N|, clamped
    h, w = shape
    yy, xx = np.mgrid[0:h, 0:w]
    c = xx if axis == "x" else yy
    a0 = 10 if slope > 0 else 250
    src = np.zeros((h, w, 4), np.uint8)
    src[..., 3] = np.clip(a0 + slope * c, 0, 255)
    assert (np.diff(src[..., 3].astype(int), axis=1 if axis == "x" else 0) == slope).all()
    ss = 2.0
    n = -2.0 * ss * slope / 255.0
    light = distant_light(0.0 if axis == "x" else 90.0, 0.0)
    lv = np.array([np.cos(np.radians(0.0 if axis == "x" else 90.0)), np.sin(np.radians(0.0 if axis == "x" else 90.0))])
    k = (n * (lv[0] if axis == "x" else lv[1])) / np.sqrt(n * n + 1.0)
    out = oracle_lighting(src, False, ss, 1.0, 1.0, (255, 255, 255), light)
    exp = int(np.floor(min(max(255.0 * k, 0.0), 255.0) + 0.5))
    assert (np.abs(out[..., :3].astype(int) - exp) <= 1).all(), (out[..., 0], exp)
    # the lit side: a rising ramp faces away from a light along the ramp's direction
    assert (exp == 0) == (slope > 0)


# ---------------------------------------------------------------------------
# displacement map (displacement_map.rs:15-62)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(1, 1), (4, 9), (17, 13)])
@pytest.mark.parametrize("chans", [(abi.CHANNEL_R, abi.CHANNEL_A), (abi.CHANNEL_B, abi.CHANNEL_G),
                                   (abi.CHANNEL_A, abi.CHANNEL_A)])
@pytest.mark.parametrize("scale", [0.0, 7.0, -12.5, 300.0])
def test_displacement_equals_spec(shape, chans, scale):
    seed = hash((shape, chans, scale)) & 0xffff
    src, map_, dst = rand_premul(*shape, seed), rand_u8(*shape, seed + 1), rand_premul(*shape, seed + 2)
    for sx, sy in ((1.0, 1.0), (1.5, 0.75)):
        got = dst.copy()
        orc.displacement_map(src, map_, got, sx, sy, scale, *chans)
        assert np.array_equal(got, npr.displacement_map(src, map_, dst, sx, sy, scale, *chans))


@pytest.mark.parametrize("k", [1, 3, -2])
def test_displacement_exact_shift(k):
    # x channel 255 (dx = +1/2), y channel 128 (dy = 1/510): scale 2k moves by exactly k
    # columns and no rows; pixels displaced outside keep the destination
    h, w = 6, 10
    src, dst = rand_premul(h, w, 1), rand_premul(h, w, 2)
    map_ = np.zeros((h, w, 4), np.uint8)
    map_[..., 0] = 255
    map_[..., 1] = 128
    got = dst.copy()
    orc.displacement_map(src, map_, got, 1.0, 1.0, 2.0 * k, abi.CHANNEL_R, abi.CHANNEL_G)
    exp = dst.copy()
    xs = np.arange(w) + k
    ok = (xs >= 0) & (xs < w)
    exp[:, ok] = src[:, xs[ok]]
    assert np.array_equal(got, exp)


# ---------------------------------------------------------------------------
# turbulence lattice (turbulence.rs:93-140)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 42, -7, 1000000, 2147483646, -2147483647])
def test_turbulence_tables_equal_spec_init(seed):
    lat, grad = orc.turbulence_init(seed)
    elat, egrad = npr.turbulence_init(seed)
    assert np.array_equal(lat, elat)
    assert np.array_equal(grad.view(np.uint64), egrad.view(np.uint64))


# ---------------------------------------------------------------------------
# Two-point conical gradients with the focal point off the centre (tiny-skia's
# focal radial gradient): the gradient parameter at a pixel centre p is the largest
# t >= 0 with |p - f - t (c - f)| = t r (circles grow from the focal point f to the
# circle (c, r)), a quadratic in t; where it has no such root the shader is
# transparent.  A black-to-white gradient shows t directly as the colour.
# ---------------------------------------------------------------------------
from resvg_amd.scene import PaintSpec, Scene, rect_path  # noqa: E402


def _focal_t(px, py, f, c, r):
    qx, qy = px - f[0], py - f[1]
    dx, dy = c[0] - f[0], c[1] - f[1]
    a = dx * dx + dy * dy - r * r
    b = qx * dx + qy * dy
    q2 = qx * qx + qy * qy
    with np.errstate(invalid="ignore", divide="ignore"):
        if abs(a) < 1e-9:
            t = q2 / (2.0 * b)
        else:
            disc = b * b - a * q2
            t = (b - np.sqrt(disc)) / a if a < 0 else (b + np.sqrt(disc)) / a
    return t


@pytest.mark.parametrize("where", ["inside", "on", "outside"])
def test_focal_gradient_parameter_is_the_conical_root(where):
    size = 96
    c, r = (50.0, 46.0), 30.0
    d = {"inside": 0.55, "on": 1.0, "outside": 1.6}[where] * r
    f = (c[0] + d * 0.8, c[1] - d * 0.6)
    f = (float(np.float32(f[0])), float(np.float32(f[1])))
    sc = Scene(size, size)
    v, p = rect_path(0.0, 0.0, float(size), float(size))
    sc.path(sc.root, v, p, fill=PaintSpec.radial(c[0], c[1], r, [(0.0, (0, 0, 0), 1.0), (1.0, (255, 255, 255), 1.0)],
                                                 fx=f[0], fy=f[1]))
    px = np.zeros((size, size, 4), np.uint8)
    orc.render(sc.to_abi().abi, abi.Transform.identity(), px)
    ys, xs = np.mgrid[0:size, 0:size].astype(np.float64) + 0.5
    t = np.vectorize(lambda x, y: _focal_t(x, y, f, c, r))(xs, ys)
    defined = np.isfinite(t) & (t > 0)
    exp = np.rint(np.clip(np.where(defined, t, 0.0), 0.0, 1.0) * 255.0)
    got_a = px[..., 3].astype(np.int32)
    # away from the boundary of the undefined region (where f32 and f64 may disagree),
    # defined pixels are opaque with colour = t and undefined ones are transparent
    interior = defined.copy()
    if not defined.all():
        from scipy.ndimage import binary_erosion
        interior = binary_erosion(defined, iterations=2) | binary_erosion(~defined, iterations=2)
    assert np.array_equal(got_a[interior], np.where(defined, 255, 0)[interior])
    diff = np.abs(px[..., 0].astype(np.int32) - exp)[interior & defined]
    assert diff.max() <= 1, (where, diff.max())
    if where == "outside":
        assert (~defined).sum() > 100  # the undefined cone is on the canvas
