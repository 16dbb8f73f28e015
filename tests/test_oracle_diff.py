"""The FP64 oracle of the differentiable form against closed forms (CPU only)."""
import numpy as np

from oracle import diff_ref as D

P = D.PATH_PARAMS


def _params(n, sigma, m=None):
    p = np.zeros(n * P + 21)
    for i in range(n):
        q = p[i * P:(i + 1) * P]
        q[0:3] = 1.0
        q[3] = 1.0
        q[20:24] = [0.1, 0.4, 0.6, 0.9]
    p[n * P] = sigma
    p[n * P + 1:] = (np.eye(4, 5) if m is None else m).reshape(-1)
    return p


def test_blur_weights_normalised_and_small_sigma_is_identity():
    for s in (0.05, 0.7, 3.3, 16.0, 40.0):
        r, w = D.blur_weights(s)
        assert abs(w.sum() - 1.0) < 1e-12 and r <= D.MAX_RADIUS
    r, w = D.blur_weights(0.05)
    assert r == 1 and w[1] > 1.0 - 1e-12


def test_single_solid_path_closed_form():
    # one fully covering path, all stops the same colour: the canvas is that colour
    # premultiplied by alpha·opacity; with a tiny σ and the identity matrix the
    # output is F + (1 - F_a) S with F = S
    h, w = 6, 7
    cov = np.full((1, h, w), 255, np.uint8)
    p = _params(1, 0.05)
    p[4:20] = np.tile([0.2, 0.4, 0.8, 0.5], 4)
    p[3] = 0.6
    out = D.forward(cov, np.array([[0, 0, 7, 0]], np.float32), p, w, h)
    a = 0.5 * 0.6
    s = np.array([0.2 * a, 0.4 * a, 0.8 * a, a])
    np.testing.assert_allclose(out, np.broadcast_to(s + (1 - a) * s, out.shape), atol=1e-12)


def test_gradient_interpolates_between_stops():
    # t = (x + 0.5) / w along the row: at each pixel the colour is the lerp of the
    # two stops around t
    h, w = 1, 10
    cov = np.full((1, h, w), 255, np.uint8)
    p = _params(1, 0.05)
    cols = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [1, 1, 1, 1]], float)
    p[4:20] = cols.reshape(-1)
    s = D._paths(cov, np.array([[0, 0, w, 0]], np.float32), p, w, h)
    off = p[20:24]
    for x in range(w):
        t = (x + 0.5) / w
        if t <= off[0]:
            exp = cols[0]
        elif t >= off[3]:
            exp = cols[3]
        else:
            k = int(np.searchsorted(off, t, side="right")) - 1
            u = (t - off[k]) / (off[k + 1] - off[k])
            exp = cols[k] + u * (cols[k + 1] - cols[k])
        np.testing.assert_allclose(s[0, x], np.concatenate([exp[:3] * exp[3], exp[3:]]), atol=1e-12)


def test_opacity_gradient_matches_closed_form():
    # S = c·op·a·rgb over nothing, then out = S + (1 - S_a)·S (tiny σ, identity
    # matrix); d(sum out)/d(op) in closed form
    h, w = 2, 3
    cov = np.full((1, h, w), 128, np.uint8)
    p = _params(1, 0.05)
    p[4:20] = np.tile([0.3, 0.6, 0.9, 0.8], 4)
    op = 0.7
    p[3] = op
    c = 128 / 255
    ends = np.array([[0, 0, 3, 0]], np.float32)
    g = D.grad_fd(cov, ends, p, w, h, np.ones((h, w, 4)), indices=[3])
    # per pixel: A = c op a ; S = (rgb A, A) ; out = S (2 - A) ; sum = (sum rgb + 1) A (2 - A)
    a, k = 0.8, 0.3 + 0.6 + 0.9 + 1.0
    dA = c * a
    A = c * op * a
    exp = h * w * k * dA * (2.0 - 2.0 * A)
    assert abs(g[3] - exp) < 1e-6 * abs(exp)


def test_matrix_identity_and_clamp():
    b = np.array([[[0.2, 0.1, 0.4, 0.5], [0.0, 0.0, 0.0, 0.0], [0.9, 0.9, 0.9, 0.9]]])
    f = D._color_matrix(b, np.eye(4, 5))
    np.testing.assert_allclose(f, b, atol=1e-12)
    m = np.eye(4, 5)
    m[0, 4] = 2.0  # red saturates at 1
    f = D._color_matrix(b, m)
    np.testing.assert_allclose(f[..., 0], b[..., 3])


def test_matrix_demultiply_floor():
    # alpha at or above 1/255 (every non-zero u8 alpha): exact demultiply, so a matrix
    # that swaps red and blue returns the swapped premultiplied colour
    swap = np.eye(4, 5)[[2, 1, 0, 3]]
    b = np.array([[[0.3, 0.2, 0.1, 0.5], [0.6 / 255, 0.2 / 255, 0.4 / 255, 1.0 / 255]]])
    np.testing.assert_allclose(D._color_matrix(b, swap), b[..., [2, 1, 0, 3]], atol=1e-15)
    # below: colour / (1/255), continuous at the floor and zero with the colour
    a = 1e-6
    b = np.array([[[0.5 * a, 0.25 * a, 0.0, a]]])
    f = D._color_matrix(b, np.eye(4, 5))
    np.testing.assert_allclose(f[0, 0], [0.5 * a * 255 * a, 0.25 * a * 255 * a, 0.0, a], rtol=1e-12)


def test_blur_of_an_impulse_is_the_kernel_outer_product():
    h, w, s = 31, 29, 2.2
    img = np.zeros((h, w, 4))
    img[15, 14] = [1.0, 0.5, 0.25, 1.0]
    out = D._blur(img, s)
    r, k = D.blur_weights(s)
    exp = np.zeros((h, w))
    exp[15 - r:15 + r + 1, 14 - r:14 + r + 1] = np.outer(k, k)
    np.testing.assert_allclose(out[..., 0], exp, atol=1e-15)
    np.testing.assert_allclose(out[..., 1], 0.5 * exp, atol=1e-15)


def _line_cubics(poly):
    """A closed polygon as a chain of straight cubics (control points at thirds)."""
    pts = [poly[0]]
    for i in range(len(poly)):
        a, b = np.array(poly[i], float), np.array(poly[(i + 1) % len(poly)], float)
        pts += [tuple(a + (b - a) / 3), tuple(a + 2 * (b - a) / 3), tuple(b)]
    return np.concatenate([[1, 0, 0, 1, 0, 0], np.array(pts).reshape(-1)]), len(poly)


def test_analytic_coverage_of_a_rectangle():
    g, k = _line_cubics([(2.25, 1.5), (7.75, 1.5), (7.75, 5.5), (2.25, 5.5)])
    cov = D.analytic_coverage(g, k, 10, 8)[0]
    col = np.array([0, 0, 0.75, 1, 1, 1, 1, 0.75, 0, 0])
    row = np.array([0, 0.5, 1, 1, 1, 0.5, 0, 0])
    np.testing.assert_allclose(cov, np.outer(row, col), atol=1e-12)


def test_analytic_coverage_sums_to_the_polygon_area_in_either_orientation():
    poly = [(3.3, 2.1), (17.6, 5.4), (8.2, 14.9), (2.7, 9.3)]
    area = 0.5 * abs(sum(poly[i][0] * poly[(i + 1) % 4][1] - poly[(i + 1) % 4][0] * poly[i][1] for i in range(4)))
    for p in (poly, poly[::-1]):
        g, k = _line_cubics(p)
        cov = D.analytic_coverage(g, k, 21, 18)[0]
        assert abs(cov.sum() - area) < 1e-9
        assert cov.max() <= 1.0 and cov.min() >= 0.0


def test_analytic_coverage_follows_the_affine():
    # a translation by whole pixels shifts the coverage by the same amount
    g, k = _line_cubics([(2.4, 2.2), (6.1, 3.0), (4.0, 6.7)])
    c0 = D.analytic_coverage(g, k, 16, 16)[0]
    g2 = g.copy()
    g2[4:6] = (5.0, 3.0)
    c1 = D.analytic_coverage(g2, k, 16, 16)[0]
    np.testing.assert_allclose(c1[3:, 5:], c0[:-3, :-5], atol=1e-12)


def _winding(px, py, poly):
    """Winding number of the closed polygon about each point (crossing rule)."""
    wn = np.zeros(px.shape)
    for i in range(len(poly)):
        (ax, ay), (bx, by) = poly[i], poly[(i + 1) % len(poly)]
        up = (ay <= py) & (by > py)
        dn = (by <= py) & (ay > py)
        side = (bx - ax) * (py - ay) - (px - ax) * (by - ay)
        wn += np.where(up & (side > 0), 1, 0) - np.where(dn & (side < 0), 1, 0)
    return wn


def test_analytic_coverage_matches_supersampled_winding():
    # random self-intersecting polygons, partly outside the canvas on every side (left
    # of it, their edges are moved onto x = 0, which changes no winding number inside
    # the canvas; right of it, above and below, they are dropped): per pixel, the mean winding number over a 64x64 grid of sample points
    # converges to the integral the oracle computes (with y pointing down, the
    # oracle's signed area is the crossing rule's winding negated)
    rng = np.random.default_rng(5)
    w, h, s = 12, 10, 64
    sub = (np.arange(s) + 0.5) / s
    for _ in range(3):
        poly = [tuple(v) for v in np.column_stack([rng.uniform(-2.0, 14.0, 7), rng.uniform(-1.0, 11.0, 7)])]
        g, k = _line_cubics(poly)
        exact = D.winding_area(D.flatten(g, k), w, h)
        for y in range(h):
            for x in range(w):
                px, py = np.meshgrid(x + sub, y + sub)
                assert abs(-_winding(px, py, poly).mean() - exact[y, x]) < 5e-3, (x, y)
