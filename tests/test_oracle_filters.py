"""Pins the oracle's filter half (CPU only).

* the reference's own fixture PNGs, end to end (fixture_scenes.py);
* the colour-space LUTs against the sRGB transfer formula;
* independent numpy formulations (numpy_ref.py) on random, ragged and tiny inputs;
* textbook values (Park–Miller), closed forms (box radii) and invariants.
"""
import numpy as np
import pytest

from oracle import bindings as orc
from resvg_amd import abi
from resvg_amd.filters import FilterGraph

import numpy_ref as npr
from fixture_scenes import compare_png, fixtures, load_png, render_filtered_rect_oracle

FIXTURES = fixtures()


def rand_premul(h, w, seed):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    a = px[..., 3:4].astype(np.uint16)
    px[..., :3] = (px[..., :3].astype(np.uint16) * a // 255).astype(np.uint8)
    return px


@pytest.mark.parametrize("fx", FIXTURES, ids=[f.name for f in FIXTURES])
def test_reference_fixture_png(fx):
    out = render_filtered_rect_oracle(fx)
    assert compare_png(out, load_png(fx.name), guides=fx.guides) == 0


def test_colorspace_tables_match_srgb_formula():
    s2l, l2s = orc.tables()
    c = np.arange(256) / 255.0
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    srgb = np.where(c <= 0.0031308, c * 12.92, 1.055 * c ** (1 / 2.4) - 0.055)
    assert np.array_equal(np.round(lin * 255).astype(int), s2l)
    assert np.array_equal(np.round(srgb * 255).astype(int), l2s)


def test_premultiply_of_opaque_and_transparent():
    px = rand_premul(7, 9, 1)
    px[..., 3] = 255
    ref = px.copy()
    orc.demultiply_alpha(px)
    assert np.array_equal(px, ref)
    orc.multiply_alpha(px)
    assert np.array_equal(px, ref)
    z = np.zeros((3, 3, 4), np.uint8)
    orc.demultiply_alpha(z)  # 0/0 = NaN -> 0
    assert not z.any()


@pytest.mark.parametrize("sigma,sizes", [(4.0, [7] * 5), (6.0, [9, 9, 9, 9, 11]), (24.0, [37] * 5),
                                         (2.0, [3] * 5), (0.0, [1] * 5), (-1.0, [1] * 5)])
def test_box_gauss_closed_form(sigma, sizes):
    assert orc.box_gauss_sizes(sigma) == sizes
    assert npr.box_sizes(sigma) == sizes


@pytest.mark.parametrize("shape", [(1, 1), (3, 50), (50, 3), (17, 9), (40, 33), (2, 7)])
@pytest.mark.parametrize("sig", [(2.0, 2.0), (4.0, 4.0), (6.0, 0.5), (0.0, 5.0), (24.0, 3.0)])
def test_box_blur_equals_window_sums(shape, sig):
    px = rand_premul(*shape, seed=hash((shape, sig)) & 0xffff)
    exp = npr.box_blur(px, *sig)
    orc.box_blur(px, *sig)
    assert np.array_equal(px, exp)


@pytest.mark.parametrize("shape", [(1, 1), (5, 1), (1, 6), (13, 11), (24, 31)])
@pytest.mark.parametrize("sig", [(0.5, 0.5), (1.5, 1.0), (1.9, 0.0), (0.0, 0.7)])
def test_iir_equals_vectorised_recurrence(shape, sig):
    px = rand_premul(*shape, seed=7)
    exp = npr.iir_blur(px, *sig)
    orc.iir_blur(px, *sig)
    assert np.array_equal(px, exp)


def test_iir_constant_image_stays_near_constant():
    px = np.full((32, 32, 4), 200, np.uint8)
    orc.iir_blur(px, 1.5, 1.5)
    # post_scale normalises the cascade's DC gain to 1: away from the borders a flat
    # field stays within 1 LSB of itself (truncation), while the borders darken
    inner = px[8:24, 8:24].astype(int)
    assert np.abs(inner - 200).max() <= 1
    assert px[0, 16, 0] < 150


@pytest.mark.parametrize("shape", [(1, 1), (9, 4), (20, 17)])
@pytest.mark.parametrize("r", [(1.0, 1.0), (2.5, 0.5), (8.0, 3.0), (20.0, 20.0)])
@pytest.mark.parametrize("erode", [True, False])
def test_morphology_bruteforce_and_order(shape, r, erode):
    px = rand_premul(*shape, seed=3)
    exp = npr.morphology(px, erode, *r)
    out = px.copy()
    orc.morphology(out, abi.MORPH_ERODE if erode else abi.MORPH_DILATE, *r)
    assert np.array_equal(out, exp)
    assert (out <= px).all() if erode else (out >= px).all()


def test_arithmetic_composite_numpy():
    s1, s2 = rand_premul(19, 23, 11), rand_premul(19, 23, 12)
    for k in [(0.1, 0.2, 0.3, 0.4), (0.0, 1.0, 0.0, 0.0), (1.5, -0.5, 0.7, -0.1), (0, 0, 0, 0)]:
        dst = rand_premul(19, 23, 13)
        exp = npr.composite_arithmetic(s1, s2, dst, *k)
        orc.composite_arithmetic(s1, s2, dst, *k)
        assert np.array_equal(dst, exp)


def test_arithmetic_k2_is_identity_on_premultiplied():
    s1, s2 = rand_premul(8, 8, 5), rand_premul(8, 8, 6)
    dst = np.zeros_like(s1)
    orc.composite_arithmetic(s1, s2, dst, 0.0, 1.0, 0.0, 0.0)
    keep = s1[..., 3] > 0
    assert np.abs(dst[keep].astype(int) - s1[keep].astype(int)).max() <= 1


def test_park_miller_textbook_value():
    # Park & Miller (1988): 10000 steps of the minimal standard generator from 1
    x = 1
    for _ in range(10000):
        x = orc.turbulence_random(x)
    assert x == 1043618065
    assert npr.park_miller(1, 10000) == 1043618065


def test_turbulence_tables_are_unit_gradients_and_a_permutation():
    lat, grad = orc.turbulence_init(0)
    assert sorted(lat[:256].tolist()) == list(range(256))
    assert np.array_equal(lat[256:512], lat[:256]) and np.array_equal(lat[512:], lat[:2])
    norms = np.hypot(grad[..., 0], grad[..., 1])
    assert np.allclose(norms[:, :256], 1.0, atol=1e-12)


def test_draw_pixmap_onto_transparent_is_identity():
    src = rand_premul(31, 17, 9)
    dst = np.zeros_like(src)
    orc.draw_pixmap(dst, 0, 0, src)
    assert np.array_equal(dst, src)


def test_filter_graph_error_clears_layer():
    g = FilterGraph((0.0, 0.0, 10.0, 10.0))
    g.gaussian_blur(3.0)
    px = rand_premul(8, 8, 2)
    # a filter region entirely outside the layer is an InvalidRegion error
    g.rect = (100.0, 100.0, 5.0, 5.0)
    err = orc.filter_apply(g, abi.Transform.identity(), px)
    assert err == abi.ERR_INVALID_REGION
    assert not px.any()
