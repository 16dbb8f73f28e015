"""Host logic of the differentiable form's harness (CPU only): the geometry ->
u8 document conversion used to score fitted parameters."""
import numpy as np

from resvg_amd import abi, diff, fit
from resvg_amd.workloads import diff_scene


def _bezier(p, t):
    u = 1.0 - t
    return u ** 3 * p[0] + 3 * u * u * t * p[1] + 3 * u * t * t * p[2] + t ** 3 * p[3]


def test_paths_from_geometry_maps_the_curves():
    # an affine image of a cubic's control points is the cubic's image: points on the
    # device-space path equal the affine applied to points on the user-space path
    k = 3
    _, _, paths, _, _, _ = diff_scene(seed=2, size=64, n_paths=2, segments=k, sigma=1.0)
    g = diff.geometry(paths, k).astype(np.float64)
    g[:, :6] = [[0.9, 0.2, -0.3, 1.1, 5.0, -2.0], [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]]
    out = fit.paths_from_geometry(g, k)
    assert len(out) == 2
    for (verbs, dev), row in zip(out, g):
        assert list(verbs) == [abi.VERB_MOVE] + [abi.VERB_CUBIC] * k + [abi.VERB_CLOSE]
        user = row[6:].reshape(-1, 2)
        sx, ky, kx, sy, tx, ty = row[:6]
        for c in range(k):
            for t in (0.0, 0.3, 0.77, 1.0):
                pu = _bezier(user[3 * c:3 * c + 4], t)
                pd = _bezier(dev[3 * c:3 * c + 4].astype(np.float64), t)
                np.testing.assert_allclose(pd, [sx * pu[0] + kx * pu[1] + tx, ky * pu[0] + sy * pu[1] + ty],
                                           atol=1e-4)


def test_u8_scene_clamps_like_the_model():
    # the u8 document of a parameter vector: stop offsets made non-decreasing and
    # clipped to [0, 1] (usvg's stop processing), σ clipped to the model's [0.05, 16],
    # the filter chain in sRGB (the f32 model has no colour-space conversion)
    w = h = 32
    _, _, paths, ends, p, _ = diff_scene(seed=1, size=w, n_paths=1, segments=2, sigma=40.0)
    p = p.astype(np.float64)
    p[20:24] = [0.6, 0.2, 1.4, -0.5]
    sc = diff.u8_scene(w, h, paths, ends, p)
    offs = [o for o, _, _ in sc.paths[0]["fill"].stops]
    assert offs == sorted(offs) and min(offs) >= 0.0 and max(offs) <= 1.0
    assert offs == [0.6, 0.6, 1.0, 1.0]
    _, off, count = sc.filters[0]  # blur -> colour matrix -> composite
    prims = sc.primitives[off:off + count]
    assert [pr.kind for pr in prims] == [abi.FE_GAUSSIAN_BLUR, abi.FE_COLOR_MATRIX, abi.FE_COMPOSITE]
    assert abs(prims[0].u.blur.std_dev_x - 16.0) < 1e-6
    assert all(pr.color_interpolation == abi.CI_SRGB for pr in prims)
