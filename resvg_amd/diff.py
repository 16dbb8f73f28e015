"""The differentiable f32 form: rvk_diff_forward / rvk_diff_backward wrapped in a
torch.autograd.Function.

A scene is a list of closed paths (verbs, points) with the end points of each
path's linear gradient.  Two coverage modes:

* planes (default): the coverage is computed once by the u8 pipeline (each path
  drawn alone with an opaque white fill through rvk_render_batch, the alpha channel
  kept); gradients flow to the paint, opacity, blur σ and colour matrix;
* analytic: every path is a closed chain of k cubics (M, k x C, Z) whose control
  points and affine transform are parameters too (geometry tensor, GEOM_FLOATS(k)
  floats per path: sx, ky, kx, sy, tx, ty, then P0 .. P3k); the coverage is the
  winding area of the flattened path per pixel and gradients also flow to the
  geometry.

Parameter layout (float32):

    per path (PATH_PARAMS = 24): tint r, g, b; opacity; 4 stops x (r, g, b, a);
                                 4 stop offsets (ascending)
    then: sigma, then the 4x5 colour matrix (row-major)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi, lib, _check, _ptr
from .scene import PaintSpec, Scene

PATH_PARAMS = 24


def geom_floats(k: int) -> int:
    return 6 + 2 * (3 * k + 1)


class DiffScene(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("path_count", C.c_uint32), ("cubics", C.c_uint32),
                ("dev_coverage", C.c_void_p), ("dev_gradient_ends", C.c_void_p)]


def _bind():
    L = lib()
    P, u32 = C.c_void_p, C.c_uint32
    L.rvk_diff_workspace_size.argtypes = [u32, u32, u32, u32, C.POINTER(C.c_size_t)]
    L.rvk_diff_workspace_size.restype = C.c_int32
    L.rvk_diff_forward.argtypes = [P, C.POINTER(DiffScene), P, P, P, P]
    L.rvk_diff_forward.restype = C.c_int32
    L.rvk_diff_backward.argtypes = [P, C.POINTER(DiffScene), P, P, P, P, P, P]
    L.rvk_diff_backward.restype = C.c_int32
    L.rvk_diff_forward_batch.argtypes = [P, C.POINTER(DiffScene), u32, P, P, P, P]
    L.rvk_diff_forward_batch.restype = C.c_int32
    L.rvk_diff_backward_batch.argtypes = [P, C.POINTER(DiffScene), u32, P, P, P, P, P, P]
    L.rvk_diff_backward_batch.restype = C.c_int32
    return L


def n_params(n_paths: int) -> int:
    return n_paths * PATH_PARAMS + 1 + 20


# A path entry of a scene is (verbs, points) or (verbs, points, style): style
# {"fill_rule": ...} fills with that rule, style {"stroke_width": ..., "join": ...,
# "cap": ..., "miter": ..., "dasharray": ..., "dashoffset": ...} makes the entry the
# path's stroke (its coverage is the stroke outline's, tiny-skia's stroker and
# dashing).  The coverage-plane mode takes every kind; the analytic mode takes
# nonzero fills of closed cubic chains.
_STROKE_KEYS = {"stroke_width": "stroke_width", "join": "stroke_join", "cap": "stroke_cap", "miter": "miter_limit",
                "dasharray": "dasharray", "dashoffset": "dashoffset"}


def _entry(path):
    verbs, points = path[0], path[1]
    style = dict(path[2]) if len(path) > 2 and path[2] else {}
    return verbs, points, style


def _draw(sc: Scene, parent: int, path, paint: PaintSpec) -> None:
    """Adds one path entry to `sc` under `parent`, painted with `paint` (as its fill,
    or as its stroke for a stroke entry)."""
    verbs, points, style = _entry(path)
    if "stroke_width" in style:
        kw = {_STROKE_KEYS[k]: v for k, v in style.items() if k in _STROKE_KEYS}
        sc.path(parent, verbs, points, stroke=paint, **kw)
    else:
        sc.path(parent, verbs, points, fill=paint, fill_rule=style.get("fill_rule", abi.FILL_NONZERO))


def one_path_scene(size, verbs, points, style=None) -> Scene:
    """A document holding one path entry drawn opaque white (its alpha = coverage)."""
    w, h = size
    sc = Scene(w, h)
    _draw(sc, sc.root, (verbs, points, style), PaintSpec.solid((255, 255, 255)))
    return sc


def geometry(paths, k: int):
    """The analytic mode's geometry parameters of closed k-cubic paths (identity
    transforms), float32 (n, GEOM_FLOATS(k))."""
    out = np.zeros((len(paths), geom_floats(k)), np.float32)
    for i, path in enumerate(paths):
        verbs, pts, style = _entry(path)
        assert "stroke_width" not in style and style.get("fill_rule", abi.FILL_NONZERO) == abi.FILL_NONZERO, \
            "analytic mode takes nonzero fills"
        v = list(verbs)
        assert v[0] == abi.VERB_MOVE and v[1:1 + k] == [abi.VERB_CUBIC] * k and len(v) in (k + 1, k + 2), \
            "analytic mode takes closed chains of k cubics"
        out[i, :6] = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        out[i, 6:] = np.asarray(pts, np.float32).reshape(-1)[:2 * (3 * k + 1)]
    return out


class Renderer:
    """The coverage source, gradient ends and workspace of one scene, or of a batch
    of scenes of one shape (`Renderer.batched`), on one context.  `cubics` > 0
    selects the analytic mode (paths of that many cubics)."""

    def __init__(self, ctx, width: int, height: int, paths, ends, cubics: int = 0):
        self._setup(ctx, width, height, [(paths, ends)], cubics, single=True)

    @classmethod
    def batched(cls, ctx, width: int, height: int, scenes, cubics: int = 0):
        """`scenes`: a list of (paths, ends), every scene with the same number of
        paths.  Parameters are then (batch, n_params), geometry (batch, paths,
        GEOM_FLOATS) and images (batch, height, width, 4); every kernel runs once for
        the whole batch."""
        r = cls.__new__(cls)
        r._setup(ctx, width, height, list(scenes), cubics, single=False)
        return r

    def _setup(self, ctx, width, height, scenes, cubics, single):
        import torch
        self.ctx, self.w, self.h, self.k = ctx, int(width), int(height), int(cubics)
        self.batch, self.single = len(scenes), single
        self.n = len(scenes[0][0])
        assert self.batch >= 1 and all(len(p) == self.n for p, _ in scenes), "scenes of one path count"
        self._L = _bind()
        dev = torch.device("cuda", ctx.device)
        if self.k == 0:
            # coverage: every path rendered alone, opaque white, through one batch call
            pxs = torch.zeros((self.batch * self.n, self.h, self.w, 4), dtype=torch.uint8, device=dev)
            docs = [one_path_scene((self.w, self.h), *_entry(path)) for paths, _ in scenes for path in paths]
            ctx.render_batch(docs, abi.Transform.identity(), [pxs[i] for i in range(len(docs))])
            self.coverage = pxs[..., 3].contiguous()
            del pxs
            self.geom0 = None
        else:
            self.coverage = None
            g = np.stack([geometry(paths, self.k) for paths, _ in scenes])
            self.geom0 = torch.as_tensor(g[0] if single else g, device=dev)
        e = np.stack([np.asarray(ends, np.float32).reshape(self.n, 4) for _, ends in scenes])
        self.ends = torch.as_tensor(e, device=dev).contiguous()
        nbytes = C.c_size_t(0)
        _check(self._L.rvk_diff_workspace_size(self.w, self.h, self.n, self.k, C.byref(nbytes)),
               "rvk_diff_workspace_size")
        self.workspace_bytes = int(nbytes.value)
        self.workspace = torch.empty(self.batch * self.workspace_bytes, dtype=torch.uint8, device=dev)
        # counts forward calls: the workspace holds the intermediates of the latest one
        self.generation = 0
        cov = self.coverage.data_ptr() if self.coverage is not None else None
        self.scene = DiffScene(self.w, self.h, self.n, self.k, cov, self.ends.data_ptr())

    def _image_shape(self):
        return (self.h, self.w, 4) if self.single else (self.batch, self.h, self.w, 4)

    def forward(self, params, geom=None):
        import torch
        assert params.dtype == torch.float32 and params.is_contiguous()
        assert params.numel() == self.batch * n_params(self.n)
        if self.k:
            assert geom is not None and geom.dtype == torch.float32 and geom.is_contiguous()
            assert geom.numel() == self.batch * self.n * geom_floats(self.k)
        out = torch.empty(self._image_shape(), dtype=torch.float32, device=params.device)
        _check(self._L.rvk_diff_forward_batch(self.ctx.h, C.byref(self.scene), self.batch, _ptr(params),
                                              _ptr(geom) if self.k else None, _ptr(self.workspace), _ptr(out)),
               "rvk_diff_forward_batch")
        self.generation += 1
        return out

    def backward(self, params, grad_out, geom=None):
        """dL/d(params), and dL/d(geom) in the analytic mode (else None).  Reads the
        intermediates the latest forward left in the workspace, so it must follow the
        forward of the same parameters and geometry (the autograd wrapper from
        render_fn re-runs that forward when another one came in between)."""
        import torch
        grad_out = grad_out.contiguous().float()
        assert grad_out.numel() == self.batch * self.w * self.h * 4
        g = torch.empty_like(params)
        gg = torch.empty_like(geom) if self.k else None
        _check(self._L.rvk_diff_backward_batch(self.ctx.h, C.byref(self.scene), self.batch, _ptr(params),
                                               _ptr(geom) if self.k else None, _ptr(self.workspace),
                                               _ptr(grad_out), _ptr(g), _ptr(gg) if self.k else None),
               "rvk_diff_backward_batch")
        return (g, gg) if self.k else g


def render_fn():
    """A torch.autograd.Function: render(params, geom, renderer) -> (H, W, 4)
    float32; geom is None in the planes mode."""
    import torch

    class DiffRender(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params, geom, renderer):
            params = params.contiguous()
            geom = geom.contiguous() if geom is not None else None
            out = renderer.forward(params, geom)
            ctx.save_for_backward(params, geom) if geom is not None else ctx.save_for_backward(params)
            ctx.renderer = renderer
            ctx.generation = renderer.generation
            return out

        @staticmethod
        def backward(ctx, grad_out):
            saved = ctx.saved_tensors
            params, geom = saved[0], (saved[1] if len(saved) == 2 else None)
            r = ctx.renderer
            # the shared workspace holds the intermediates of the renderer's latest
            # forward; when a later forward replaced them, recompute this one's (the
            # forward is deterministic, so they are the same bits)
            if r.generation != ctx.generation:
                r.forward(params, geom)
            if geom is not None:
                gp, gg = r.backward(params, grad_out, geom)
                return gp, gg, None
            return r.backward(params, grad_out), None, None

    return DiffRender.apply


def u8_scene(width, height, paths, ends, params) -> Scene:
    """The same document as an ordinary u8 scene (rvk_render's input): each path in
    an opacity group with its gradient (stops premultiplied by nothing, colours
    quantised to u8), all inside a group filtered by blur -> colour matrix ->
    composite over the source, in sRGB (color-interpolation-filters="sRGB": the f32
    model has no colour-space conversion).  σ is clamped to the model's [0.05, 16]
    and the offsets made non-decreasing, as usvg does for stops.  Used to time the
    plain forward of the same work and to score fitted parameters."""
    from .filters import SOURCE_GRAPHIC, FilterGraph
    p = np.asarray(params, np.float64)
    n = len(paths)
    sc = Scene(width, height)
    g = FilterGraph((0.0, 0.0, float(width), float(height)))
    sigma = float(np.clip(p[n * PATH_PARAMS], 0.05, 16.0))
    cs = abi.CI_SRGB
    b = g.gaussian_blur(sigma, cs=cs)
    cm = g.color_matrix(abi.COLOR_MATRIX_MATRIX, p[n * PATH_PARAMS + 1:].astype(np.float32).tolist(), in1=b, cs=cs)
    g.composite(abi.COMPOSITE_OVER, in1=cm, in2=SOURCE_GRAPHIC, cs=cs)
    root = sc.group(sc.root, filters=[g])
    for i, path in enumerate(paths):
        q = p[i * PATH_PARAMS:(i + 1) * PATH_PARAMS]
        tint, opacity, cols = q[0:3], float(q[3]), q[4:20].reshape(4, 4)
        offs = np.maximum.accumulate(np.clip(q[20:24], 0.0, 1.0))
        stops = []
        for k in range(4):
            rgb = tuple(int(np.clip(round(cols[k, c] * tint[c] * 255.0), 0, 255)) for c in range(3))
            stops.append((float(offs[k]), rgb, float(np.clip(cols[k, 3], 0.0, 1.0))))
        x0, y0, x1, y1 = ends[i]
        grp = sc.group(root, opacity=float(np.clip(opacity, 0.0, 1.0)))
        _draw(sc, grp, path, PaintSpec.linear(float(x0), float(y0), float(x1), float(y1), stops))
    return sc
