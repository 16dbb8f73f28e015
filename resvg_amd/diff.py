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
    return L


def n_params(n_paths: int) -> int:
    return n_paths * PATH_PARAMS + 1 + 20


def one_path_scene(size, verbs, points) -> Scene:
    """A document holding one path with an opaque white fill (its alpha = coverage)."""
    w, h = size
    sc = Scene(w, h)
    sc.path(sc.root, verbs, points, fill=PaintSpec.solid((255, 255, 255)), fill_rule=abi.FILL_NONZERO)
    return sc


def geometry(paths, k: int):
    """The analytic mode's geometry parameters of closed k-cubic paths (identity
    transforms), float32 (n, GEOM_FLOATS(k))."""
    out = np.zeros((len(paths), geom_floats(k)), np.float32)
    for i, (verbs, pts) in enumerate(paths):
        v = list(verbs)
        assert v[0] == abi.VERB_MOVE and v[1:1 + k] == [abi.VERB_CUBIC] * k and len(v) in (k + 1, k + 2), \
            "analytic mode takes closed chains of k cubics"
        out[i, :6] = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        out[i, 6:] = np.asarray(pts, np.float32).reshape(-1)[:2 * (3 * k + 1)]
    return out


class Renderer:
    """The coverage source, gradient ends and workspace of one scene on one context.
    `cubics` > 0 selects the analytic mode (paths of that many cubics)."""

    def __init__(self, ctx, width: int, height: int, paths, ends, cubics: int = 0):
        import torch
        self.ctx, self.w, self.h, self.n, self.k = ctx, int(width), int(height), len(paths), int(cubics)
        self._L = _bind()
        dev = torch.device("cuda", ctx.device)
        if self.k == 0:
            # coverage: every path rendered alone, opaque white, through one batch call
            pxs = torch.zeros((self.n, self.h, self.w, 4), dtype=torch.uint8, device=dev)
            scenes = [one_path_scene((self.w, self.h), v, p) for v, p in paths]
            ctx.render_batch(scenes, abi.Transform.identity(), [pxs[i] for i in range(self.n)])
            self.coverage = pxs[..., 3].contiguous()
            del pxs
            self.geom0 = None
        else:
            self.coverage = None
            self.geom0 = torch.as_tensor(geometry(paths, self.k), device=dev)
        self.ends = torch.as_tensor(np.asarray(ends, np.float32).reshape(self.n, 4), device=dev).contiguous()
        nbytes = C.c_size_t(0)
        _check(self._L.rvk_diff_workspace_size(self.w, self.h, self.n, self.k, C.byref(nbytes)),
               "rvk_diff_workspace_size")
        self.workspace = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
        cov = self.coverage.data_ptr() if self.coverage is not None else None
        self.scene = DiffScene(self.w, self.h, self.n, self.k, cov, self.ends.data_ptr())

    def forward(self, params, geom=None):
        import torch
        assert params.dtype == torch.float32 and params.is_contiguous() and params.numel() == n_params(self.n)
        if self.k:
            assert geom is not None and geom.is_contiguous() and geom.numel() == self.n * geom_floats(self.k)
        out = torch.empty((self.h, self.w, 4), dtype=torch.float32, device=params.device)
        _check(self._L.rvk_diff_forward(self.ctx._h, C.byref(self.scene), _ptr(params),
                                        _ptr(geom) if self.k else None, _ptr(self.workspace), _ptr(out)),
               "rvk_diff_forward")
        return out

    def backward(self, params, grad_out, geom=None):
        """dL/d(params), and dL/d(geom) in the analytic mode (else None)."""
        import torch
        grad_out = grad_out.contiguous().float()
        g = torch.empty_like(params)
        gg = torch.empty_like(geom) if self.k else None
        _check(self._L.rvk_diff_backward(self.ctx._h, C.byref(self.scene), _ptr(params),
                                         _ptr(geom) if self.k else None, _ptr(self.workspace), _ptr(grad_out),
                                         _ptr(g), _ptr(gg) if self.k else None), "rvk_diff_backward")
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
            return out

        @staticmethod
        def backward(ctx, grad_out):
            saved = ctx.saved_tensors
            # the workspace holds the forward intermediates of these parameters
            if len(saved) == 2:
                gp, gg = ctx.renderer.backward(saved[0], grad_out, saved[1])
                return gp, gg, None
            return ctx.renderer.backward(saved[0], grad_out), None, None

    return DiffRender.apply


def u8_scene(width, height, paths, ends, params) -> Scene:
    """The same document as an ordinary u8 scene (rvk_render's input): each path in
    an opacity group with its gradient (stops premultiplied by nothing, colours
    quantised to u8), all inside a group filtered by blur -> colour matrix ->
    composite over the source.  Used to time the plain forward of the same work."""
    from .filters import SOURCE_GRAPHIC, FilterGraph
    p = np.asarray(params, np.float64)
    n = len(paths)
    sc = Scene(width, height)
    g = FilterGraph((0.0, 0.0, float(width), float(height)))
    sigma = float(p[n * PATH_PARAMS])
    b = g.gaussian_blur(sigma)
    cm = g.color_matrix(abi.COLOR_MATRIX_MATRIX, p[n * PATH_PARAMS + 1:].astype(np.float32).tolist(), in1=b)
    g.composite(abi.COMPOSITE_OVER, in1=cm, in2=SOURCE_GRAPHIC)
    root = sc.group(sc.root, filters=[g])
    for i, (verbs, pts) in enumerate(paths):
        q = p[i * PATH_PARAMS:(i + 1) * PATH_PARAMS]
        tint, opacity, cols, offs = q[0:3], float(q[3]), q[4:20].reshape(4, 4), q[20:24]
        stops = []
        for k in range(4):
            rgb = tuple(int(np.clip(round(cols[k, c] * tint[c] * 255.0), 0, 255)) for c in range(3))
            stops.append((float(np.clip(offs[k], 0.0, 1.0)), rgb, float(np.clip(cols[k, 3], 0.0, 1.0))))
        x0, y0, x1, y1 = ends[i]
        grp = sc.group(root, opacity=float(np.clip(opacity, 0.0, 1.0)))
        sc.path(grp, verbs, pts, fill=PaintSpec.linear(float(x0), float(y0), float(x1), float(y1), stops))
    return sc
