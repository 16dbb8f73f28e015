"""resvg_amd — the resvg pixel pipeline on MI355X (gfx950).

Python bindings of librvk.so (include/resvg_amd.h).  The library is required: a
missing or unloadable librvk.so raises at import of the first entry point, there is
no CPU fallback.  Pixel buffers are torch uint8 CUDA tensors of shape (H, W, 4)
holding premultiplied RGBA8 (the tiny_skia::Pixmap layout).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from . import abi
from .abi import Transform  # noqa: F401

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "librvk.so"
_lib = None


class RvkError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: rvk status {status}")


def lib() -> C.CDLL:
    """Loads librvk.so; raises when it is missing (build with resvg_amd.build)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m resvg_amd.build` (no CPU fallback exists)")
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def _declare(L: C.CDLL) -> None:
    P = C.c_void_p
    u32, i32, f32, f64 = C.c_uint32, C.c_int32, C.c_float, C.c_double
    sigs = {
        "rvk_status_string": ([i32], C.c_char_p),
        "rvk_context_create": ([C.c_int, C.POINTER(P)], i32),
        "rvk_context_destroy": ([P], i32),
        "rvk_context_set_stream": ([P, P], i32),
        "rvk_context_stream": ([P], P),
        "rvk_synchronize": ([P], i32),
        "rvk_device_alloc": ([P, C.c_size_t, C.POINTER(P)], i32),
        "rvk_device_free": ([P, P], i32),
        "rvk_upload": ([P, P, P, C.c_size_t], i32),
        "rvk_download": ([P, P, P, C.c_size_t], i32),
        "rvk_into_linear_rgb": ([P, P, u32, u32], i32),
        "rvk_into_srgb": ([P, P, u32, u32], i32),
        "rvk_demultiply_alpha": ([P, P, u32, u32], i32),
        "rvk_multiply_alpha": ([P, P, u32, u32], i32),
        "rvk_box_blur": ([P, P, u32, u32, f64, f64], i32),
        "rvk_box_blur_noalloc": ([P, P, u32, u32, f64, f64], i32),
        "rvk_iir_blur": ([P, P, u32, u32, f64, f64], i32),
        "rvk_morphology": ([P, P, u32, u32, u32, f32, f32], i32),
        "rvk_convolve_matrix": ([P, P, u32, u32, C.POINTER(abi.Primitive), C.POINTER(f32)], i32),
        "rvk_color_matrix": ([P, P, u32, u32, C.POINTER(abi.Primitive)], i32),
        "rvk_component_transfer": ([P, P, u32, u32, C.POINTER(abi.Primitive), C.POINTER(f32)], i32),
        "rvk_composite_arithmetic": ([P, P, P, P, u32, u32, f32, f32, f32, f32], i32),
        "rvk_turbulence": ([P, P, u32, u32, f64, f64, f64, f64, f64, f64, u32, i32, u32, u32], i32),
        "rvk_lighting": ([P, P, P, u32, u32, u32, f32, f32, f32, C.POINTER(C.c_uint8), C.POINTER(abi.LightSource)],
                         i32),
        "rvk_displacement_map": ([P, P, P, P, u32, u32, f32, f32, f32, u32, u32], i32),
        "rvk_draw_pixmap": ([P, P, u32, u32, i32, i32, P, u32, u32, f32, u32], i32),
        "rvk_fill": ([P, P, u32, u32, C.POINTER(C.c_uint8)], i32),
        "rvk_filter_apply": ([P, C.POINTER(abi.Filter), C.POINTER(abi.FilterData), abi.Transform, P, u32, u32,
                              C.POINTER(i32)], i32),
        "rvk_png_filter_rows": ([P, P, u32, u32, P], i32),
        "rvk_get_stats": ([P, C.POINTER(abi.Stats)], i32),
        "rvk_profile_begin": ([P], i32),
        "rvk_profile_end": ([P, C.POINTER(abi.KernelStat), C.c_uint32, C.POINTER(C.c_uint32)], i32),
    }
    optional = {
        "rvk_render": ([P, C.POINTER(abi.Scene), abi.Transform, u32, u32, P], i32),
        "rvk_render_host": ([P, C.POINTER(abi.Scene), abi.Transform, u32, u32, P], i32),
        "rvk_render_batch": ([P, u32, C.POINTER(C.POINTER(abi.Scene)), C.POINTER(abi.Transform), C.POINTER(u32),
                              C.POINTER(u32), C.POINTER(C.c_void_p), C.POINTER(i32)], i32),
        "rvk_render_node": ([P, C.POINTER(abi.Scene), abi.NodeRef, C.POINTER(abi.Rect), abi.Transform, u32, u32, P,
                             C.POINTER(i32)], i32),
        "rvk_fill_path": ([P, P, u32, u32, C.POINTER(abi.PathGeometry), u32, C.POINTER(abi.Paint),
                           C.POINTER(abi.Stop), abi.Transform, u32, u32], i32),
        "rvk_stroke_path": ([P, P, u32, u32, C.POINTER(abi.PathGeometry), C.POINTER(abi.StrokeStyle),
                             C.POINTER(abi.Paint), C.POINTER(abi.Stop), abi.Transform, u32], i32),
        "rvk_mask_fill_path": ([P, P, u32, u32, C.POINTER(abi.PathGeometry), u32, u32, abi.Transform], i32),
        "rvk_mask_from_pixmap": ([P, P, u32, u32, u32, P], i32),
        "rvk_mask_invert": ([P, P, u32, u32], i32),
        "rvk_apply_mask": ([P, P, u32, u32, P], i32),
        "rvk_clear_rect": ([P, P, u32, u32, i32, i32, u32, u32], i32),
    }
    for name, (args, res) in {**sigs, **optional}.items():
        fn = getattr(L, name, None)
        if fn is None:
            if name in optional:
                continue
            raise RuntimeError(f"librvk.so does not export {name}")
        fn.argtypes = args
        fn.restype = res


def exported_symbols() -> list[str]:
    """Names declared in include/resvg_amd.h that the loaded library exports."""
    L = lib()
    hdr = (_PKG.parent / "include" / "resvg_amd.h").read_text()
    import re
    names = sorted(set(re.findall(r"\b(rvk_[a-z0-9_]+)\s*\(", hdr)))
    return [n for n in names if hasattr(L, n)]


def _check(status: int, what: str) -> None:
    if status != abi.OK:
        raise RvkError(status, what)


def _ptr(t) -> int:
    """Device pointer of a torch tensor (or a raw int pointer)."""
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _dims(px) -> tuple[int, int]:
    h, w, c = px.shape
    assert c == 4 and px.dtype.itemsize == 1 and px.is_contiguous(), "expected a contiguous (H, W, 4) uint8 tensor"
    return w, h


class Context:
    """An rvk_context on one HIP device.  With `use_torch_stream` every call is
    ordered on torch's current stream at the time of the call (the stream is read
    again on each entry point, so work issued under `with torch.cuda.stream(s)` runs
    on `s`); otherwise on the context's own stream."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        L = lib()
        h = C.c_void_p()
        _check(L.rvk_context_create(device, C.byref(h)), "rvk_context_create")
        self._h = h
        self.device = device
        self._torch_stream = use_torch_stream
        self._bound_stream = None
        self._sync_stream()

    def _sync_stream(self) -> None:
        """Binds the context to torch's current stream when that has changed."""
        if not self._torch_stream:
            return
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._bound_stream:
            _check(lib().rvk_context_set_stream(self._h, C.c_void_p(s)), "rvk_context_set_stream")
            self._bound_stream = s

    @property
    def h(self):
        """The handle, bound to torch's current stream (used by every entry point)."""
        self._sync_stream()
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().rvk_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def synchronize(self) -> None:
        _check(lib().rvk_synchronize(self.h), "rvk_synchronize")

    def stats(self) -> abi.Stats:
        s = abi.Stats()
        _check(lib().rvk_get_stats(self.h, C.byref(s)), "rvk_get_stats")
        return s

    def profile_begin(self) -> None:
        """Start timing every kernel launch with HIP events on the context stream."""
        _check(lib().rvk_profile_begin(self.h), "rvk_profile_begin")

    def profile_end(self) -> list[dict]:
        """Wait for the stream; per-kernel launches, total ms and algorithmic bytes."""
        cap = 64
        out = (abi.KernelStat * cap)()
        n = C.c_uint32(0)
        _check(lib().rvk_profile_end(self.h, out, cap, C.byref(n)), "rvk_profile_end")
        return [dict(name=out[i].name.decode(), launches=int(out[i].launches), total_ms=float(out[i].total_ms),
                     bytes=float(out[i].bytes)) for i in range(n.value)]

    # --- filter primitives (in place on a (H, W, 4) uint8 device tensor) ---
    def into_linear_rgb(self, px):
        w, h = _dims(px)
        _check(lib().rvk_into_linear_rgb(self.h, _ptr(px), w, h), "rvk_into_linear_rgb")

    def into_srgb(self, px):
        w, h = _dims(px)
        _check(lib().rvk_into_srgb(self.h, _ptr(px), w, h), "rvk_into_srgb")

    def demultiply_alpha(self, px):
        w, h = _dims(px)
        _check(lib().rvk_demultiply_alpha(self.h, _ptr(px), w, h), "rvk_demultiply_alpha")

    def multiply_alpha(self, px):
        w, h = _dims(px)
        _check(lib().rvk_multiply_alpha(self.h, _ptr(px), w, h), "rvk_multiply_alpha")

    def box_blur(self, px, sigma_x: float, sigma_y: float):
        w, h = _dims(px)
        _check(lib().rvk_box_blur(self.h, _ptr(px), w, h, sigma_x, sigma_y), "rvk_box_blur")

    def box_blur_noalloc(self, px, sigma_x: float, sigma_y: float) -> int:
        """rvk_box_blur_noalloc: the box blur without any device allocation (its back
        buffer is the library's 64 MiB device array); returns the status."""
        w, h = _dims(px)
        return lib().rvk_box_blur_noalloc(self.h, _ptr(px), w, h, sigma_x, sigma_y)

    def iir_blur(self, px, sigma_x: float, sigma_y: float):
        w, h = _dims(px)
        _check(lib().rvk_iir_blur(self.h, _ptr(px), w, h, sigma_x, sigma_y), "rvk_iir_blur")

    def morphology(self, px, op: int, rx: float, ry: float):
        w, h = _dims(px)
        _check(lib().rvk_morphology(self.h, _ptr(px), w, h, op, rx, ry), "rvk_morphology")

    def convolve_matrix(self, px, prim: abi.Primitive, values):
        w, h = _dims(px)
        vals = abi.array(C.c_float, values)
        _check(lib().rvk_convolve_matrix(self.h, _ptr(px), w, h, C.byref(prim), vals), "rvk_convolve_matrix")

    def color_matrix(self, px, prim: abi.Primitive):
        w, h = _dims(px)
        _check(lib().rvk_color_matrix(self.h, _ptr(px), w, h, C.byref(prim)), "rvk_color_matrix")

    def component_transfer(self, px, prim: abi.Primitive, values):
        w, h = _dims(px)
        vals = abi.array(C.c_float, values)
        _check(lib().rvk_component_transfer(self.h, _ptr(px), w, h, C.byref(prim), vals),
               "rvk_component_transfer")

    def composite_arithmetic(self, src1, src2, dst, k1, k2, k3, k4):
        w, h = _dims(dst)
        _check(lib().rvk_composite_arithmetic(self.h, _ptr(src1), _ptr(src2), _ptr(dst), w, h, k1, k2, k3, k4),
               "rvk_composite_arithmetic")

    def turbulence(self, dst, offset_x, offset_y, sx, sy, bfx, bfy, octaves, seed, stitch, fractal):
        w, h = _dims(dst)
        _check(lib().rvk_turbulence(self.h, _ptr(dst), w, h, offset_x, offset_y, sx, sy, bfx, bfy, octaves, seed,
                                    int(stitch), int(fractal)), "rvk_turbulence")

    def lighting(self, src, dst, specular, surface_scale, constant, specular_exponent, color, light):
        w, h = _dims(dst)
        col = (C.c_uint8 * 4)(*color[:3], 255)
        _check(lib().rvk_lighting(self.h, _ptr(src), _ptr(dst), w, h, int(specular), surface_scale, constant,
                                  specular_exponent, col, C.byref(light)), "rvk_lighting")

    def displacement_map(self, src, map_, dst, sx, sy, scale, x_channel, y_channel):
        w, h = _dims(dst)
        _check(lib().rvk_displacement_map(self.h, _ptr(src), _ptr(map_), _ptr(dst), w, h, sx, sy, scale, x_channel,
                                          y_channel), "rvk_displacement_map")

    def draw_pixmap(self, dst, x: int, y: int, src, opacity: float = 1.0, blend_mode: int = abi.BLEND["source_over"]):
        dw, dh = _dims(dst)
        sw, sh = _dims(src)
        _check(lib().rvk_draw_pixmap(self.h, _ptr(dst), dw, dh, x, y, _ptr(src), sw, sh, opacity, blend_mode),
               "rvk_draw_pixmap")

    def fill(self, px, rgba):
        w, h = _dims(px)
        col = (C.c_uint8 * 4)(*rgba)
        _check(lib().rvk_fill(self.h, _ptr(px), w, h, col), "rvk_fill")

    def filter_apply(self, graph, ts: abi.Transform, px) -> int:
        """filter::apply on a layer; returns the filter error status (0 when none)."""
        w, h = _dims(px)
        f, d, keep = graph.to_abi()
        err = C.c_int32(0)
        _check(lib().rvk_filter_apply(self.h, C.byref(f), C.byref(d), ts, _ptr(px), w, h, C.byref(err)),
               "rvk_filter_apply")
        del keep
        return err.value

    # --- PNG output (tiny-skia Pixmap::encode_png) ---
    def png_filter_rows(self, px, out):
        """rvk_png_filter_rows: the IDAT stream's uncompressed rows of a (H, W, 4) uint8
        device tensor written to `out`, a contiguous uint8 device tensor of at least
        H * (4W + 1) bytes."""
        w, h = _dims(px)
        if out.dtype.itemsize != 1 or not out.is_contiguous() or out.numel() < h * (4 * w + 1):
            raise ValueError("out must be a contiguous uint8 tensor of H * (4W + 1) bytes")
        _check(lib().rvk_png_filter_rows(self.h, _ptr(px), w, h, _ptr(out)), "rvk_png_filter_rows")

    def encode_png(self, px, level: int = 6) -> bytes:
        """Pixmap::encode_png of a premultiplied (H, W, 4) uint8 device tensor: the
        device demultiplies and filters the rows, the host deflates them (zlib
        `level`) and returns the PNG file's bytes."""
        import torch

        from . import png
        w, h = _dims(px)
        rows = torch.empty(h * (4 * w + 1), dtype=torch.uint8, device=px.device)
        self.png_filter_rows(px, rows)
        self.synchronize()
        return png.from_rows(w, h, rows.cpu().numpy().tobytes(), level)

    # --- rendering ---
    def render(self, scene, ts: abi.Transform, pixmap):
        """resvg::render onto a (H, W, 4) uint8 device tensor (composites, never clears)."""
        w, h = _dims(pixmap)
        sc = scene.to_abi() if hasattr(scene, "to_abi") else scene
        _check(lib().rvk_render(self.h, C.byref(sc.abi), ts, w, h, _ptr(pixmap)), "rvk_render")

    def render_batch(self, scenes, transforms, pixmaps, check: bool = True) -> list[int]:
        """Several resvg::render calls in one: scenes[i] onto the (H, W, 4) uint8 device
        tensor pixmaps[i] with transforms[i] (one Transform for all, or a list).  Returns
        the per-document statuses; raises on the first failure when `check`."""
        n = len(scenes)
        if isinstance(transforms, abi.Transform):
            transforms = [transforms] * n
        assert len(transforms) == n and len(pixmaps) == n
        abis = [sc.to_abi() if hasattr(sc, "to_abi") else sc for sc in scenes]
        dims = [_dims(px) for px in pixmaps]
        sc_arr = (C.POINTER(abi.Scene) * n)(*[C.pointer(a.abi) for a in abis])
        ts_arr = (abi.Transform * n)(*transforms)
        w_arr = (C.c_uint32 * n)(*[d[0] for d in dims])
        h_arr = (C.c_uint32 * n)(*[d[1] for d in dims])
        px_arr = (C.c_void_p * n)(*[_ptr(px) for px in pixmaps])
        st_arr = (C.c_int32 * n)()
        status = lib().rvk_render_batch(self.h, n, sc_arr, ts_arr, w_arr, h_arr, px_arr, st_arr)
        if check:
            _check(status, "rvk_render_batch")
        return list(st_arr)

    def render_node(self, scene, node, ts: abi.Transform, pixmap, abs_layer_bbox=None) -> bool:
        """resvg::render_node: draws `node` ((abi.NODE_GROUP | abi.NODE_PATH, index))
        with `ts` pre-translated by minus its absolute layer bounding box's origin.
        The box is usvg's Node::abs_layer_bounding_box(); by default it is taken from
        the scene builder.  Returns False (and draws nothing) when the node has none."""
        w, h = _dims(pixmap)
        kind, index = node
        if abs_layer_bbox is None and hasattr(scene, "abs_layer_bbox"):
            abs_layer_bbox = scene.abs_layer_bbox(kind, index)
        sc = scene.to_abi() if hasattr(scene, "to_abi") else scene
        bb = C.byref(abi.Rect(*abs_layer_bbox)) if abs_layer_bbox is not None else None
        done = C.c_int32(0)
        _check(lib().rvk_render_node(self.h, C.byref(sc.abi), abi.NodeRef(kind, index), bb, ts, w, h, _ptr(pixmap),
                                     C.byref(done)), "rvk_render_node")
        return bool(done.value)

    # --- per-operation replacements of tiny-skia's pixel calls ---
    @staticmethod
    def _geometry(verbs, points):
        import numpy as np
        v = abi.array(C.c_uint8, [int(x) for x in verbs])
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1))
        g = abi.PathGeometry(C.cast(v, C.POINTER(C.c_uint8)), len(verbs),
                             pts.ctypes.data_as(C.POINTER(C.c_float)), pts.size // 2)
        return g, (v, pts)

    def fill_path(self, px, verbs, points, paint, fill_rule: int = abi.FILL_NONZERO,
                  ts: abi.Transform | None = None, blend_mode: int = abi.BLEND["source_over"],
                  anti_alias: bool = True):
        """path::fill_path (path.rs:26-75) into a device pixmap; `paint` a scene.PaintSpec."""
        from .scene import paint_to_abi
        w, h = _dims(px)
        g, keep = self._geometry(verbs, points)
        stops = []
        pa = paint_to_abi(paint, stops)
        st = abi.array(abi.Stop, stops)
        _check(lib().rvk_fill_path(self.h, _ptr(px), w, h, C.byref(g), fill_rule, C.byref(pa), st,
                                   ts or abi.Transform.identity(), blend_mode, int(anti_alias)), "rvk_fill_path")
        del keep

    def stroke_path(self, px, verbs, points, paint, width: float, cap: int = abi.CAP_BUTT,
                    join: int = abi.JOIN_MITER, miter_limit: float = 4.0, dasharray=None, dashoffset: float = 0.0,
                    ts: abi.Transform | None = None, anti_alias: bool = True):
        """path::stroke_path (path.rs:77-116) into a device pixmap."""
        from .scene import paint_to_abi
        w, h = _dims(px)
        g, keep = self._geometry(verbs, points)
        stops = []
        pa = paint_to_abi(paint, stops)
        st = abi.array(abi.Stop, stops)
        d = abi.array(C.c_float, dasharray or ())
        style = abi.StrokeStyle(width, miter_limit, cap, join, C.cast(d, C.POINTER(C.c_float)),
                                len(dasharray or ()), dashoffset)
        _check(lib().rvk_stroke_path(self.h, _ptr(px), w, h, C.byref(g), C.byref(style), C.byref(pa), st,
                                     ts or abi.Transform.identity(), int(anti_alias)), "rvk_stroke_path")
        del keep

    def mask_fill_path(self, mask, verbs, points, fill_rule: int = abi.FILL_NONZERO, anti_alias: bool = True,
                       ts: abi.Transform | None = None):
        """tiny_skia::Mask::fill_path on a (H, W) uint8 device mask: writes the coverage."""
        h, w = mask.shape
        g, keep = self._geometry(verbs, points)
        _check(lib().rvk_mask_fill_path(self.h, _ptr(mask), w, h, C.byref(g), fill_rule, int(anti_alias),
                                        ts or abi.Transform.identity()), "rvk_mask_fill_path")
        del keep

    def mask_from_pixmap(self, px, mask, mask_type: int = abi.MASK_TYPE_ALPHA):
        w, h = _dims(px)
        _check(lib().rvk_mask_from_pixmap(self.h, _ptr(px), w, h, mask_type, _ptr(mask)), "rvk_mask_from_pixmap")

    def mask_invert(self, mask):
        h, w = mask.shape
        _check(lib().rvk_mask_invert(self.h, _ptr(mask), w, h), "rvk_mask_invert")

    def apply_mask(self, px, mask):
        w, h = _dims(px)
        _check(lib().rvk_apply_mask(self.h, _ptr(px), w, h, _ptr(mask)), "rvk_apply_mask")

    def clear_rect(self, px, x: int, y: int, rw: int, rh: int):
        w, h = _dims(px)
        _check(lib().rvk_clear_rect(self.h, _ptr(px), w, h, x, y, rw, rh), "rvk_clear_rect")

    def render_host(self, scene, ts: abi.Transform, pixmap_np):
        """resvg_render twin: host numpy (H, W, 4) uint8 in and out."""
        import numpy as np
        assert pixmap_np.dtype == np.uint8 and pixmap_np.flags.c_contiguous
        h, w, _ = pixmap_np.shape
        sc = scene.to_abi() if hasattr(scene, "to_abi") else scene
        _check(lib().rvk_render_host(self.h, C.byref(sc.abi), ts, w, h, pixmap_np.ctypes.data), "rvk_render_host")
