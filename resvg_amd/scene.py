"""Flattened scene builder: a usvg::Tree as the rvk_scene display list.

The builder plays usvg's role for synthetic documents: it stores groups (transform,
opacity, blend mode, filters, layer bounding box), paths (verbs, points, fill and
stroke paints) and gradient stops exactly as resvg's render walk consumes them
(crates/usvg/src/tree/mod.rs).  Layer bounding boxes are computed conservatively
from the points of every path (control points included) grown by the stroke's
reach, so a path never extends past its layer.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .filters import FilterGraph

F32 = np.float32


def f32(v: float) -> float:
    return float(F32(v))


@dataclass
class PaintSpec:
    kind: int = abi.PAINT_NONE
    color: tuple = (0, 0, 0)
    opacity: float = 1.0
    spread: int = abi.SPREAD_PAD
    x1: float = 0.0
    y1: float = 0.0
    x2: float = 1.0
    y2: float = 0.0
    cx: float = 0.0
    cy: float = 0.0
    r: float = 1.0
    fx: float = 0.0
    fy: float = 0.0
    fr: float = 0.0
    transform: tuple = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
    stops: list = field(default_factory=list)  # (offset, (r, g, b), opacity)
    rect: tuple = (0.0, 0.0, 0.0, 0.0)  # pattern tile (x, y, w, h)
    root: int = 0                        # pattern content group

    @staticmethod
    def solid(color, opacity=1.0) -> "PaintSpec":
        return PaintSpec(kind=abi.PAINT_COLOR, color=tuple(color), opacity=opacity)

    @staticmethod
    def linear(x1, y1, x2, y2, stops, opacity=1.0, spread=abi.SPREAD_PAD, transform=None) -> "PaintSpec":
        return PaintSpec(kind=abi.PAINT_LINEAR, x1=x1, y1=y1, x2=x2, y2=y2, stops=list(stops), opacity=opacity,
                         spread=spread, transform=transform or (1.0, 0.0, 0.0, 1.0, 0.0, 0.0))

    @staticmethod
    def radial(cx, cy, r, stops, opacity=1.0, spread=abi.SPREAD_PAD, fx=None, fy=None, fr=0.0,
               transform=None) -> "PaintSpec":
        return PaintSpec(kind=abi.PAINT_RADIAL, cx=cx, cy=cy, r=r, fx=cx if fx is None else fx,
                         fy=cy if fy is None else fy, fr=fr, stops=list(stops), opacity=opacity, spread=spread,
                         transform=transform or (1.0, 0.0, 0.0, 1.0, 0.0, 0.0))

    @staticmethod
    def pattern(rect, root, opacity=1.0, transform=None) -> "PaintSpec":
        """A usvg::Pattern with its units resolved: the tile rect in user space, the
        content group (Scene.content_group, children in tile space) and patternTransform."""
        return PaintSpec(kind=abi.PAINT_PATTERN, rect=tuple(f32(v) for v in rect), root=root, opacity=opacity,
                         transform=transform or (1.0, 0.0, 0.0, 1.0, 0.0, 0.0))


def paint_to_abi(ps: PaintSpec, stops: list) -> abi.Paint:
    """The rvk_paint of a PaintSpec; its stops are appended to `stops` (abi.Stop)."""
    P = abi.Paint()
    P.kind = ps.kind
    for k in range(3):
        P.color[k] = ps.color[k]
    P.opacity = ps.opacity
    P.spread = ps.spread
    P.x1, P.y1, P.x2, P.y2 = ps.x1, ps.y1, ps.x2, ps.y2
    P.cx, P.cy, P.r, P.fx, P.fy, P.fr = ps.cx, ps.cy, ps.r, ps.fx, ps.fy, ps.fr
    P.transform = abi.Transform(*ps.transform)
    P.rect = abi.Rect(*ps.rect)
    P.root = ps.root
    P.stop_offset = len(stops)
    P.stop_count = len(ps.stops)
    for off, col, op in ps.stops:
        s = abi.Stop()
        s.offset = off
        for k in range(3):
            s.color[k] = col[k]
        s.opacity = op
        stops.append(s)
    return P


class Scene:
    """Builder; node 0 is the root group."""

    def __init__(self, width: float, height: float):
        self.width, self.height = width, height
        self.groups: list[dict] = []
        self.paths: list[dict] = []
        self.verbs: list[int] = []
        self._point_chunks: list[np.ndarray] = []  # float32 x, y pairs of every path, in order
        self._n_point_values = 0
        self.stops: list[tuple] = []
        self.filters: list[tuple] = []  # (rect, prim_offset, prim_count)
        self.primitives: list = []
        self.values: list[float] = []
        self.inputs: list[int] = []
        self.dashes: list[float] = []
        self.clip_paths: list[dict] = []
        self.masks: list[dict] = []
        self.images: list[dict] = []
        self.root = self._new_group(None, (1.0, 0.0, 0.0, 1.0, 0.0, 0.0), 1.0, abi.MIX["normal"], False, [])
        self._abi = None

    # ---- construction ----
    def _new_group(self, parent, transform, opacity, blend, isolate, filters, clip_path=-1, mask=-1):
        gid = len(self.groups)
        self.groups.append(dict(transform=tuple(transform), opacity=opacity, blend=blend, isolate=isolate,
                                filters=list(filters), children=[], bbox=None, clip_path=clip_path, mask=mask))
        if parent is not None:
            self.groups[parent]["children"].append((abi.NODE_GROUP, gid))
        self._abi = None
        return gid

    def clip_path(self, transform=(1.0, 0.0, 0.0, 1.0, 0.0, 0.0), clip_path: int = -1) -> tuple[int, int]:
        """A usvg::ClipPath; returns (clip path index, group to add its children to).
        Children are drawn as fills (with the clip rule as their fill rule) and groups."""
        root = self._new_group(None, (1.0, 0.0, 0.0, 1.0, 0.0, 0.0), 1.0, abi.MIX["normal"], False, [])
        self.clip_paths.append(dict(transform=tuple(transform), clip_path=clip_path, root=root))
        self._abi = None
        return len(self.clip_paths) - 1, root

    def mask(self, rect, kind: int = abi.MASK_LUMINANCE, mask: int = -1) -> tuple[int, int]:
        """A usvg::Mask with its region in user space; returns (mask index, content group)."""
        root = self._new_group(None, (1.0, 0.0, 0.0, 1.0, 0.0, 0.0), 1.0, abi.MIX["normal"], False, [])
        self.masks.append(dict(rect=tuple(f32(v) for v in rect), kind=kind, mask=mask, root=root))
        self._abi = None
        return len(self.masks) - 1, root

    def content_group(self) -> int:
        """A group outside the tree (usvg filter::Image::root, the element an feImage
        links to): only an feImage primitive draws its children."""
        return self._new_group(None, (1.0, 0.0, 0.0, 1.0, 0.0, 0.0), 1.0, abi.MIX["normal"], False, [])

    def group(self, parent: int, transform=(1.0, 0.0, 0.0, 1.0, 0.0, 0.0), opacity: float = 1.0,
              blend: str | int = "normal", isolate: bool = False, filters: list[FilterGraph] = (),
              clip_path: int = -1, mask: int = -1) -> int:
        mode = abi.MIX[blend] if isinstance(blend, str) else blend
        fids = []
        for g in filters:
            off = len(self.primitives)
            base_v, base_i = len(self.values), len(self.inputs)
            for p in g.primitives:
                q = abi.Primitive.from_buffer_copy(p)
                if q.kind == abi.FE_CONVOLVE_MATRIX:
                    q.u.convolve.data_offset += base_v
                elif q.kind == abi.FE_COMPONENT_TRANSFER:
                    for k in range(4):
                        q.u.component_transfer.func[k].table_offset += base_v
                elif q.kind == abi.FE_MERGE:
                    q.u.merge.input_offset += base_i
                self.primitives.append(q)
            self.values.extend(g.values)
            self.inputs.extend(g.inputs)
            fids.append(len(self.filters))
            self.filters.append((tuple(g.rect), off, len(g.primitives)))
        return self._new_group(parent, transform, opacity, mode, isolate, fids, clip_path, mask)

    def path(self, parent: int, verbs, points, fill: PaintSpec | None = None, fill_rule: int = abi.FILL_NONZERO,
             stroke: PaintSpec | None = None, stroke_width: float = 1.0, stroke_join: int = abi.JOIN_MITER,
             stroke_cap: int = abi.CAP_BUTT, miter_limit: float = 4.0, anti_alias: bool = True,
             visible: bool = True, stroke_first: bool = False, dasharray=None, dashoffset: float = 0.0) -> int:
        """`dasharray` is usvg's resolved list (non-negative, even length, non-zero sum)."""
        arr = np.asarray(points, dtype=np.float64).reshape(-1).astype(np.float32)  # f32(v) of every value
        pid = len(self.paths)
        self.paths.append(dict(verb_offset=len(self.verbs), verb_count=len(verbs), point_offset=self._n_point_values // 2,
                               fill=fill or PaintSpec(), fill_rule=fill_rule, stroke=stroke or PaintSpec(),
                               stroke_width=stroke_width, join=stroke_join, cap=stroke_cap, miter=miter_limit,
                               aa=anti_alias, visible=visible, stroke_first=stroke_first,
                               npts=len(arr) // 2, pts=arr.reshape(-1, 2),
                               dash_offset=len(self.dashes), dash_count=len(dasharray or ()),
                               dash_phase=f32(dashoffset)))
        self.dashes.extend(f32(v) for v in (dasharray or ()))
        self.verbs.extend(int(v) for v in verbs)
        self._point_chunks.append(arr)
        self._n_point_values += len(arr)
        self.groups[parent]["children"].append((abi.NODE_PATH, pid))
        self._abi = None
        return pid

    def image(self, parent: int, pixels, quality: int = abi.QUALITY_BICUBIC, visible: bool = True) -> int:
        """A raster image node: `pixels` is the decoded image as tiny-skia holds it
        (H x W x 4 premultiplied RGBA8), drawn over (0, 0, W, H) in the parent's space
        (usvg puts the image's view box transform on a group of its own)."""
        px = np.ascontiguousarray(pixels, dtype=np.uint8)
        if px.ndim != 3 or px.shape[2] != 4:
            raise ValueError("pixels must be H x W x 4 uint8")
        iid = len(self.images)
        self.images.append(dict(pixels=px, quality=int(quality), visible=bool(visible)))
        self.groups[parent]["children"].append((abi.NODE_IMAGE, iid))
        self._abi = None
        return iid

    # ---- bounding boxes (usvg's role) ----
    def _path_bounds(self, p):
        pts = p["pts"]
        if len(pts) == 0:
            return None
        x0, y0 = float(pts[:, 0].min()), float(pts[:, 1].min())
        x1, y1 = float(pts[:, 0].max()), float(pts[:, 1].max())
        return (x0, y0, x1, y1)

    def _path_layer_bounds(self, p):
        b = self._path_bounds(p)
        if b is None:
            return None
        if p["stroke"].kind != abi.PAINT_NONE:
            reach = p["stroke_width"] * 0.5
            if p["join"] in (abi.JOIN_MITER, abi.JOIN_MITER_CLIP):
                reach *= max(1.0, p["miter"])
            if p["cap"] == abi.CAP_SQUARE:
                reach *= math.sqrt(2.0)
            reach += 1.0
            b = (b[0] - reach, b[1] - reach, b[2] + reach, b[3] + reach)
        return b

    @staticmethod
    def _transform_box(b, t):
        a, bb, c, d, e, f = t
        xs, ys = [], []
        for x, y in ((b[0], b[1]), (b[2], b[1]), (b[2], b[3]), (b[0], b[3])):
            xs.append(a * x + c * y + e)
            ys.append(bb * x + d * y + f)
        return (min(xs), min(ys), max(xs), max(ys))

    def _group_layer_bounds(self, gid):
        g = self.groups[gid]
        if g["filters"]:
            box = None
            for fi in g["filters"]:
                x, y, w, h = self.filters[fi][0]
                fb = (x, y, x + w, y + h)
                box = fb if box is None else (min(box[0], fb[0]), min(box[1], fb[1]), max(box[2], fb[2]),
                                              max(box[3], fb[3]))
            return box
        box = None
        for kind, idx in g["children"]:
            if kind == abi.NODE_PATH:
                cb = self._path_layer_bounds(self.paths[idx])
            elif kind == abi.NODE_IMAGE:  # Image::bounding_box: the size at the origin
                h, w = self.images[idx]["pixels"].shape[:2]
                cb = (0.0, 0.0, float(w), float(h))
            else:
                cb = self._group_layer_bounds(idx)
                if cb is not None:
                    cb = self._transform_box(cb, self.groups[idx]["transform"])
            if cb is None:
                continue
            box = cb if box is None else (min(box[0], cb[0]), min(box[1], cb[1]), max(box[2], cb[2]),
                                          max(box[3], cb[3]))
        return box

    # ---- absolute bounding boxes (usvg's Node::abs_layer_bounding_box) ----
    @staticmethod
    def _concat(a, b):
        """a.pre_concat(b): maps by b first, then a (6-tuples sx, ky, kx, sy, tx, ty)."""
        return (a[0] * b[0] + a[2] * b[1], a[1] * b[0] + a[3] * b[1], a[0] * b[2] + a[2] * b[3],
                a[1] * b[2] + a[3] * b[3], a[0] * b[4] + a[2] * b[5] + a[4], a[1] * b[4] + a[3] * b[5] + a[5])

    def _parent_transforms(self):
        """Per group, the product of its ancestors' transforms (its parent's abs transform)."""
        out = {self.root: (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)}
        stack = [self.root]
        while stack:
            g = stack.pop()
            abs_g = self._concat(out[g], self.groups[g]["transform"])
            for kind, idx in self.groups[g]["children"]:
                if kind == abi.NODE_GROUP:
                    out[idx] = abs_g
                    stack.append(idx)
        return out

    def abs_layer_bbox(self, kind: int, index: int):
        """(x, y, w, h) of the node's layer bounding box in canvas coordinates, or None:
        a group's layer box mapped by its absolute transform (its ancestors' and its
        own), a path's fill bounds or an image's rect mapped by its parent's (None for a
        zero width or height, as NonZeroRect)."""
        parents = self._parent_transforms()
        if kind == abi.NODE_GROUP:
            t = self._concat(parents[index], self.groups[index]["transform"])
            b = self._group_layer_bounds(index)
        else:
            owner = next(g for g, grp in enumerate(self.groups) if (kind, index) in grp["children"])
            t = self._concat(parents.get(owner, (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)), self.groups[owner]["transform"])
            if kind == abi.NODE_IMAGE:
                h, w = self.images[index]["pixels"].shape[:2]
                b = (0.0, 0.0, float(w), float(h))
            else:
                b = self._path_bounds(self.paths[index])
        if b is None:
            return None
        x0, y0, x1, y1 = self._transform_box(b, t)
        w, h = f32(x1 - x0), f32(y1 - y0)
        if not (w > 0 and h > 0):
            return None
        return (f32(x0), f32(y0), w, h)

    # ---- ABI ----
    def to_abi(self):
        if self._abi is not None:
            return self._abi
        groups = (abi.Group * len(self.groups))()
        children = []
        for i, g in enumerate(self.groups):
            G = groups[i]
            G.transform = abi.Transform(*g["transform"])
            G.opacity = g["opacity"]
            G.blend_mode = g["blend"]
            G.isolate = int(g["isolate"])
            G.clip_path = g["clip_path"]
            G.mask = g["mask"]
            G.filter_offset = g["filters"][0] if g["filters"] else 0
            G.filter_count = len(g["filters"])
            bb = self._group_layer_bounds(i)
            if bb is None or not (bb[2] > bb[0] and bb[3] > bb[1]):
                bb = (0.0, 0.0, 1.0, 1.0)
            G.layer_bbox = abi.Rect(bb[0], bb[1], f32(bb[2] - bb[0]), f32(bb[3] - bb[1]))
            G.child_offset = len(children)
            G.child_count = len(g["children"])
            children.extend(abi.NodeRef(k, j) for k, j in g["children"])
        paths = (abi.Path * max(1, len(self.paths)))()
        stops = []

        def paint(ps: PaintSpec) -> abi.Paint:
            return paint_to_abi(ps, stops)

        for i, p in enumerate(self.paths):
            Q = paths[i]
            Q.verb_offset, Q.verb_count, Q.point_offset = p["verb_offset"], p["verb_count"], p["point_offset"]
            Q.visible, Q.anti_alias, Q.paint_order_stroke_first = int(p["visible"]), int(p["aa"]), int(p["stroke_first"])
            Q.fill = paint(p["fill"])
            Q.fill_rule = p["fill_rule"]
            Q.stroke = paint(p["stroke"])
            Q.stroke_width, Q.stroke_miter_limit = p["stroke_width"], p["miter"]
            Q.stroke_cap, Q.stroke_join = p["cap"], p["join"]
            b = self._path_bounds(p) or (0.0, 0.0, 0.0, 0.0)
            Q.bounds = abi.Rect(b[0], b[1], f32(b[2] - b[0]), f32(b[3] - b[1]))
            Q.dash_offset, Q.dash_count, Q.dash_phase = p["dash_offset"], p["dash_count"], p["dash_phase"]
        ch = abi.array(abi.NodeRef, children)
        verbs = abi.array(C.c_uint8, self.verbs)
        pts = np.ascontiguousarray(np.concatenate(self._point_chunks) if self._point_chunks
                                   else np.zeros(1, np.float32))
        st = abi.array(abi.Stop, stops)
        filt = abi.array(abi.Filter, [abi.Filter(abi.Rect(*r), o, n) for r, o, n in self.filters])
        prims = abi.array(abi.Primitive, self.primitives)
        vals = abi.array(C.c_float, self.values)
        ins = abi.array(C.c_int32, self.inputs)
        sc = abi.Scene()
        sc.groups = C.cast(groups, C.POINTER(abi.Group))
        sc.group_count = len(self.groups)
        sc.paths = C.cast(paths, C.POINTER(abi.Path))
        sc.path_count = len(self.paths)
        sc.children = C.cast(ch, C.POINTER(abi.NodeRef))
        sc.verbs = C.cast(verbs, C.POINTER(C.c_uint8))
        sc.points = pts.ctypes.data_as(C.POINTER(C.c_float))
        sc.stops = C.cast(st, C.POINTER(abi.Stop))
        sc.filters = C.cast(filt, C.POINTER(abi.Filter))
        sc.filter_data = abi.FilterData(C.cast(prims, C.POINTER(abi.Primitive)), C.cast(vals, C.POINTER(C.c_float)),
                                        C.cast(ins, C.POINTER(C.c_int32)))
        cps = abi.array(abi.ClipPath, [abi.ClipPath(abi.Transform(*c["transform"]), c["clip_path"], c["root"])
                                       for c in self.clip_paths])
        msk = abi.array(abi.Mask, [abi.Mask(abi.Rect(*m["rect"]), m["kind"], m["mask"], m["root"]) for m in self.masks])
        sc.clip_paths = C.cast(cps, C.POINTER(abi.ClipPath))
        sc.clip_path_count = len(self.clip_paths)
        sc.masks = C.cast(msk, C.POINTER(abi.Mask))
        sc.mask_count = len(self.masks)
        dsh = abi.array(C.c_float, self.dashes)
        sc.dashes = C.cast(dsh, C.POINTER(C.c_float))
        imgs = (abi.Image * max(1, len(self.images)))()
        for i, im in enumerate(self.images):
            h, w = im["pixels"].shape[:2]
            imgs[i] = abi.Image(im["pixels"].ctypes.data_as(C.POINTER(C.c_uint8)), w, h, im["quality"],
                                int(im["visible"]))
        sc.images = C.cast(imgs, C.POINTER(abi.Image))
        sc.image_count = len(self.images)

        class _SceneABI:
            pass

        out = _SceneABI()
        out.abi = sc
        out.keep = (groups, paths, ch, verbs, pts, st, filt, prims, vals, ins, cps, msk, dsh, imgs,
                    [im["pixels"] for im in self.images])
        self._abi = out
        return out


# ---------------------------------------------------------------------------
# path helpers
# ---------------------------------------------------------------------------
def rect_path(x, y, w, h):
    verbs = [abi.VERB_MOVE, abi.VERB_LINE, abi.VERB_LINE, abi.VERB_LINE, abi.VERB_CLOSE]
    pts = [(x, y), (x + w, y), (x + w, y + h), (x, y + h)]
    return verbs, pts


def ellipse_path(cx, cy, rx, ry):
    """Four cubic arcs (the kappa approximation usvg uses for circles/ellipses)."""
    k = 0.5522847498307936
    verbs = [abi.VERB_MOVE] + [abi.VERB_CUBIC] * 4 + [abi.VERB_CLOSE]
    pts = [(cx + rx, cy),
           (cx + rx, cy + k * ry), (cx + k * rx, cy + ry), (cx, cy + ry),
           (cx - k * rx, cy + ry), (cx - rx, cy + k * ry), (cx - rx, cy),
           (cx - rx, cy - k * ry), (cx - k * rx, cy - ry), (cx, cy - ry),
           (cx + k * rx, cy - ry), (cx + rx, cy - k * ry), (cx + rx, cy)]
    return verbs, pts
