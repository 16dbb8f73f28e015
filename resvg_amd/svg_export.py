"""Writes a flattened Scene back out as an SVG document, so that the seeded
workloads can be rendered by the reference itself (``resvg`` reads the file through
usvg, which rebuilds the same tree).

Every number is printed with ``%.9g``, which round-trips exactly to the f32 value the
scene holds (usvg parses in f64 and stores f32).  The document uses user-space units
throughout: gradients and patterns are ``userSpaceOnUse``, filters have explicit
``userSpaceOnUse`` regions and primitive subregions and name every input and result,
and each primitive carries its ``color-interpolation-filters``.  Paint opacity becomes
``fill-opacity`` / ``stroke-opacity`` (usvg multiplies it into every stop, as the
scene's stops expect).  Groups inside a clip path, which SVG does not allow there,
are flattened into their paths' ``transform``.  Raster images are embedded as PNG
data after demultiplying their premultiplied pixels.
"""
from __future__ import annotations

import base64

import numpy as np

from . import abi, png

_JOIN = {abi.JOIN_MITER: "miter", abi.JOIN_MITER_CLIP: "miter-clip", abi.JOIN_ROUND: "round",
         abi.JOIN_BEVEL: "bevel"}
_CAP = {abi.CAP_BUTT: "butt", abi.CAP_ROUND: "round", abi.CAP_SQUARE: "square"}
_SPREAD = {abi.SPREAD_PAD: "pad", abi.SPREAD_REFLECT: "reflect", abi.SPREAD_REPEAT: "repeat"}
_RULE = {abi.FILL_NONZERO: "nonzero", abi.FILL_EVENODD: "evenodd"}
_COMPOSITE = {abi.COMPOSITE_OVER: "over", abi.COMPOSITE_IN: "in", abi.COMPOSITE_OUT: "out",
              abi.COMPOSITE_ATOP: "atop", abi.COMPOSITE_XOR: "xor", abi.COMPOSITE_ARITHMETIC: "arithmetic"}
_CM = {abi.COLOR_MATRIX_MATRIX: "matrix", abi.COLOR_MATRIX_SATURATE: "saturate",
       abi.COLOR_MATRIX_HUE_ROTATE: "hueRotate", abi.COLOR_MATRIX_LUMINANCE_TO_ALPHA: "luminanceToAlpha"}
_CM_COUNT = {abi.COLOR_MATRIX_MATRIX: 20, abi.COLOR_MATRIX_SATURATE: 1, abi.COLOR_MATRIX_HUE_ROTATE: 1,
             abi.COLOR_MATRIX_LUMINANCE_TO_ALPHA: 0}
_EDGE = {abi.EDGE_NONE: "none", abi.EDGE_DUPLICATE: "duplicate", abi.EDGE_WRAP: "wrap"}
_CHANNEL = "RGBA"
_QUALITY = {abi.QUALITY_NEAREST: "optimizeSpeed", abi.QUALITY_BILINEAR: "smooth",
            abi.QUALITY_BICUBIC: "optimizeQuality"}
# the tiny-skia blend mode of an feBlend primitive, by its SVG name
_FE_BLEND = {abi.BLEND["source_over"]: "normal", abi.BLEND["multiply"]: "multiply", abi.BLEND["screen"]: "screen",
             abi.BLEND["overlay"]: "overlay", abi.BLEND["darken"]: "darken", abi.BLEND["lighten"]: "lighten",
             abi.BLEND["color_dodge"]: "color-dodge", abi.BLEND["color_burn"]: "color-burn",
             abi.BLEND["hard_light"]: "hard-light", abi.BLEND["soft_light"]: "soft-light",
             abi.BLEND["difference"]: "difference", abi.BLEND["exclusion"]: "exclusion", abi.BLEND["hue"]: "hue",
             abi.BLEND["saturation"]: "saturation", abi.BLEND["color"]: "color",
             abi.BLEND["luminosity"]: "luminosity"}
_TRANSFER = {abi.TRANSFER_IDENTITY: "identity", abi.TRANSFER_TABLE: "table", abi.TRANSFER_DISCRETE: "discrete",
             abi.TRANSFER_LINEAR: "linear", abi.TRANSFER_GAMMA: "gamma"}
IDENTITY = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)


def num(v) -> str:
    """The decimal form of an f32 value that parses back to the same f32."""
    return "%.9g" % float(np.float32(v))


def nums(vals) -> str:
    return " ".join(num(v) for v in vals)


def _matrix(t) -> str:
    return f"matrix({nums(t)})"


def _concat(a, b):
    """a * b (b applied first)."""
    return (a[0] * b[0] + a[2] * b[1], a[1] * b[0] + a[3] * b[1], a[0] * b[2] + a[2] * b[3],
            a[1] * b[2] + a[3] * b[3], a[0] * b[4] + a[2] * b[5] + a[4], a[1] * b[4] + a[3] * b[5] + a[5])


def _color(c) -> str:
    return "#%02x%02x%02x" % (int(c[0]), int(c[1]), int(c[2]))


def png_bytes(rgba: np.ndarray) -> bytes:
    """A PNG file (8-bit RGBA, filter type None, zlib level 6) of straight-alpha pixels."""
    rgba = np.ascontiguousarray(rgba, np.uint8)
    h, w, _ = rgba.shape
    return png.from_rows(w, h, b"".join(b"\x00" + rgba[y].tobytes() for y in range(h)))


def demultiply(px: np.ndarray) -> np.ndarray:
    """Premultiplied RGBA8 to straight alpha, c / (a / 255) + 0.5 truncated (0 where
    a = 0), in f64."""
    px = np.asarray(px, np.uint8)
    a = px[..., 3:4].astype(np.float64) / 255.0
    safe = np.where(a > 0, a, 1.0)
    c = np.where(a > 0, px[..., :3] / safe + 0.5, 0.0)
    out = px.copy()
    out[..., :3] = np.clip(c, 0, 255).astype(np.uint8)
    return out


def _path_data(verbs, pts) -> str:
    out, i = [], 0
    for v in verbs:
        if v == abi.VERB_MOVE:
            out.append("M" + nums(pts[i]))
            i += 1
        elif v == abi.VERB_LINE:
            out.append("L" + nums(pts[i]))
            i += 1
        elif v == abi.VERB_QUAD:
            out.append("Q" + nums(pts[i:i + 2].reshape(-1)))
            i += 2
        elif v == abi.VERB_CUBIC:
            out.append("C" + nums(pts[i:i + 3].reshape(-1)))
            i += 3
        elif v == abi.VERB_CLOSE:
            out.append("Z")
        else:
            raise ValueError(f"unknown path verb {v}")
    return " ".join(out)


class _Writer:
    def __init__(self, sc):
        self.sc = sc
        self.defs: list[str] = []
        self.n_paint = 0
        self.done_content: set[int] = set()

    # ---- paints ----
    def paint(self, ps, what: str) -> str:
        """The fill or stroke attributes of a paint (its gradient or pattern goes to defs)."""
        if ps.kind == abi.PAINT_NONE:
            return f' {what}="none"'
        op = "" if float(np.float32(ps.opacity)) == 1.0 else f' {what}-opacity="{num(ps.opacity)}"'
        if ps.kind == abi.PAINT_COLOR:
            return f' {what}="{_color(ps.color)}"{op}'
        pid = f"paint{self.n_paint}"
        self.n_paint += 1
        ts = "" if tuple(float(v) for v in ps.transform) == IDENTITY else f' {{}}="{_matrix(ps.transform)}"'
        if ps.kind == abi.PAINT_PATTERN:
            x, y, w, h = ps.rect
            body = self.children(ps.root, IDENTITY)
            self.defs.append(f'<pattern id="{pid}" patternUnits="userSpaceOnUse" patternContentUnits="userSpaceOnUse" '
                             f'x="{num(x)}" y="{num(y)}" width="{num(w)}" height="{num(h)}"'
                             f'{ts.format("patternTransform")}>{body}</pattern>')
        else:
            stops = "".join(
                f'<stop offset="{num(off)}" stop-color="{_color(col)}"'
                + ("" if float(np.float32(sop)) == 1.0 else f' stop-opacity="{num(sop)}"') + "/>"
                for off, col, sop in ps.stops)
            common = (f' gradientUnits="userSpaceOnUse" spreadMethod="{_SPREAD[ps.spread]}"'
                      f'{ts.format("gradientTransform")}')
            if ps.kind == abi.PAINT_LINEAR:
                self.defs.append(f'<linearGradient id="{pid}" x1="{num(ps.x1)}" y1="{num(ps.y1)}" x2="{num(ps.x2)}" '
                                 f'y2="{num(ps.y2)}"{common}>{stops}</linearGradient>')
            else:
                self.defs.append(f'<radialGradient id="{pid}" cx="{num(ps.cx)}" cy="{num(ps.cy)}" r="{num(ps.r)}" '
                                 f'fx="{num(ps.fx)}" fy="{num(ps.fy)}" fr="{num(ps.fr)}"{common}>{stops}'
                                 f'</radialGradient>')
        return f' {what}="url(#{pid})"{op}'

    # ---- nodes ----
    def path(self, pid: int, transform=IDENTITY, clip_rule: bool = False) -> str:
        sc = self.sc
        p = sc.paths[pid]
        verbs = sc.verbs[p["verb_offset"]:p["verb_offset"] + p["verb_count"]]
        attrs = f' d="{_path_data(verbs, p["pts"])}"'
        if tuple(float(v) for v in transform) != IDENTITY:
            attrs += f' transform="{_matrix(transform)}"'
        attrs += self.paint(p["fill"], "fill")
        rule = _RULE[p["fill_rule"]]
        attrs += f' clip-rule="{rule}"' if clip_rule else f' fill-rule="{rule}"'
        if p["stroke"].kind != abi.PAINT_NONE:
            attrs += self.paint(p["stroke"], "stroke")
            attrs += (f' stroke-width="{num(p["stroke_width"])}" stroke-linejoin="{_JOIN[p["join"]]}"'
                      f' stroke-linecap="{_CAP[p["cap"]]}" stroke-miterlimit="{num(p["miter"])}"')
            if p["dash_count"]:
                dash = sc.dashes[p["dash_offset"]:p["dash_offset"] + p["dash_count"]]
                attrs += f' stroke-dasharray="{nums(dash)}" stroke-dashoffset="{num(p["dash_phase"])}"'
        if not p["aa"]:
            attrs += ' shape-rendering="crispEdges"'
        if not p["visible"]:
            attrs += ' visibility="hidden"'
        if p["stroke_first"]:
            attrs += ' paint-order="stroke"'
        return f"<path{attrs}/>"

    def image(self, iid: int) -> str:
        im = self.sc.images[iid]
        h, w = im["pixels"].shape[:2]
        data = base64.b64encode(png_bytes(demultiply(im["pixels"]))).decode("ascii")
        vis = "" if im["visible"] else ' visibility="hidden"'
        return (f'<image x="0" y="0" width="{w}" height="{h}" image-rendering="{_QUALITY[im["quality"]]}"{vis} '
                f'xlink:href="data:image/png;base64,{data}"/>')

    def children(self, gid: int, transform) -> str:
        return "".join(self.node(kind, idx) for kind, idx in self.sc.groups[gid]["children"])

    def node(self, kind: int, idx: int) -> str:
        if kind == abi.NODE_PATH:
            return self.path(idx)
        if kind == abi.NODE_IMAGE:
            return self.image(idx)
        return self.group(idx)

    def group(self, gid: int) -> str:
        g = self.sc.groups[gid]
        attrs = ""
        if tuple(float(v) for v in g["transform"]) != IDENTITY:
            attrs += f' transform="{_matrix(g["transform"])}"'
        if float(np.float32(g["opacity"])) != 1.0:
            attrs += f' opacity="{num(g["opacity"])}"'
        if g["blend"] != abi.MIX["normal"]:
            attrs += f' style="mix-blend-mode:{abi.MIX_MODES[g["blend"]].replace("_", "-")}"'
        if g["isolate"]:
            attrs += ' isolation="isolate"'
        if g["filters"]:
            attrs += ' filter="' + " ".join(f"url(#{self.filter(fid)})" for fid in g["filters"]) + '"'
        if g["clip_path"] >= 0:
            attrs += f' clip-path="url(#{self.clip_path(g["clip_path"])})"'
        if g["mask"] >= 0:
            attrs += f' mask="url(#{self.mask(g["mask"])})"'
        return f"<g{attrs}>{self.children(gid, IDENTITY)}</g>"

    # ---- clip paths, masks, content groups ----
    def clip_path(self, cid: int) -> str:
        name = f"clip{cid}"
        if ("clip", cid) in self.done_content:
            return name
        self.done_content.add(("clip", cid))
        c = self.sc.clip_paths[cid]
        attrs = ' clipPathUnits="userSpaceOnUse"'
        if tuple(float(v) for v in c["transform"]) != IDENTITY:
            attrs += f' transform="{_matrix(c["transform"])}"'
        if c["clip_path"] >= 0:
            attrs += f' clip-path="url(#{self.clip_path(c["clip_path"])})"'
        body = self.clip_children(c["root"], IDENTITY)
        self.defs.append(f'<clipPath id="{name}"{attrs}>{body}</clipPath>')
        return name

    def clip_children(self, gid: int, transform) -> str:
        out = []
        for kind, idx in self.sc.groups[gid]["children"]:
            if kind == abi.NODE_PATH:
                out.append(self.path(idx, transform, clip_rule=True))
            elif kind == abi.NODE_GROUP:
                g = self.sc.groups[idx]
                out.append(self.clip_children(idx, _concat(transform, g["transform"])))
        return "".join(out)

    def mask(self, mid: int) -> str:
        name = f"mask{mid}"
        if ("mask", mid) in self.done_content:
            return name
        self.done_content.add(("mask", mid))
        m = self.sc.masks[mid]
        x, y, w, h = m["rect"]
        kind = "alpha" if m["kind"] == abi.MASK_ALPHA else "luminance"
        attrs = (f' maskUnits="userSpaceOnUse" maskContentUnits="userSpaceOnUse" x="{num(x)}" y="{num(y)}" '
                 f'width="{num(w)}" height="{num(h)}" mask-type="{kind}"')
        if m["mask"] >= 0:
            attrs += f' mask="url(#{self.mask(m["mask"])})"'
        self.defs.append(f'<mask id="{name}"{attrs}>{self.children(m["root"], IDENTITY)}</mask>')
        return name

    def content(self, gid: int) -> str:
        """An feImage's element: the content group as a group in defs."""
        name = f"content{gid}"
        if ("content", gid) not in self.done_content:
            self.done_content.add(("content", gid))
            self.defs.append(f'<g id="{name}">{self.children(gid, IDENTITY)}</g>')
        return name

    # ---- filters ----
    def filter(self, fid: int) -> str:
        name = f"filter{fid}"
        if ("filter", fid) in self.done_content:
            return name
        self.done_content.add(("filter", fid))
        sc = self.sc
        rect, off, count = sc.filters[fid]
        prims = "".join(self.primitive(sc.primitives[off + i]) for i in range(count))
        x, y, w, h = rect
        self.defs.append(f'<filter id="{name}" filterUnits="userSpaceOnUse" primitiveUnits="userSpaceOnUse" '
                         f'x="{num(x)}" y="{num(y)}" width="{num(w)}" height="{num(h)}">{prims}</filter>')
        return name

    @staticmethod
    def _input(i: int) -> str:
        if i == abi.INPUT_SOURCE_GRAPHIC:
            return "SourceGraphic"
        if i == abi.INPUT_SOURCE_ALPHA:
            return "SourceAlpha"
        return f"r{i}"

    def primitive(self, p) -> str:
        sc, u = self.sc, p.u
        r = p.rect
        head = (f' x="{num(r.x)}" y="{num(r.y)}" width="{num(r.w)}" height="{num(r.h)}"'
                f' color-interpolation-filters="{"linearRGB" if p.color_interpolation == abi.CI_LINEAR_RGB else "sRGB"}"'
                f' result="r{p.result}"')
        in1 = f' in="{self._input(p.in1)}"'
        in2 = f' in2="{self._input(p.in2)}"'
        k = p.kind
        if k == abi.FE_GAUSSIAN_BLUR:
            return f'<feGaussianBlur{head}{in1} stdDeviation="{num(u.blur.std_dev_x)} {num(u.blur.std_dev_y)}"/>'
        if k == abi.FE_OFFSET:
            return f'<feOffset{head}{in1} dx="{num(u.offset.dx)}" dy="{num(u.offset.dy)}"/>'
        if k == abi.FE_COMPOSITE:
            c = u.composite
            ks = (f' k1="{num(c.k1)}" k2="{num(c.k2)}" k3="{num(c.k3)}" k4="{num(c.k4)}"'
                  if c.op == abi.COMPOSITE_ARITHMETIC else "")
            return f'<feComposite{head}{in1}{in2} operator="{_COMPOSITE[c.op]}"{ks}/>'
        if k == abi.FE_BLEND:
            return f'<feBlend{head}{in1}{in2} mode="{_FE_BLEND[u.blend.mode]}"/>'
        if k == abi.FE_FLOOD:
            return (f'<feFlood{head} flood-color="{_color(u.flood.color)}" '
                    f'flood-opacity="{num(u.flood.opacity)}"/>')
        if k == abi.FE_TILE:
            return f'<feTile{head}{in1}/>'
        if k == abi.FE_MORPHOLOGY:
            m = u.morphology
            op = "erode" if m.op == abi.MORPH_ERODE else "dilate"
            return f'<feMorphology{head}{in1} operator="{op}" radius="{num(m.radius_x)} {num(m.radius_y)}"/>'
        if k == abi.FE_COLOR_MATRIX:
            cm = u.color_matrix
            vals = list(cm.values)[:_CM_COUNT[cm.kind]]
            v = f' values="{nums(vals)}"' if vals else ""
            return f'<feColorMatrix{head}{in1} type="{_CM[cm.kind]}"{v}/>'
        if k == abi.FE_COMPONENT_TRANSFER:
            funcs = []
            for tag, f in zip("RGBA", u.component_transfer.func):
                a = f' type="{_TRANSFER[f.kind]}"'
                if f.kind in (abi.TRANSFER_TABLE, abi.TRANSFER_DISCRETE):
                    a += f' tableValues="{nums(sc.values[f.table_offset:f.table_offset + f.table_len])}"'
                elif f.kind == abi.TRANSFER_LINEAR:
                    a += f' slope="{num(f.a)}" intercept="{num(f.b)}"'
                elif f.kind == abi.TRANSFER_GAMMA:
                    a += f' amplitude="{num(f.a)}" exponent="{num(f.b)}" offset="{num(f.c)}"'
                funcs.append(f"<feFunc{tag}{a}/>")
            return f'<feComponentTransfer{head}{in1}>{"".join(funcs)}</feComponentTransfer>'
        if k == abi.FE_CONVOLVE_MATRIX:
            c = u.convolve
            kern = sc.values[c.data_offset:c.data_offset + c.columns * c.rows]
            return (f'<feConvolveMatrix{head}{in1} order="{c.columns} {c.rows}" kernelMatrix="{nums(kern)}" '
                    f'divisor="{num(c.divisor)}" bias="{num(c.bias)}" targetX="{c.target_x}" targetY="{c.target_y}" '
                    f'edgeMode="{_EDGE[c.edge_mode]}" preserveAlpha="{"true" if c.preserve_alpha else "false"}"/>')
        if k == abi.FE_TURBULENCE:
            t = u.turbulence
            return (f'<feTurbulence{head} baseFrequency="{num(t.base_frequency_x)} {num(t.base_frequency_y)}" '
                    f'numOctaves="{t.num_octaves}" seed="{t.seed}" '
                    f'stitchTiles="{"stitch" if t.stitch_tiles else "noStitch"}" '
                    f'type="{"fractalNoise" if t.fractal_noise else "turbulence"}"/>')
        if k in (abi.FE_DIFFUSE_LIGHTING, abi.FE_SPECULAR_LIGHTING):
            lg = u.lighting
            L = lg.light
            if L.kind == abi.LIGHT_DISTANT:
                light = f'<feDistantLight azimuth="{num(L.azimuth)}" elevation="{num(L.elevation)}"/>'
            elif L.kind == abi.LIGHT_POINT:
                light = f'<fePointLight x="{num(L.x)}" y="{num(L.y)}" z="{num(L.z)}"/>'
            else:
                cone = f' limitingConeAngle="{num(L.limiting_cone_angle)}"' if L.has_limiting_cone_angle else ""
                light = (f'<feSpotLight x="{num(L.x)}" y="{num(L.y)}" z="{num(L.z)}" pointsAtX="{num(L.points_at_x)}" '
                         f'pointsAtY="{num(L.points_at_y)}" pointsAtZ="{num(L.points_at_z)}" '
                         f'specularExponent="{num(L.specular_exponent)}"{cone}/>')
            color = f' lighting-color="{_color(lg.color)}" surfaceScale="{num(lg.surface_scale)}"'
            if k == abi.FE_DIFFUSE_LIGHTING:
                return f'<feDiffuseLighting{head}{in1}{color} diffuseConstant="{num(lg.constant)}">{light}</feDiffuseLighting>'
            return (f'<feSpecularLighting{head}{in1}{color} specularConstant="{num(lg.constant)}" '
                    f'specularExponent="{num(lg.specular_exponent)}">{light}</feSpecularLighting>')
        if k == abi.FE_DISPLACEMENT_MAP:
            d = u.displacement
            return (f'<feDisplacementMap{head}{in1}{in2} scale="{num(d.scale)}" '
                    f'xChannelSelector="{_CHANNEL[d.x_channel]}" yChannelSelector="{_CHANNEL[d.y_channel]}"/>')
        if k == abi.FE_DROP_SHADOW:
            d = u.drop_shadow
            return (f'<feDropShadow{head}{in1} dx="{num(d.dx)}" dy="{num(d.dy)}" '
                    f'stdDeviation="{num(d.std_dev_x)} {num(d.std_dev_y)}" flood-color="{_color(d.color)}" '
                    f'flood-opacity="{num(d.opacity)}"/>')
        if k == abi.FE_MERGE:
            m = u.merge
            nodes = "".join(f'<feMergeNode in="{self._input(sc.inputs[m.input_offset + i])}"/>'
                            for i in range(m.input_count))
            return f"<feMerge{head}>{nodes}</feMerge>"
        if k == abi.FE_IMAGE:
            return f'<feImage{head} xlink:href="#{self.content(u.image.root)}"/>'
        raise ValueError(f"unknown filter primitive kind {k}")


def scene_to_svg(sc) -> str:
    """The scene as an SVG document of its own size (viewBox = canvas)."""
    w = _Writer(sc)
    body = w.children(sc.root, IDENTITY)
    defs = f"<defs>{''.join(w.defs)}</defs>" if w.defs else ""
    return (f'<svg xmlns="http://www.w3.org/2000/svg" xmlns:xlink="http://www.w3.org/1999/xlink" '
            f'width="{num(sc.width)}" height="{num(sc.height)}" viewBox="0 0 {num(sc.width)} {num(sc.height)}">'
            f"{defs}{body}</svg>\n")
