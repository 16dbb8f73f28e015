"""Builder for flattened filter graphs (a usvg::filter::Filter as rvk_filter +
rvk_filter_data), mirroring the primitive kinds of crates/usvg/src/tree/filter.rs.

Each ``add_*`` method appends one primitive and returns its result id, which later
primitives pass as ``in1``/``in2`` (usvg's ``Input::Reference``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import abi
from .abi import INPUT_SOURCE_ALPHA, INPUT_SOURCE_GRAPHIC  # noqa: F401  (re-exported)

SOURCE_GRAPHIC = INPUT_SOURCE_GRAPHIC
SOURCE_ALPHA = INPUT_SOURCE_ALPHA


@dataclass
class FilterGraph:
    """One filter with its primitives; ``rect`` is the filter region in user space."""

    rect: tuple[float, float, float, float]
    primitives: list = field(default_factory=list)
    values: list = field(default_factory=list)
    inputs: list = field(default_factory=list)
    _next_result: int = 0

    # -- generic -----------------------------------------------------------
    def add(self, kind: int, *, rect=None, cs: int = abi.CI_LINEAR_RGB, in1: int = SOURCE_GRAPHIC,
            in2: int = SOURCE_GRAPHIC, result: int | None = None) -> tuple[int, abi.Primitive]:
        p = abi.Primitive()
        p.kind = kind
        p.color_interpolation = cs
        p.rect = abi.Rect(*(rect if rect is not None else self.rect))
        p.in1 = in1
        p.in2 = in2
        if result is None:
            result = self._next_result
        self._next_result = max(self._next_result, result + 1)
        p.result = result
        self.primitives.append(p)
        return result, p

    # -- primitives ----------------------------------------------------------
    def gaussian_blur(self, sx: float, sy: float | None = None, **kw) -> int:
        r, p = self.add(abi.FE_GAUSSIAN_BLUR, **kw)
        p.u.blur.std_dev_x = sx
        p.u.blur.std_dev_y = sx if sy is None else sy
        return r

    def offset(self, dx: float, dy: float, **kw) -> int:
        r, p = self.add(abi.FE_OFFSET, **kw)
        p.u.offset.dx, p.u.offset.dy = dx, dy
        return r

    def composite(self, op: int, k=(0.0, 0.0, 0.0, 0.0), **kw) -> int:
        r, p = self.add(abi.FE_COMPOSITE, **kw)
        p.u.composite.op = op
        p.u.composite.k1, p.u.composite.k2, p.u.composite.k3, p.u.composite.k4 = k
        return r

    def blend(self, mode: str | int, **kw) -> int:
        r, p = self.add(abi.FE_BLEND, **kw)
        p.u.blend.mode = abi.BLEND[mode] if isinstance(mode, str) else mode
        return r

    def flood(self, color=(0, 0, 0), opacity: float = 1.0, **kw) -> int:
        r, p = self.add(abi.FE_FLOOD, **kw)
        for i in range(3):
            p.u.flood.color[i] = color[i]
        p.u.flood.opacity = opacity
        return r

    def tile(self, **kw) -> int:
        """feTile: repeats ``in1``'s subregion over the filter region."""
        r, _ = self.add(abi.FE_TILE, **kw)
        return r

    def morphology(self, op: int, rx: float, ry: float | None = None, **kw) -> int:
        r, p = self.add(abi.FE_MORPHOLOGY, **kw)
        p.u.morphology.op = op
        p.u.morphology.radius_x = rx
        p.u.morphology.radius_y = rx if ry is None else ry
        return r

    def color_matrix(self, kind: int, values=(), **kw) -> int:
        r, p = self.add(abi.FE_COLOR_MATRIX, **kw)
        p.u.color_matrix.kind = kind
        for i, v in enumerate(values):
            p.u.color_matrix.values[i] = v
        return r

    def component_transfer(self, funcs, **kw) -> int:
        """funcs: 4 entries (r, g, b, a), each None/("identity",), ("table", [..]),
        ("discrete", [..]), ("linear", slope, intercept), ("gamma", amp, exp, off)."""
        r, p = self.add(abi.FE_COMPONENT_TRANSFER, **kw)
        for i, f in enumerate(funcs):
            tf = p.u.component_transfer.func[i]
            if f is None or f[0] == "identity":
                tf.kind = abi.TRANSFER_IDENTITY
            elif f[0] in ("table", "discrete"):
                tf.kind = abi.TRANSFER_TABLE if f[0] == "table" else abi.TRANSFER_DISCRETE
                tf.table_offset = len(self.values)
                tf.table_len = len(f[1])
                self.values.extend(float(v) for v in f[1])
            elif f[0] == "linear":
                tf.kind, tf.a, tf.b = abi.TRANSFER_LINEAR, f[1], f[2]
            elif f[0] == "gamma":
                tf.kind, tf.a, tf.b, tf.c = abi.TRANSFER_GAMMA, f[1], f[2], f[3]
            else:
                raise ValueError(f"unknown transfer function {f[0]!r}")
        return r

    def convolve_matrix(self, columns: int, rows: int, kernel, divisor: float | None = None, bias: float = 0.0,
                        target_x: int | None = None, target_y: int | None = None, edge_mode: int = abi.EDGE_DUPLICATE,
                        preserve_alpha: bool = False, **kw) -> int:
        r, p = self.add(abi.FE_CONVOLVE_MATRIX, **kw)
        kernel = [float(v) for v in kernel]
        assert len(kernel) == columns * rows
        c = p.u.convolve
        c.columns, c.rows = columns, rows
        c.target_x = columns // 2 if target_x is None else target_x
        c.target_y = rows // 2 if target_y is None else target_y
        if divisor is None:  # usvg default: the kernel sum, or 1 when it is 0 (parser/filter.rs:678-688)
            s = sum(kernel)
            divisor = 1.0 if s == 0.0 else s
        c.divisor, c.bias = divisor, bias
        c.edge_mode = edge_mode
        c.preserve_alpha = 1 if preserve_alpha else 0
        c.data_offset = len(self.values)
        self.values.extend(kernel)
        return r

    def turbulence(self, base_frequency_x: float, base_frequency_y: float, num_octaves: int = 1, seed: int = 0,
                   stitch_tiles: bool = False, fractal_noise: bool = False, **kw) -> int:
        r, p = self.add(abi.FE_TURBULENCE, **kw)
        t = p.u.turbulence
        t.base_frequency_x, t.base_frequency_y = base_frequency_x, base_frequency_y
        t.num_octaves, t.seed = num_octaves, seed
        t.stitch_tiles, t.fractal_noise = int(stitch_tiles), int(fractal_noise)
        return r

    def image(self, root: int, **kw) -> int:
        """feImage referencing an element (usvg filter::Image): the children of scene
        group ``root`` (a content group, see Scene.content_group) drawn into the
        primitive's pixmap with the layer's scale, moved to the subregion's origin."""
        r, p = self.add(abi.FE_IMAGE, **kw)
        p.u.image.root = root
        return r

    def lighting(self, specular: bool, light: abi.LightSource, surface_scale: float = 1.0, constant: float = 1.0,
                 specular_exponent: float = 1.0, color=(255, 255, 255), **kw) -> int:
        r, p = self.add(abi.FE_SPECULAR_LIGHTING if specular else abi.FE_DIFFUSE_LIGHTING, **kw)
        lg = p.u.lighting
        lg.surface_scale, lg.constant, lg.specular_exponent = surface_scale, constant, specular_exponent
        for i in range(3):
            lg.color[i] = color[i]
        lg.light = light
        return r

    def displacement_map(self, scale: float, x_channel: int = abi.CHANNEL_A, y_channel: int = abi.CHANNEL_A,
                         **kw) -> int:
        r, p = self.add(abi.FE_DISPLACEMENT_MAP, **kw)
        p.u.displacement.scale, p.u.displacement.x_channel, p.u.displacement.y_channel = scale, x_channel, y_channel
        return r

    def drop_shadow(self, dx: float, dy: float, sx: float, sy: float, color=(0, 0, 0), opacity: float = 1.0,
                    **kw) -> int:
        r, p = self.add(abi.FE_DROP_SHADOW, **kw)
        d = p.u.drop_shadow
        d.dx, d.dy, d.std_dev_x, d.std_dev_y = dx, dy, sx, sy
        for i in range(3):
            d.color[i] = color[i]
        d.opacity = opacity
        return r

    def merge(self, inputs, **kw) -> int:
        r, p = self.add(abi.FE_MERGE, **kw)
        p.u.merge.input_offset = len(self.inputs)
        p.u.merge.input_count = len(inputs)
        self.inputs.extend(inputs)
        return r

    # -- ABI --------------------------------------------------------------------
    def to_abi(self, base_primitive: int = 0):
        """Returns (rvk_filter, rvk_filter_data, keepalive) for a standalone filter."""
        prims = abi.array(abi.Primitive, self.primitives)
        vals = abi.array(C.c_float, self.values)
        ins = abi.array(C.c_int32, self.inputs)
        f = abi.Filter(abi.Rect(*self.rect), base_primitive, len(self.primitives))
        d = abi.FilterData(C.cast(prims, C.POINTER(abi.Primitive)), C.cast(vals, C.POINTER(C.c_float)),
                           C.cast(ins, C.POINTER(C.c_int32)))
        return f, d, (prims, vals, ins)


def distant_light(azimuth: float, elevation: float) -> abi.LightSource:
    return abi.LightSource(kind=abi.LIGHT_DISTANT, azimuth=azimuth, elevation=elevation)


def point_light(x: float, y: float, z: float) -> abi.LightSource:
    return abi.LightSource(kind=abi.LIGHT_POINT, x=x, y=y, z=z)


def spot_light(x, y, z, points_at_x, points_at_y, points_at_z, specular_exponent=1.0, cone=None) -> abi.LightSource:
    return abi.LightSource(kind=abi.LIGHT_SPOT, x=x, y=y, z=z, points_at_x=points_at_x, points_at_y=points_at_y,
                           points_at_z=points_at_z, specular_exponent=specular_exponent,
                           has_limiting_cone_angle=0 if cone is None else 1,
                           limiting_cone_angle=0.0 if cone is None else cone)
