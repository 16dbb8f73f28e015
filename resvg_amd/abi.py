"""ctypes mirror of include/resvg_amd.h (structure layouts and constants).

Shared by the product bindings (``resvg_amd``) and by the test-side oracle
bindings, which take the same structures.
"""
from __future__ import annotations

import ctypes as C

# status
OK, ERR_INVALID_ARGUMENT, ERR_OUT_OF_MEMORY, ERR_HIP, ERR_INVALID_REGION, ERR_UNSUPPORTED, ERR_NO_DEVICE, \
    ERR_TOO_LARGE = range(8)

# tiny_skia::BlendMode (declaration order)
BLEND_MODES = [
    "clear", "source", "destination", "source_over", "destination_over", "source_in", "destination_in",
    "source_out", "destination_out", "source_atop", "destination_atop", "xor", "plus", "modulate", "screen",
    "overlay", "darken", "lighten", "color_dodge", "color_burn", "hard_light", "soft_light", "difference",
    "exclusion", "multiply", "hue", "saturation", "color", "luminosity",
]
BLEND = {n: i for i, n in enumerate(BLEND_MODES)}
# usvg::BlendMode
MIX_MODES = [
    "normal", "multiply", "screen", "overlay", "darken", "lighten", "color_dodge", "color_burn", "hard_light",
    "soft_light", "difference", "exclusion", "hue", "saturation", "color", "luminosity",
]
MIX = {n: i for i, n in enumerate(MIX_MODES)}

CI_SRGB, CI_LINEAR_RGB = 0, 1
INPUT_SOURCE_GRAPHIC, INPUT_SOURCE_ALPHA = -1, -2

(FE_BLEND, FE_COLOR_MATRIX, FE_COMPONENT_TRANSFER, FE_COMPOSITE, FE_CONVOLVE_MATRIX, FE_DIFFUSE_LIGHTING,
 FE_DISPLACEMENT_MAP, FE_DROP_SHADOW, FE_FLOOD, FE_GAUSSIAN_BLUR, FE_IMAGE, FE_MERGE, FE_MORPHOLOGY, FE_OFFSET,
 FE_SPECULAR_LIGHTING, FE_TILE, FE_TURBULENCE) = range(17)

COMPOSITE_OVER, COMPOSITE_IN, COMPOSITE_OUT, COMPOSITE_ATOP, COMPOSITE_XOR, COMPOSITE_ARITHMETIC = range(6)
COLOR_MATRIX_MATRIX, COLOR_MATRIX_SATURATE, COLOR_MATRIX_HUE_ROTATE, COLOR_MATRIX_LUMINANCE_TO_ALPHA = range(4)
TRANSFER_IDENTITY, TRANSFER_TABLE, TRANSFER_DISCRETE, TRANSFER_LINEAR, TRANSFER_GAMMA = range(5)
EDGE_NONE, EDGE_DUPLICATE, EDGE_WRAP = range(3)
MORPH_ERODE, MORPH_DILATE = range(2)
CHANNEL_R, CHANNEL_G, CHANNEL_B, CHANNEL_A = range(4)
LIGHT_DISTANT, LIGHT_POINT, LIGHT_SPOT = range(3)

VERB_MOVE, VERB_LINE, VERB_QUAD, VERB_CUBIC, VERB_CLOSE = range(5)
PAINT_NONE, PAINT_COLOR, PAINT_LINEAR, PAINT_RADIAL, PAINT_PATTERN = range(5)
SPREAD_PAD, SPREAD_REFLECT, SPREAD_REPEAT = range(3)
FILL_NONZERO, FILL_EVENODD = range(2)
CAP_BUTT, CAP_ROUND, CAP_SQUARE = range(3)
JOIN_MITER, JOIN_MITER_CLIP, JOIN_ROUND, JOIN_BEVEL = range(4)
NODE_GROUP, NODE_PATH, NODE_IMAGE = range(3)
QUALITY_NEAREST, QUALITY_BILINEAR, QUALITY_BICUBIC = range(3)


class Transform(C.Structure):
    _fields_ = [(n, C.c_float) for n in "abcdef"]

    @staticmethod
    def identity() -> "Transform":
        return Transform(1, 0, 0, 1, 0, 0)

    @staticmethod
    def scale(sx: float, sy: float) -> "Transform":
        return Transform(sx, 0, 0, sy, 0, 0)


class Rect(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("w", C.c_float), ("h", C.c_float)]


class TransferFn(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("a", C.c_float), ("b", C.c_float), ("c", C.c_float),
                ("table_offset", C.c_uint32), ("table_len", C.c_uint32)]


class LightSource(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("azimuth", C.c_float), ("elevation", C.c_float), ("x", C.c_float),
                ("y", C.c_float), ("z", C.c_float), ("points_at_x", C.c_float), ("points_at_y", C.c_float),
                ("points_at_z", C.c_float), ("specular_exponent", C.c_float),
                ("has_limiting_cone_angle", C.c_uint32), ("limiting_cone_angle", C.c_float)]


class _Blend(C.Structure):
    _fields_ = [("mode", C.c_uint32)]


class _ColorMatrix(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("values", C.c_float * 20)]


class _ComponentTransfer(C.Structure):
    _fields_ = [("func", TransferFn * 4)]


class _Composite(C.Structure):
    _fields_ = [("op", C.c_uint32), ("k1", C.c_float), ("k2", C.c_float), ("k3", C.c_float), ("k4", C.c_float)]


class _Convolve(C.Structure):
    _fields_ = [("columns", C.c_uint32), ("rows", C.c_uint32), ("target_x", C.c_uint32), ("target_y", C.c_uint32),
                ("divisor", C.c_float), ("bias", C.c_float), ("edge_mode", C.c_uint32),
                ("preserve_alpha", C.c_uint32), ("data_offset", C.c_uint32)]


class _Lighting(C.Structure):
    _fields_ = [("surface_scale", C.c_float), ("constant", C.c_float), ("specular_exponent", C.c_float),
                ("color", C.c_uint8 * 4), ("light", LightSource)]


class _Displacement(C.Structure):
    _fields_ = [("scale", C.c_float), ("x_channel", C.c_uint32), ("y_channel", C.c_uint32)]


class _DropShadow(C.Structure):
    _fields_ = [("dx", C.c_float), ("dy", C.c_float), ("std_dev_x", C.c_float), ("std_dev_y", C.c_float),
                ("color", C.c_uint8 * 4), ("opacity", C.c_float)]


class _Flood(C.Structure):
    _fields_ = [("color", C.c_uint8 * 4), ("opacity", C.c_float)]


class _Blur(C.Structure):
    _fields_ = [("std_dev_x", C.c_float), ("std_dev_y", C.c_float)]


class _Merge(C.Structure):
    _fields_ = [("input_offset", C.c_uint32), ("input_count", C.c_uint32)]


class _Morphology(C.Structure):
    _fields_ = [("op", C.c_uint32), ("radius_x", C.c_float), ("radius_y", C.c_float)]


class _Offset(C.Structure):
    _fields_ = [("dx", C.c_float), ("dy", C.c_float)]


class _Turbulence(C.Structure):
    _fields_ = [("base_frequency_x", C.c_float), ("base_frequency_y", C.c_float), ("num_octaves", C.c_uint32),
                ("seed", C.c_int32), ("stitch_tiles", C.c_uint32), ("fractal_noise", C.c_uint32)]


class _Image(C.Structure):
    _fields_ = [("root", C.c_uint32)]


class _PrimU(C.Union):
    _fields_ = [("blend", _Blend), ("color_matrix", _ColorMatrix), ("component_transfer", _ComponentTransfer),
                ("composite", _Composite), ("convolve", _Convolve), ("lighting", _Lighting),
                ("displacement", _Displacement), ("drop_shadow", _DropShadow), ("flood", _Flood), ("blur", _Blur),
                ("merge", _Merge), ("morphology", _Morphology), ("offset", _Offset), ("turbulence", _Turbulence),
                ("image", _Image), ("pad_", C.c_float * 48)]


class Primitive(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("color_interpolation", C.c_uint32), ("rect", Rect), ("in1", C.c_int32),
                ("in2", C.c_int32), ("result", C.c_int32), ("u", _PrimU)]


class Filter(C.Structure):
    _fields_ = [("rect", Rect), ("primitive_offset", C.c_uint32), ("primitive_count", C.c_uint32)]


class FilterData(C.Structure):
    _fields_ = [("primitives", C.POINTER(Primitive)), ("values", C.POINTER(C.c_float)),
                ("inputs", C.POINTER(C.c_int32))]


class Stop(C.Structure):
    _fields_ = [("offset", C.c_float), ("color", C.c_uint8 * 4), ("opacity", C.c_float)]


class Paint(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("color", C.c_uint8 * 4), ("opacity", C.c_float), ("spread", C.c_uint32),
                ("x1", C.c_float), ("y1", C.c_float), ("x2", C.c_float), ("y2", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("r", C.c_float), ("fx", C.c_float), ("fy", C.c_float),
                ("fr", C.c_float), ("transform", Transform), ("stop_offset", C.c_uint32),
                ("stop_count", C.c_uint32), ("rect", Rect), ("root", C.c_uint32)]


class Path(C.Structure):
    _fields_ = [("verb_offset", C.c_uint32), ("verb_count", C.c_uint32), ("point_offset", C.c_uint32),
                ("visible", C.c_uint32), ("anti_alias", C.c_uint32), ("paint_order_stroke_first", C.c_uint32),
                ("fill", Paint), ("fill_rule", C.c_uint32), ("stroke", Paint), ("stroke_width", C.c_float),
                ("stroke_miter_limit", C.c_float), ("stroke_cap", C.c_uint32), ("stroke_join", C.c_uint32),
                ("bounds", Rect), ("dash_offset", C.c_uint32), ("dash_count", C.c_uint32), ("dash_phase", C.c_float)]


class Group(C.Structure):
    _fields_ = [("transform", Transform), ("opacity", C.c_float), ("blend_mode", C.c_uint32),
                ("isolate", C.c_uint32), ("clip_path", C.c_int32), ("mask", C.c_int32),
                ("filter_offset", C.c_uint32), ("filter_count", C.c_uint32), ("layer_bbox", Rect),
                ("child_offset", C.c_uint32), ("child_count", C.c_uint32)]


class NodeRef(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("index", C.c_uint32)]


MASK_LUMINANCE, MASK_ALPHA = 0, 1


class ClipPath(C.Structure):
    _fields_ = [("transform", Transform), ("clip_path", C.c_int32), ("root", C.c_uint32)]


class Mask(C.Structure):
    _fields_ = [("rect", Rect), ("kind", C.c_uint32), ("mask", C.c_int32), ("root", C.c_uint32)]


class Image(C.Structure):
    _fields_ = [("pixels", C.POINTER(C.c_uint8)), ("width", C.c_uint32), ("height", C.c_uint32),
                ("quality", C.c_uint32), ("visible", C.c_uint32)]


class Scene(C.Structure):
    _fields_ = [("groups", C.POINTER(Group)), ("group_count", C.c_uint32), ("paths", C.POINTER(Path)),
                ("path_count", C.c_uint32), ("children", C.POINTER(NodeRef)), ("verbs", C.POINTER(C.c_uint8)),
                ("points", C.POINTER(C.c_float)), ("stops", C.POINTER(Stop)), ("filters", C.POINTER(Filter)),
                ("filter_data", FilterData), ("clip_paths", C.POINTER(ClipPath)), ("clip_path_count", C.c_uint32),
                ("masks", C.POINTER(Mask)), ("mask_count", C.c_uint32), ("dashes", C.POINTER(C.c_float)),
                ("images", C.POINTER(Image)), ("image_count", C.c_uint32)]


class PathGeometry(C.Structure):
    _fields_ = [("verbs", C.POINTER(C.c_uint8)), ("verb_count", C.c_uint32), ("points", C.POINTER(C.c_float)),
                ("point_count", C.c_uint32)]


class StrokeStyle(C.Structure):
    _fields_ = [("width", C.c_float), ("miter_limit", C.c_float), ("cap", C.c_uint32), ("join", C.c_uint32),
                ("dashes", C.POINTER(C.c_float)), ("dash_count", C.c_uint32), ("dash_offset", C.c_float)]


MASK_TYPE_ALPHA, MASK_TYPE_LUMINANCE = 0, 1


class Stats(C.Structure):
    _fields_ = [("kernel_launches", C.c_uint64), ("layers", C.c_uint64), ("paths", C.c_uint64),
                ("filters", C.c_uint64), ("scratch_bytes", C.c_uint64), ("edge_pieces", C.c_uint64),
                ("band_entries", C.c_uint64), ("upload_bytes", C.c_uint64), ("flushes", C.c_uint64),
                ("host_build_ms", C.c_double), ("host_wait_ms", C.c_double),
                ("host_geom_ms", C.c_double), ("validated_flushes", C.c_uint64)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("launches", C.c_uint64), ("total_ms", C.c_double), ("bytes", C.c_double)]


def array(ctype, items):
    """A ctypes array from a sequence (keeps an explicit reference to the buffer)."""
    items = list(items)
    arr = (ctype * max(1, len(items)))()
    for i, v in enumerate(items):
        arr[i] = v
    return arr
