/*
 * resvg_amd.h — C-ABI of the MI355X (gfx950) pixel pipeline that sits below
 * resvg::render / resvg_render.
 *
 * Every entry point returns an rvk_status and never aborts.  Pixel buffers are
 * premultiplied RGBA8, row-major, stride = 4 * width, no padding — the layout of
 * tiny_skia::Pixmap and of the `char *pixmap` argument of resvg_render
 * (crates/c-api/resvg.h:476-489).  Buffers named `dev_*` are device pointers
 * (hipMalloc'd, or a torch CUDA tensor's data_ptr); buffers named `host_*` are
 * host pointers.  Work is enqueued on the context's stream; a call returns once
 * the work is enqueued unless it says otherwise.
 *
 * Reference interfaces replaced (file:line relative to the resvg 0.48.1 tree):
 *   resvg::render / resvg_render      crates/resvg/src/lib.rs:34-43, crates/c-api/lib.rs:874-897
 *   resvg::render_node                crates/resvg/src/lib.rs:55-70
 *   filter::apply                     crates/resvg/src/filter/mod.rs:335-355
 *   per-primitive filter functions    crates/resvg/src/filter/<primitive>.rs (cited per entry)
 *   tiny-skia 0.12 pixel calls        fill_path / stroke_path / draw_pixmap / fill /
 *                                     fill_rect / Mask::fill_path / Mask::from_pixmap /
 *                                     Mask::invert / apply_mask
 *                                     (call sites: crates/resvg/src/path.rs:73,113,
 *                                      render.rs:127-140, clip.rs:11-27, mask.rs:13-45,
 *                                      filter/mod.rs:457-508, 624-1111)
 */
#ifndef RESVG_AMD_H
#define RESVG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes                                                              */
/* ------------------------------------------------------------------------- */
typedef int32_t rvk_status;
#define RVK_OK                   0
#define RVK_ERR_INVALID_ARGUMENT 1  /* null pointer, zero/oversized dimension, bad enum */
#define RVK_ERR_OUT_OF_MEMORY    2  /* device allocation failed */
#define RVK_ERR_HIP              3  /* a HIP runtime call failed */
#define RVK_ERR_INVALID_REGION   4  /* filter::Error::InvalidRegion (filter/mod.rs:86-90) */
#define RVK_ERR_UNSUPPORTED      5  /* feature outside this library's scope (nested SVG images) */
#define RVK_ERR_NO_DEVICE        6  /* no HIP device / library built without a device */
#define RVK_ERR_TOO_LARGE        7  /* input beyond a fixed limit of the entry point (the no-allocation form) */

/* Returns a static string naming a status code. */
const char *rvk_status_string(rvk_status status);

/* ------------------------------------------------------------------------- */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------- */
/* Same memory layout and meaning as resvg_transform (crates/c-api/resvg.h:101-108):
 * tiny_skia::Transform::from_row(a, b, c, d, e, f) = (sx, ky, kx, sy, tx, ty). */
typedef struct rvk_transform {
    float a, b, c, d, e, f;
} rvk_transform;

/* usvg::NonZeroRect / tiny_skia::Rect as x, y, width, height (f32). */
typedef struct rvk_rect {
    float x, y, w, h;
} rvk_rect;

/* ------------------------------------------------------------------------- */
/* Enumerations (values are this ABI's; the mapping follows the reference)   */
/* ------------------------------------------------------------------------- */
/* tiny_skia::BlendMode, in tiny-skia's declaration order. */
enum {
    RVK_BLEND_CLEAR = 0, RVK_BLEND_SOURCE, RVK_BLEND_DESTINATION, RVK_BLEND_SOURCE_OVER,
    RVK_BLEND_DESTINATION_OVER, RVK_BLEND_SOURCE_IN, RVK_BLEND_DESTINATION_IN,
    RVK_BLEND_SOURCE_OUT, RVK_BLEND_DESTINATION_OUT, RVK_BLEND_SOURCE_ATOP,
    RVK_BLEND_DESTINATION_ATOP, RVK_BLEND_XOR, RVK_BLEND_PLUS, RVK_BLEND_MODULATE,
    RVK_BLEND_SCREEN, RVK_BLEND_OVERLAY, RVK_BLEND_DARKEN, RVK_BLEND_LIGHTEN,
    RVK_BLEND_COLOR_DODGE, RVK_BLEND_COLOR_BURN, RVK_BLEND_HARD_LIGHT, RVK_BLEND_SOFT_LIGHT,
    RVK_BLEND_DIFFERENCE, RVK_BLEND_EXCLUSION, RVK_BLEND_MULTIPLY, RVK_BLEND_HUE,
    RVK_BLEND_SATURATION, RVK_BLEND_COLOR, RVK_BLEND_LUMINOSITY, RVK_BLEND_COUNT
};

/* usvg::BlendMode (crates/usvg/src/tree/mod.rs:216-233); mapped to tiny-skia
 * modes by convert_blend_mode (crates/resvg/src/render.rs:145-164). */
enum {
    RVK_MIX_NORMAL = 0, RVK_MIX_MULTIPLY, RVK_MIX_SCREEN, RVK_MIX_OVERLAY, RVK_MIX_DARKEN,
    RVK_MIX_LIGHTEN, RVK_MIX_COLOR_DODGE, RVK_MIX_COLOR_BURN, RVK_MIX_HARD_LIGHT,
    RVK_MIX_SOFT_LIGHT, RVK_MIX_DIFFERENCE, RVK_MIX_EXCLUSION, RVK_MIX_HUE,
    RVK_MIX_SATURATION, RVK_MIX_COLOR, RVK_MIX_LUMINOSITY, RVK_MIX_COUNT
};

/* usvg::filter::ColorInterpolation (crates/usvg/src/tree/filter.rs:141-146). */
enum { RVK_CI_SRGB = 0, RVK_CI_LINEAR_RGB = 1 };

/* usvg::filter::Input (filter.rs:130-134).  A non-negative input value is a
 * result name id: Reference(name) resolves to the latest earlier primitive whose
 * `result` equals that id (filter/mod.rs:554-562). */
#define RVK_INPUT_SOURCE_GRAPHIC (-1)
#define RVK_INPUT_SOURCE_ALPHA   (-2)

/* usvg::filter::Kind (filter.rs:82-100). */
enum {
    RVK_FE_BLEND = 0, RVK_FE_COLOR_MATRIX, RVK_FE_COMPONENT_TRANSFER, RVK_FE_COMPOSITE,
    RVK_FE_CONVOLVE_MATRIX, RVK_FE_DIFFUSE_LIGHTING, RVK_FE_DISPLACEMENT_MAP,
    RVK_FE_DROP_SHADOW, RVK_FE_FLOOD, RVK_FE_GAUSSIAN_BLUR, RVK_FE_IMAGE, RVK_FE_MERGE,
    RVK_FE_MORPHOLOGY, RVK_FE_OFFSET, RVK_FE_SPECULAR_LIGHTING, RVK_FE_TILE,
    RVK_FE_TURBULENCE, RVK_FE_KIND_COUNT
};

/* CompositeOperator (filter.rs:332-338). */
enum { RVK_COMPOSITE_OVER = 0, RVK_COMPOSITE_IN, RVK_COMPOSITE_OUT, RVK_COMPOSITE_ATOP,
       RVK_COMPOSITE_XOR, RVK_COMPOSITE_ARITHMETIC };
/* ColorMatrixKind (filter.rs:208-213). */
enum { RVK_COLOR_MATRIX_MATRIX = 0, RVK_COLOR_MATRIX_SATURATE, RVK_COLOR_MATRIX_HUE_ROTATE,
       RVK_COLOR_MATRIX_LUMINANCE_TO_ALPHA };
/* TransferFunction (filter.rs:269-296). */
enum { RVK_TRANSFER_IDENTITY = 0, RVK_TRANSFER_TABLE, RVK_TRANSFER_DISCRETE,
       RVK_TRANSFER_LINEAR, RVK_TRANSFER_GAMMA };
/* EdgeMode (filter.rs:484-488). */
enum { RVK_EDGE_NONE = 0, RVK_EDGE_DUPLICATE, RVK_EDGE_WRAP };
/* MorphologyOperator (filter.rs:947-950). */
enum { RVK_MORPH_ERODE = 0, RVK_MORPH_DILATE };
/* ColorChannel (filter.rs:542-547). */
enum { RVK_CHANNEL_R = 0, RVK_CHANNEL_G, RVK_CHANNEL_B, RVK_CHANNEL_A };
/* LightSource (filter.rs:792-796). */
enum { RVK_LIGHT_DISTANT = 0, RVK_LIGHT_POINT, RVK_LIGHT_SPOT };

/* ------------------------------------------------------------------------- */
/* Filter description: a flattened usvg::filter::Filter                      */
/* ------------------------------------------------------------------------- */
typedef struct rvk_transfer_fn {          /* TransferFunction */
    uint32_t kind;                         /* RVK_TRANSFER_* */
    float a, b, c;                         /* linear: slope, intercept; gamma: amplitude, exponent, offset */
    uint32_t table_offset, table_len;      /* table/discrete values in rvk_filter_data.values */
} rvk_transfer_fn;

typedef struct rvk_light_source {          /* LightSource (filter.rs:792-884) */
    uint32_t kind;                         /* RVK_LIGHT_* */
    float azimuth, elevation;              /* distant */
    float x, y, z;                         /* point, spot */
    float points_at_x, points_at_y, points_at_z; /* spot */
    float specular_exponent;               /* spot (PositiveF32) */
    uint32_t has_limiting_cone_angle;
    float limiting_cone_angle;
} rvk_light_source;

typedef struct rvk_primitive {
    uint32_t kind;                         /* RVK_FE_* */
    uint32_t color_interpolation;          /* RVK_CI_* */
    rvk_rect rect;                         /* primitive subregion, user space (Primitive::rect) */
    int32_t in1, in2;                      /* RVK_INPUT_* or a result id */
    int32_t result;                        /* result name id (any value; compared for equality) */
    union {
        struct { uint32_t mode; } blend;                            /* RVK_BLEND_* (tiny-skia mode) */
        struct { uint32_t kind; float values[20]; } color_matrix;   /* saturate/hue: values[0] */
        struct { rvk_transfer_fn func[4]; } component_transfer;     /* r, g, b, a */
        struct { uint32_t op; float k1, k2, k3, k4; } composite;
        struct { uint32_t columns, rows, target_x, target_y;
                 float divisor, bias; uint32_t edge_mode, preserve_alpha;
                 uint32_t data_offset; } convolve;                  /* row-major columns*rows in values */
        struct { float surface_scale, constant, specular_exponent;
                 uint8_t color[4]; rvk_light_source light; } lighting; /* constant = diffuse or specular */
        struct { float scale; uint32_t x_channel, y_channel; } displacement;
        struct { float dx, dy, std_dev_x, std_dev_y; uint8_t color[4]; float opacity; } drop_shadow;
        struct { uint8_t color[4]; float opacity; } flood;          /* color[3] unused */
        struct { float std_dev_x, std_dev_y; } blur;
        struct { uint32_t input_offset, input_count; } merge;       /* inputs in rvk_filter_data.inputs */
        struct { uint32_t op; float radius_x, radius_y; } morphology;
        struct { float dx, dy; } offset;
        struct { float base_frequency_x, base_frequency_y; uint32_t num_octaves;
                 int32_t seed; uint32_t stitch_tiles, fractal_noise; } turbulence;
        struct { uint32_t root; } image;  /* feImage (usvg filter::Image::root): a scene group whose
                                             children are drawn; only inside rvk_render and its
                                             relatives (rvk_filter_apply has no scene: unsupported) */
        float pad_[48];
    } u;
} rvk_primitive;

typedef struct rvk_filter {                /* usvg::filter::Filter (filter.rs:14-42) */
    rvk_rect rect;                         /* filter region, user space */
    uint32_t primitive_offset, primitive_count;
} rvk_filter;

typedef struct rvk_filter_data {
    const rvk_primitive *primitives;
    const float *values;                   /* convolve kernels, transfer tables */
    const int32_t *inputs;                 /* merge input lists */
} rvk_filter_data;

/* ------------------------------------------------------------------------- */
/* Scene: a flattened usvg::Tree (crates/usvg/src/tree/mod.rs)                */
/* ------------------------------------------------------------------------- */
enum { RVK_VERB_MOVE = 0, RVK_VERB_LINE, RVK_VERB_QUAD, RVK_VERB_CUBIC, RVK_VERB_CLOSE };
enum { RVK_PAINT_NONE = 0, RVK_PAINT_COLOR, RVK_PAINT_LINEAR, RVK_PAINT_RADIAL, RVK_PAINT_PATTERN };
enum { RVK_SPREAD_PAD = 0, RVK_SPREAD_REFLECT, RVK_SPREAD_REPEAT };
enum { RVK_FILL_NONZERO = 0, RVK_FILL_EVENODD };
enum { RVK_CAP_BUTT = 0, RVK_CAP_ROUND, RVK_CAP_SQUARE };
enum { RVK_JOIN_MITER = 0, RVK_JOIN_MITER_CLIP, RVK_JOIN_ROUND, RVK_JOIN_BEVEL };
enum { RVK_NODE_GROUP = 0, RVK_NODE_PATH, RVK_NODE_IMAGE };
/* tiny_skia::FilterQuality of a raster image (image.rs:176-183: OptimizeSpeed, CrispEdges
 * and Pixelated sample nearest, Smooth bilinear, OptimizeQuality and HighQuality bicubic) */
enum { RVK_QUALITY_NEAREST = 0, RVK_QUALITY_BILINEAR, RVK_QUALITY_BICUBIC };

typedef struct rvk_stop {                  /* usvg::Stop */
    float offset;
    uint8_t color[4];                      /* r, g, b, unused */
    float opacity;
} rvk_stop;

typedef struct rvk_paint {                 /* usvg::Paint + fill/stroke opacity */
    uint32_t kind;                         /* RVK_PAINT_* */
    uint8_t color[4];                      /* solid colour r, g, b, unused */
    float opacity;                         /* Fill::opacity / Stroke::opacity */
    uint32_t spread;                       /* RVK_SPREAD_* */
    float x1, y1, x2, y2;                  /* linear */
    float cx, cy, r, fx, fy, fr;           /* radial */
    rvk_transform transform;               /* gradient transform; a pattern's Pattern::transform */
    uint32_t stop_offset, stop_count;
    rvk_rect rect;                         /* pattern: Pattern::rect (user space, units resolved) */
    uint32_t root;                         /* pattern: group holding Pattern::root's children */
} rvk_paint;

typedef struct rvk_path {                  /* usvg::Path (mod.rs:1282-1440) */
    uint32_t verb_offset, verb_count;
    uint32_t point_offset;                 /* in points (x,y pairs) */
    uint32_t visible, anti_alias, paint_order_stroke_first;
    rvk_paint fill;
    uint32_t fill_rule;                    /* RVK_FILL_* */
    rvk_paint stroke;                      /* kind NONE = no stroke */
    float stroke_width, stroke_miter_limit;
    uint32_t stroke_cap, stroke_join;
    rvk_rect bounds;                       /* path.data().bounds(): bounds of all points */
    uint32_t dash_offset, dash_count;      /* Stroke::dasharray in rvk_scene.dashes (count 0: none) */
    float dash_phase;                      /* Stroke::dashoffset */
} rvk_path;

typedef struct rvk_group {                 /* usvg::Group (mod.rs:1035-1225) */
    rvk_transform transform;
    float opacity;
    uint32_t blend_mode;                   /* RVK_MIX_* */
    uint32_t isolate;
    int32_t clip_path, mask;               /* into rvk_scene.clip_paths / masks; -1: none */
    uint32_t filter_offset, filter_count;
    rvk_rect layer_bbox;                   /* layer_bounding_box(), object coordinates */
    uint32_t child_offset, child_count;    /* into rvk_scene.children */
} rvk_group;

typedef struct rvk_node_ref { uint32_t kind, index; } rvk_node_ref;

typedef struct rvk_image {                 /* usvg::Image of a raster kind (image.rs:4-35, 159-206) */
    const uint8_t *pixels;                 /* the decoded image as tiny-skia holds it: premultiplied
                                              RGBA8, row-major, width * 4 bytes a row, host memory */
    uint32_t width, height;                /* drawn over the rect (0, 0, width, height) with the
                                              node's transform (usvg puts the view box transform on
                                              the parent group) */
    uint32_t quality;                      /* RVK_QUALITY_* */
    uint32_t visible;                      /* Image::is_visible */
} rvk_image;

enum { RVK_MASK_LUMINANCE = 0, RVK_MASK_ALPHA };

typedef struct rvk_clip_path {             /* usvg::ClipPath (mod.rs:793-840) */
    rvk_transform transform;               /* ClipPath::transform (objectBoundingBox units folded in) */
    int32_t clip_path;                     /* the clip path's own clip-path; -1: none */
    uint32_t root;                         /* group holding its children (drawn with their
                                              transforms only: fills of paths, nested groups) */
} rvk_clip_path;

typedef struct rvk_mask {                  /* usvg::Mask (mod.rs:857-910) */
    rvk_rect rect;                         /* Mask::rect, user space */
    uint32_t kind;                         /* RVK_MASK_* */
    int32_t mask;                          /* the mask's own mask; -1: none */
    uint32_t root;                         /* group holding its content (rendered normally;
                                              the group's own transform is not applied) */
} rvk_mask;

typedef struct rvk_scene {
    const rvk_group *groups;    uint32_t group_count;   /* groups[0] is the root */
    const rvk_path *paths;      uint32_t path_count;
    const rvk_node_ref *children;
    const uint8_t *verbs;
    const float *points;
    const rvk_stop *stops;
    const rvk_filter *filters;
    rvk_filter_data filter_data;
    const rvk_clip_path *clip_paths;  uint32_t clip_path_count;
    const rvk_mask *masks;            uint32_t mask_count;
    const float *dashes;              /* dash arrays of the paths' strokes */
    const rvk_image *images;          uint32_t image_count;  /* raster images (RVK_NODE_IMAGE) */
} rvk_scene;

/* ------------------------------------------------------------------------- */
/* Context, memory, streams                                                  */
/* ------------------------------------------------------------------------- */
typedef struct rvk_context rvk_context;

/* Creates a context on HIP device `device` with its own stream. */
rvk_status rvk_context_create(int device, rvk_context **out);
rvk_status rvk_context_destroy(rvk_context *ctx);
/* Uses an existing hipStream_t (e.g. torch.cuda.current_stream().cuda_stream) instead. */
rvk_status rvk_context_set_stream(rvk_context *ctx, void *hip_stream);
void *rvk_context_stream(rvk_context *ctx);
rvk_status rvk_synchronize(rvk_context *ctx);
/* Device memory helpers for C callers without another allocator. */
rvk_status rvk_device_alloc(rvk_context *ctx, size_t bytes, void **out);
rvk_status rvk_device_free(rvk_context *ctx, void *dev_ptr);
rvk_status rvk_upload(rvk_context *ctx, void *dev_dst, const void *host_src, size_t bytes);
rvk_status rvk_download(rvk_context *ctx, void *host_dst, const void *dev_src, size_t bytes);

/* ------------------------------------------------------------------------- */
/* Filter primitives on device pixmaps                                       */
/* ------------------------------------------------------------------------- */
/* Image::into_color_space: demultiply -> LUT -> premultiply (filter/mod.rs:115-125,286-308). */
rvk_status rvk_into_linear_rgb(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h);
rvk_status rvk_into_srgb(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h);
/* demultiply_alpha / multiply_alpha (filter/mod.rs:129-146). */
rvk_status rvk_demultiply_alpha(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h);
rvk_status rvk_multiply_alpha(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h);

/* box_blur::apply (filter/box_blur.rs:23-34); in place. */
rvk_status rvk_box_blur(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h,
                        double sigma_x, double sigma_y);
/* No counterpart in the reference: rvk_box_blur for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The back buffer is a 64 MiB device array of the library (one per GPU, held until the
 * process ends), so w * h * 4 must not exceed 64 MiB (4096 x 4096); larger images return
 * RVK_ERR_TOO_LARGE.  No hipMalloc, hipFree or other allocation is made.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (several streams or threads)
 * race and must be ordered by the caller. */
rvk_status rvk_box_blur_noalloc(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h,
                                double sigma_x, double sigma_y);
/* iir_blur::apply (filter/iir_blur.rs:47-65); in place. */
rvk_status rvk_iir_blur(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h,
                        double sigma_x, double sigma_y);
/* morphology::apply (filter/morphology.rs:15-73); in place. rx, ry already scaled. */
rvk_status rvk_morphology(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h,
                          uint32_t op, float rx, float ry);
/* convolve_matrix::apply (filter/convolve_matrix.rs:15-111); in place.
 * `prim` must be an RVK_FE_CONVOLVE_MATRIX primitive; its kernel is read from host `values`. */
rvk_status rvk_convolve_matrix(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h,
                               const rvk_primitive *prim, const float *values);
/* color_matrix::apply (filter/color_matrix.rs:11-110) on UNPREMULTIPLIED pixels; in place. */
rvk_status rvk_color_matrix(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h,
                            const rvk_primitive *prim);
/* component_transfer::apply (filter/component_transfer.rs:10-72) on UNPREMULTIPLIED pixels. */
rvk_status rvk_component_transfer(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h,
                                  const rvk_primitive *prim, const float *values);
/* composite::arithmetic (filter/composite.rs:14-50): dst = f(src1, src2); dst untouched where alpha ~ 0. */
rvk_status rvk_composite_arithmetic(rvk_context *ctx, const uint8_t *dev_src1, const uint8_t *dev_src2,
                                    uint8_t *dev_dst, uint32_t w, uint32_t h,
                                    float k1, float k2, float k3, float k4);
/* turbulence::apply (filter/turbulence.rs:33-91): writes UNPREMULTIPLIED noise to dst. */
rvk_status rvk_turbulence(rvk_context *ctx, uint8_t *dev_dst, uint32_t w, uint32_t h,
                          double offset_x, double offset_y, double sx, double sy,
                          double base_frequency_x, double base_frequency_y, uint32_t num_octaves,
                          int32_t seed, uint32_t stitch_tiles, uint32_t fractal_noise);
/* lighting::diffuse_lighting / specular_lighting (filter/lighting.rs:132-225); `light` is
 * already in region pixel coordinates (transform_light_source, filter/mod.rs:1063-1098). */
rvk_status rvk_lighting(rvk_context *ctx, const uint8_t *dev_src, uint8_t *dev_dst, uint32_t w, uint32_t h,
                        uint32_t specular, float surface_scale, float constant, float specular_exponent,
                        const uint8_t color[4], const rvk_light_source *light);
/* displacement_map::apply (filter/displacement_map.rs:15-62); sx, sy = canvas scale. */
rvk_status rvk_displacement_map(rvk_context *ctx, const uint8_t *dev_src, const uint8_t *dev_map,
                                uint8_t *dev_dst, uint32_t w, uint32_t h, float sx, float sy,
                                float scale, uint32_t x_channel, uint32_t y_channel);

/* tiny_skia::PixmapMut::draw_pixmap(x, y, src, PixmapPaint{opacity, blend_mode, Nearest},
 * identity) — the call made by render.rs:127-140 and filter/mod.rs:624-1111. */
rvk_status rvk_draw_pixmap(rvk_context *ctx, uint8_t *dev_dst, uint32_t dw, uint32_t dh,
                           int32_t x, int32_t y, const uint8_t *dev_src, uint32_t sw, uint32_t sh,
                           float opacity, uint32_t blend_mode);
/* tiny_skia::Pixmap::fill(Color::from_rgba8(r, g, b, a)). */
rvk_status rvk_fill(rvk_context *ctx, uint8_t *dev_px, uint32_t w, uint32_t h, const uint8_t rgba[4]);

/* filter::apply (filter/mod.rs:335-355): runs one filter over the layer `dev_px` (w x h)
 * with the layer transform `ts`.  On a filter error the layer is cleared (as the reference
 * does) and RVK_OK is returned; `*out_error` (if non-null) receives the filter error
 * (RVK_OK, RVK_ERR_INVALID_REGION).  Scratch images are allocated from the context. */
rvk_status rvk_filter_apply(rvk_context *ctx, const rvk_filter *filter, const rvk_filter_data *data,
                            rvk_transform ts, uint8_t *dev_px, uint32_t w, uint32_t h,
                            rvk_status *out_error);

/* ------------------------------------------------------------------------- */
/* Rendering                                                                 */
/* ------------------------------------------------------------------------- */
/* resvg::render (lib.rs:34-43) onto a device pixmap: composites the scene onto the
 * existing content of `dev_pixmap` (it never clears it). */
rvk_status rvk_render(rvk_context *ctx, const rvk_scene *scene, rvk_transform transform,
                      uint32_t width, uint32_t height, uint8_t *dev_pixmap);
/* resvg_render (crates/c-api/lib.rs:874-897) twin: host pixmap in and out; uploads,
 * renders, downloads and synchronizes. */
rvk_status rvk_render_host(rvk_context *ctx, const rvk_scene *scene, rvk_transform transform,
                           uint32_t width, uint32_t height, uint8_t *host_pixmap);
/* PNG output, the device part of tiny-skia's Pixmap::encode_png / save_png that
 * resvg's CLI writes its result with (crates/resvg/src/main.rs:104-117): every pixel
 * of the premultiplied RGBA8 pixmap `dev_px` demultiplied like
 * PremultipliedColorU8::demultiply (c / (a / 255) + 0.5 in f64, cast to u8; restated
 * by the reference's test harness, crates/resvg/tests/integration/main.rs:204-211),
 * each row PNG-filtered with the filter type (0-4) of least signed-magnitude sum, and
 * written to `dev_out` as the IDAT stream's uncompressed data: per row one filter-type
 * byte, then 4 * width bytes (height * (4 * width + 1) bytes in all, device memory).
 * The caller deflates it and adds the chunks (resvg_amd.Context.encode_png). */
rvk_status rvk_png_filter_rows(rvk_context *ctx, const uint8_t *dev_px, uint32_t width, uint32_t height,
                               uint8_t *dev_out);
/* resvg::render_node (lib.rs:55-70) / resvg_render_node (crates/c-api/lib.rs:907-942)
 * onto a device pixmap.  `node` is the group or path to draw (the FFI crate resolves
 * usvg's node id to it); `abs_layer_bbox` is usvg's Node::abs_layer_bounding_box()
 * (NULL, or a zero, negative or non-finite width or height, when the node has none).
 * The node is drawn with `transform` pre-translated by (-bbox.x, -bbox.y), so the
 * node's layer bounding box lands at the pixmap origin; a group node applies its own
 * transform, opacity, blend mode, filters, clip path and mask (render.rs:21-47).
 * *rendered (required) receives 1 when the node was drawn and 0 when it has no bounding
 * box (resvg_render_node's `false`), in which case nothing is drawn and RVK_OK is
 * returned. */
rvk_status rvk_render_node(rvk_context *ctx, const rvk_scene *scene, rvk_node_ref node,
                           const rvk_rect *abs_layer_bbox, rvk_transform transform,
                           uint32_t width, uint32_t height, uint8_t *dev_pixmap, int32_t *rendered);

/* ------------------------------------------------------------------------- */
/* Per-operation replacements of the tiny-skia calls the render walk makes     */
/* (for a caller that keeps resvg's own walk and swaps only the pixel calls)  */
/* ------------------------------------------------------------------------- */
/* A tiny_skia_path::Path: verbs (RVK_VERB_*) and their points (x, y pairs, f32). */
typedef struct rvk_path_geometry {
    const uint8_t *verbs;   uint32_t verb_count;
    const float *points;    uint32_t point_count;   /* in points, not floats */
} rvk_path_geometry;

/* tiny_skia::Stroke as usvg::Stroke::to_tiny_skia builds it (crates/usvg/src/tree/mod.rs:638-662). */
typedef struct rvk_stroke_style {
    float width, miter_limit;
    uint32_t cap, join;                    /* RVK_CAP_*, RVK_JOIN_* */
    const float *dashes;                   /* resolved dash array (even length, sum > 0), or NULL */
    uint32_t dash_count;
    float dash_offset;
} rvk_stroke_style;

/* resvg path::fill_path (path.rs:26-75): tiny_skia Pixmap::fill_path(path, paint, rule,
 * transform, None) with the paint's shader built from `paint` (solid colour, or a
 * linear / radial gradient whose stops are paint->stop_offset .. + stop_count of
 * `stops`; the paint's opacity folded in as path.rs does), `blend_mode`
 * RVK_BLEND_SOURCE_OVER (path::render) or RVK_BLEND_CLEAR (clip path children,
 * clip.rs:30-70), anti-aliasing `anti_alias`.  A path whose bounds have zero width or
 * height draws nothing (path.rs:36-38).  Other blend modes return RVK_ERR_UNSUPPORTED. */
rvk_status rvk_fill_path(rvk_context *ctx, uint8_t *dev_px, uint32_t width, uint32_t height,
                         const rvk_path_geometry *path, uint32_t fill_rule, const rvk_paint *paint,
                         const rvk_stop *stops, rvk_transform transform, uint32_t blend_mode,
                         uint32_t anti_alias);
/* resvg path::stroke_path (path.rs:77-116): tiny_skia Pixmap::stroke_path(path, paint,
 * stroke, transform, None), SourceOver; a stroke at most one device pixel wide is
 * drawn as tiny-skia's anti-aliased hairline. */
rvk_status rvk_stroke_path(rvk_context *ctx, uint8_t *dev_px, uint32_t width, uint32_t height,
                           const rvk_path_geometry *path, const rvk_stroke_style *stroke,
                           const rvk_paint *paint, const rvk_stop *stops, rvk_transform transform,
                           uint32_t anti_alias);

/* Masks: a u8 plane of width * height bytes, row-major (tiny_skia::Mask).  The mask
 * entry points take 4-byte aligned masks and 16-byte aligned pixmaps (device
 * allocations and torch tensors are) and return RVK_ERR_INVALID_ARGUMENT otherwise. */
enum { RVK_MASK_TYPE_ALPHA = 0, RVK_MASK_TYPE_LUMINANCE = 1 };
/* tiny_skia::Mask::fill_path(path, rule, anti_alias, transform) on a new (zeroed) mask
 * (mask.rs:20-30): writes the path's coverage into dev_mask. */
rvk_status rvk_mask_fill_path(rvk_context *ctx, uint8_t *dev_mask, uint32_t width, uint32_t height,
                              const rvk_path_geometry *path, uint32_t fill_rule, uint32_t anti_alias,
                              rvk_transform transform);
/* tiny_skia::Mask::from_pixmap(pixmap, MaskType) (clip.rs:25, mask.rs:44): alpha, or the
 * luminance of the demultiplied colour (Rec. 709 weights) times alpha, rounded up. */
rvk_status rvk_mask_from_pixmap(rvk_context *ctx, const uint8_t *dev_px, uint32_t width, uint32_t height,
                                uint32_t mask_type, uint8_t *dev_mask);
/* tiny_skia::Mask::invert (clip.rs:26): m = 255 - m. */
rvk_status rvk_mask_invert(rvk_context *ctx, uint8_t *dev_mask, uint32_t width, uint32_t height);
/* tiny_skia::Pixmap::apply_mask (clip.rs:27, mask.rs:32,45): every channel c becomes
 * (c * m + 255) >> 8. */
rvk_status rvk_apply_mask(rvk_context *ctx, uint8_t *dev_px, uint32_t width, uint32_t height,
                          const uint8_t *dev_mask);
/* tiny_skia::Pixmap::fill_rect(rect, Paint{BlendMode::Clear}, identity, None) on a
 * pixel-aligned rect (filter/mod.rs:457-508, the subregion crop): zeroes
 * [x, x + rect_w) x [y, y + rect_h) intersected with the pixmap. */
rvk_status rvk_clear_rect(rvk_context *ctx, uint8_t *dev_px, uint32_t width, uint32_t height,
                          int32_t x, int32_t y, uint32_t rect_w, uint32_t rect_h);
/* A loop of resvg::render calls (lib.rs:34-43), one per document, in one call: document i
 * is composited onto dev_pixmaps[i] (widths[i] x heights[i]) with transforms[i], with one
 * host wait at the end.  The documents are spread over up to 16 internal streams
 * (RVK_BATCH_LANES) that start after the work already queued on the context stream and
 * are joined back into it, so later work on the context stream sees every document;
 * documents whose pixmaps overlap are drawn in order on one stream, as the loop would.
 * statuses[i] (if non-null) receives each document's status; the return value is the
 * first non-OK status, RVK_OK when every document rendered.  Arguments are checked for
 * every document before any is drawn. */
rvk_status rvk_render_batch(rvk_context *ctx, uint32_t count, const rvk_scene *const *scenes,
                            const rvk_transform *transforms, const uint32_t *widths, const uint32_t *heights,
                            uint8_t *const *dev_pixmaps, rvk_status *statuses);

/* ------------------------------------------------------------------------- */
/* Differentiable f32 form (no counterpart in the reference, which renders u8 only) */
/* ------------------------------------------------------------------------- */
/* A scene of the differentiable renderer: per path a u8 coverage plane (the
 * anti-aliased fill coverage the u8 pipeline computes for the path, width*height
 * bytes, path order) and the end points (x0, y0, x1, y1) of its pad-spread linear
 * gradient in pixels.  The paint, group opacity, blur σ and colour matrix are
 * parameters.  Model: paths composite SourceOver (path.rs:26-75, render.rs:133) with
 * source (rgb·tint·A, A), A = coverage·opacity·stop alpha, the stop colour
 * interpolated between 4 stops (path.rs:118-177); then feGaussianBlur (separable
 * Gaussian, radius ceil(3σ) <= 48, zero outside) -> feColorMatrix on the
 * demultiplied colour (color_matrix.rs:11-110) -> composite over the paths' canvas
 * (filter/mod.rs:731-793).  Everything is f32 premultiplied RGBA. */
typedef struct rvk_diff_scene {
    uint32_t width, height, path_count;
    /* 0: coverage from dev_coverage; k > 0: analytic mode, every path a closed chain
     * of k cubics given by the geometry parameters (coverage = min(1, |winding
     * area|) of the path flattened into 16 segments per cubic), with gradients to
     * the control points and the path's affine transform */
    uint32_t cubics;
    const uint8_t *dev_coverage;      /* cubics == 0: path_count x height x width */
    const float *dev_gradient_ends;   /* path_count x 4 floats, 16-byte aligned */
} rvk_diff_scene;
/* Parameters per path: tint r, g, b; opacity; 4 stop colours (r, g, b, a); 4 stop
 * offsets (ascending).  After the paths: σ, then the 4x5 colour matrix row-major. */
#define RVK_DIFF_PATH_PARAMS 24
/* Geometry parameters per path (analytic mode): the affine (sx, ky, kx, sy, tx, ty;
 * device = (sx x + kx y + tx, ky x + sy y + ty)) and the 3k + 1 user-space points
 * P0 .. P3k (x, y); the polygon closes from P3k back to P0. */
#define RVK_DIFF_GEOM_FLOATS(k) (6 + 2 * (3 * (k) + 1))
rvk_status rvk_diff_workspace_size(uint32_t width, uint32_t height, uint32_t path_count, uint32_t cubics,
                                   size_t *bytes);
/* Renders the scene with `dev_params` (and `dev_geom` in the analytic mode, else
 * NULL) into dev_out (width x height x 4 floats) and keeps what the backward pass
 * needs (canvas checkpoints every 8 paths, the blur intermediates, the coverage
 * planes) in dev_workspace (rvk_diff_workspace_size bytes). */
rvk_status rvk_diff_forward(rvk_context *ctx, const rvk_diff_scene *scene, const float *dev_params,
                            const float *dev_geom, void *dev_workspace, float *dev_out);
/* From dL/d(out) (width x height x 4 floats) writes dL/d(params) and, in the
 * analytic mode, dL/d(geom).  Must follow rvk_diff_forward with the same scene,
 * parameters, geometry and workspace. */
rvk_status rvk_diff_backward(rvk_context *ctx, const rvk_diff_scene *scene, const float *dev_params,
                             const float *dev_geom, void *dev_workspace, const float *dev_grad_out,
                             float *dev_grad_params, float *dev_grad_geom);
/* The same for `batch` (1 .. 65535) scenes of one shape (`scene` gives the shape and
 * the base pointers), every kernel launched once for the whole batch.  Each
 * per-scene array is stacked, scene b at b times its per-scene size: parameters
 * (path_count * 24 + 21 floats), geometry (path_count * RVK_DIFF_GEOM_FLOATS(cubics)
 * floats), gradient ends (path_count x 4 floats), coverage planes (path_count x
 * height x width bytes), workspace (rvk_diff_workspace_size bytes, so batch times
 * that in all), images and their gradients (width x height x 4 floats).  The scenes'
 * results are those of one call per scene. */
rvk_status rvk_diff_forward_batch(rvk_context *ctx, const rvk_diff_scene *scene, uint32_t batch,
                                  const float *dev_params, const float *dev_geom, void *dev_workspace,
                                  float *dev_out);
rvk_status rvk_diff_backward_batch(rvk_context *ctx, const rvk_diff_scene *scene, uint32_t batch,
                                   const float *dev_params, const float *dev_geom, void *dev_workspace,
                                   const float *dev_grad_out, float *dev_grad_params, float *dev_grad_geom);

/* Per-call statistics of the last rvk_render / rvk_filter_apply on this context. */
typedef struct rvk_stats {
    uint64_t kernel_launches;
    uint64_t layers;
    uint64_t paths;
    uint64_t filters;
    uint64_t scratch_bytes;
    uint64_t edge_pieces;   /* line pieces built for coverage */
    uint64_t band_entries;  /* piece references in the row bands */
    uint64_t upload_bytes;  /* draw-list metadata sent to the device */
    uint64_t flushes;       /* draw-list submissions */
    double host_build_ms;   /* host time building edges, bands and shaders */
    double host_wait_ms;    /* host time waiting for the device */
    double host_geom_ms;    /* part of host_build_ms: path geometry (edges, strokes), on all host threads */
    uint64_t validated_flushes; /* flush blocks index-checked on the host (RVK_VALIDATE set) */
} rvk_stats;
rvk_status rvk_get_stats(rvk_context *ctx, rvk_stats *out);

/* Kernel timing (no counterpart in the reference, which has no device kernels):
 * between rvk_profile_begin and rvk_profile_end every kernel launch on the
 * context's stream is bracketed by a pair of HIP events and charged with the
 * algorithmic bytes it must move (the bytes of its inputs and outputs, each read or
 * written once).  rvk_profile_end waits for the stream and returns one record per
 * kernel name, sorted by total time (largest first). */
typedef struct rvk_kernel_stat {
    char name[40];
    uint64_t launches;
    double total_ms;
    double bytes;
} rvk_kernel_stat;
rvk_status rvk_profile_begin(rvk_context *ctx);
rvk_status rvk_profile_end(rvk_context *ctx, rvk_kernel_stat *out, uint32_t capacity, uint32_t *count);

#ifdef __cplusplus
}
#endif
#endif /* RESVG_AMD_H */
