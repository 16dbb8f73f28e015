// Oracle: CPU restatement of resvg 0.48.1's pixel pipeline, used ONLY as test
// infrastructure (pytest, the smoke check and bench.py's cpu_baseline leg).  The
// product library never links, loads or calls anything in this directory.
//
// Every function cites the reference file:line it restates (paths relative to
// the resvg tree).  Pixels are premultiplied RGBA8 unless stated otherwise.
// Rust float semantics are kept: no FMA contraction (built with
// -ffp-contract=off), `as u8` casts saturate and map NaN to 0.
#pragma once
#include <cstdint>
#include <cstddef>
#include "../include/resvg_amd.h"

extern "C" {
// colour space (filter/mod.rs:115-238)
void orc_into_linear_rgb(uint8_t *px, uint32_t w, uint32_t h);
void orc_into_srgb(uint8_t *px, uint32_t w, uint32_t h);
void orc_demultiply_alpha(uint8_t *px, uint32_t w, uint32_t h);
void orc_multiply_alpha(uint8_t *px, uint32_t w, uint32_t h);
const uint8_t *orc_srgb_to_linear_table(void);
const uint8_t *orc_linear_to_srgb_table(void);

// primitives
void orc_box_blur(uint8_t *px, uint32_t w, uint32_t h, double sx, double sy);
void orc_box_gauss_sizes(float sigma, int32_t out[5]);
void orc_iir_blur(uint8_t *px, uint32_t w, uint32_t h, double sx, double sy);
void orc_morphology(uint8_t *px, uint32_t w, uint32_t h, uint32_t op, float rx, float ry);
void orc_convolve_matrix(uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p, const float *values);
void orc_color_matrix(uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p);
void orc_component_transfer(uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p, const float *values);
void orc_composite_arithmetic(const uint8_t *s1, const uint8_t *s2, uint8_t *dst, uint32_t w, uint32_t h,
                              float k1, float k2, float k3, float k4);
void orc_turbulence(uint8_t *dst, uint32_t w, uint32_t h, double ox, double oy, double sx, double sy,
                    double bfx, double bfy, uint32_t octaves, int32_t seed, uint32_t stitch, uint32_t fractal);
int32_t orc_turbulence_random(int32_t seed);
void orc_turbulence_init(int32_t seed, int32_t *lattice_out /*514*/, double *gradient_out /*4*514*2*/);
void orc_lighting(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, uint32_t specular,
                  float surface_scale, float constant, float spec_exp, const uint8_t color[4],
                  const rvk_light_source *light);
void orc_displacement_map(const uint8_t *src, const uint8_t *map, uint8_t *dst, uint32_t w, uint32_t h,
                          float sx, float sy, float scale, uint32_t xch, uint32_t ych);

// tiny-skia pixel operations
void orc_draw_pixmap(uint8_t *dst, uint32_t dw, uint32_t dh, int32_t x, int32_t y, const uint8_t *src,
                     uint32_t sw, uint32_t sh, float opacity, uint32_t blend_mode);
void orc_fill(uint8_t *px, uint32_t w, uint32_t h, const uint8_t rgba[4]);

// filter::apply (filter/mod.rs:335-355); returns the filter error status
int32_t orc_filter_apply(const rvk_filter *filter, const rvk_filter_data *data, rvk_transform ts,
                         uint8_t *px, uint32_t w, uint32_t h);

// resvg::render (lib.rs:34-43)
int32_t orc_render(const rvk_scene *scene, rvk_transform ts, uint32_t w, uint32_t h, uint8_t *px);
// fill a path into a pixmap (tiny-skia fill_path + resvg path.rs:26-75), exposed for tests
int32_t orc_fill_path(const rvk_scene *scene, uint32_t path_index, rvk_transform ts,
                      uint8_t *px, uint32_t w, uint32_t h);
// stroke a path (tiny-skia stroke_path + resvg path.rs:77-116); fill it with BlendMode::Clear
int32_t orc_stroke_path(const rvk_scene *scene, uint32_t path_index, rvk_transform ts,
                        uint8_t *px, uint32_t w, uint32_t h);
int32_t orc_clear_path(const rvk_scene *scene, uint32_t path_index, rvk_transform ts,
                       uint8_t *px, uint32_t w, uint32_t h);
// resvg::render_node (lib.rs:55-70); returns 1 when drawn, 0 without a bounding box
int32_t orc_render_node(const rvk_scene *scene, uint32_t kind, uint32_t index, const rvk_rect *bbox,
                        rvk_transform ts, uint32_t w, uint32_t h, uint8_t *px);
// tiny_skia::Mask::from_pixmap (alpha, or luminance when `luminance`) and Pixmap::apply_mask
void orc_mask_from_pixmap(const uint8_t *px, uint32_t w, uint32_t h, uint32_t luminance, uint8_t *mask);
void orc_apply_mask(uint8_t *px, uint32_t w, uint32_t h, const uint8_t *mask);
// tiny-skia's EdgeClipper against [0, w] x [0, h] on device-space segments, each 9
// floats (point count 2..4, then x0 y0 .. x3 y3); writes the clipped pieces in the
// same layout and returns their number, -1 on a bad count or too small an output
int32_t orc_edge_clip(const float *in, uint32_t n, uint32_t w, uint32_t h, float *out, uint32_t cap);
// the IDAT stream's rows of tiny-skia's Pixmap::encode_png (png.cpp): h * (4w + 1) bytes
void orc_png_filter_rows(const uint8_t *px, uint32_t w, uint32_t h, uint8_t *out);
// the u8 coverage mask of a path (for coverage-only tests)
int32_t orc_path_coverage(const rvk_scene *scene, uint32_t path_index, rvk_transform ts, uint32_t stroke,
                          uint8_t *cov, uint32_t w, uint32_t h);
}
