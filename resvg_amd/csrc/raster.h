// Host-side rasterizer front end and device data layouts shared with raster.hip.
//
// A fill or stroke becomes a coverage job: a list of line pieces in 16.16 fixed
// point on the supersampled grid (Skia's edge stepping, tiny-skia edge.rs, run to
// completion on the host), bucketed by bands of supersampled rows, plus the pixel
// rectangle whose coverage the device computes.  Shading parameters are plain
// structs evaluated per pixel by the compositor.
#pragma once
#include <cstdint>
#include <vector>
#include "host_geom.h"

namespace rvkr {

// One line piece of an edge: active on supersampled rows [first_y, last_y]; its
// 16.16 x at row y is x + (y - first_y) * dx.  `id` orders pieces with equal x.
struct Piece {
    int32_t x, dx, first_y, last_y;
    int32_t winding, id;
};

constexpr int kBandShift = 4;  // 16 supersampled rows (4 pixel rows) per band

// A coverage job computes u8 coverage for the pixel rect [x0, x0+w) x [y0, y0+h)
// of its pixmap (coordinates of the edge space) into cov[(y - y0) * w + (x - x0)].
struct CovJob {
    int32_t x0, y0, w, h;
    uint32_t piece_offset, piece_count;
    uint32_t band_offset;  // into band_start: bands cover rows y0*scale .. (y0+h)*scale
    uint32_t band_count;
    int32_t band_y0;       // first supersampled row of band 0
    uint32_t rule, aa;
    uint64_t cov_offset;   // byte offset into the coverage arena
    uint32_t row_start;    // prefix over jobs of h (wave index of the first row)
    uint32_t pad_;
};

// Shader parameters (tiny-skia lowp stages for solid, linear and concentric radial).
constexpr int kMaxStops = 16;
struct ShaderParams {
    uint32_t kind;       // 0 none, 1 solid, 2 gradient
    uint32_t radial, spread, pad_clamp, two_stop, opaque, r_twist, nstops;
    uint32_t rgba;       // solid premultiplied colour
    // Opaque shader (solid alpha 1, linear gradient with opaque stops): tiny-skia
    // reduces SourceOver to Source, so partial coverage lerps towards the colour.
    uint32_t as_source;
    float sx, ky, kx, sy, tx, ty;  // device -> unit transform
    float r_scale, r_bias;  // focal radius: t = (t - r_scale) / r_bias
    float t[kMaxStops + 2];
    float f[kMaxStops + 2][4];
    float b[kMaxStops + 2][4];
};

// A draw item of a draw list: fill and/or stroke coverage shaded and composited
// into the target, either directly (lowp SourceOver) or through an elided layer
// (the layer pixel is built in registers, then composited with opacity and blend
// mode in f32 like draw_pixmap).
struct DrawItem {
    int32_t x0, y0, x1, y1;     // target-pixel rect touched (clipped to the target)
    int32_t fill_job, stroke_job;  // coverage jobs (-1: none); their rects are in layer space
    int32_t lx, ly;             // layer origin in target pixels (0,0 for direct items)
    int32_t fill_shader, stroke_shader;
    uint32_t stroke_first;
    uint32_t layer;             // 0 direct, 1 elided layer
    float opacity;
    uint32_t blend_mode;
};

struct Pt { float x, y; };
struct Seg { int n; Pt p[4]; };

// Path segments in device space with every contour closed (fill semantics).
void fill_segments(const rvk_scene &sc, const rvk_path &path, rvkh::Ts ts, std::vector<Seg> &out);
// Stroke outline polygons (this build's polyline stroker) in device space.
void stroke_segments(const rvk_scene &sc, const rvk_path &path, rvkh::Ts ts, std::vector<Seg> &out);
// Skia edges of `segs` on the supersampled grid, expanded into line pieces; returns
// the supersampled row range [ymin, ymax] (ymin > ymax when empty).
void build_pieces(const std::vector<Seg> &segs, int shift, std::vector<Piece> &out, int32_t &ymin, int32_t &ymax);
// tiny-skia shader setup; false when the paint draws nothing.
bool make_shader(const rvk_scene &sc, const rvk_paint &paint, rvkh::Ts ts, ShaderParams &out);
// Control-point bounds of device-space segments nearly zero (tiny-skia skips such fills).
bool bounds_nearly_zero(const std::vector<Seg> &segs);

}  // namespace rvkr
