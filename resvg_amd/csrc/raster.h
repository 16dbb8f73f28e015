// Host-side rasterizer front end and device data layouts shared with raster.hip.
//
// A fill or stroke becomes a coverage job: the Skia edges of its path on the
// supersampled grid (tiny-skia edge.rs), set up on the host up to their first line
// piece, and the pixel rectangle whose coverage the device computes.  The device
// expands every curve edge into its remaining line pieces (k_expand, the
// forward-differencing update loop) and then computes coverage row by row.
#pragma once
#include <cstdint>
#include <vector>
#include "host_geom.h"

namespace rvkr {

// One edge as the host hands it over: its first line piece (16.16 x and slope on
// the supersampled rows first_y..last_y) and, for curves, the stepping state after
// that piece (tiny-skia QuadraticEdge / CubicEdge fields).
struct EdgeRec {
    int32_t x, dx, first_y, last_y;
    int32_t type;                 // 0 line, 1 quadratic, 2 cubic
    int32_t count;                // curve_count after the first piece
    int32_t shift, dshift;        // curve_shift; cubic_dshift
    int32_t cx, cy, lx, ly;       // current point, last point (16.16)
    int32_t d1x, d1y, d2x, d2y;   // quad: qdx qdy qddx qddy; cubic: cdx cdy cddx cddy
    int32_t d3x, d3y;             // cubic: cdddx cdddy
    uint32_t piece_offset;        // first slot of this edge's pieces
    int32_t winding;              // +1 downwards, -1 upwards
};

// What the coverage kernel reads per edge.  Edges of a job are sorted by first_y;
// ymax is an upper bound of the last covered row.
struct EdgeCov {
    int32_t first_y, ymax;
    int32_t winding, id;          // id: position in tiny-skia's edge list (orders equal x)
    uint32_t piece_offset, pad_;
};

// A line piece written by k_expand.
struct PieceDev {
    int32_t x, dx, first_y, last_y;
};

// Coverage is stored per job as a grid of kTile x kTile pixel tiles aligned to the
// target pixmap (so that one compositor wave, an 8 x 8 pixel square, reads one tile
// of each job).  A tile whose pixels all have coverage 0 or all 255 is only an entry
// of the job's tile table; the other (partial) tiles hold their 64 coverage bytes in
// the job's slot region of the arena.
constexpr int kTile = 8;
constexpr int kBandRows = kTile;    // pixel rows of one coverage band (one tile row)
constexpr int kChunkTiles = 1024;   // tiles of one band a coverage block handles at a time
// Tile table entries (u16): empty, full, or kTileSlot0 + k for the band's k-th partial
// tile, whose bytes (8 rows of 8, row-major) sit at slot (band * tiles_x + k).
constexpr uint32_t kTileEmpty = 0, kTileFull = 1, kTileSlot0 = 2;
constexpr uint32_t kMaxTilesX = 65279u;  // widest job grid (a u16 table entry and tile state address its slots)
// crossings per sub-scanline that the coverage kernel's variants hold (more keys,
// more LDS per wave, fewer waves per CU)
constexpr uint32_t kCoverageKeys[3] = {128, 256, 512};

// A coverage job computes u8 coverage for the pixel rect [x0, x0+w) x [y0, y0+h)
// of its pixmap (coordinates of the edge space; coverage is 0 outside the rect).
// Its tile grid starts at (gx0, gy0) (edge space, a multiple of kTile in target
// pixels) and has tiles_x x tiles_y tiles.  For every band (tile row) the window table
// holds the range of the job's edges that can touch the band.
struct CovJob {
    int32_t x0, y0, w, h;
    int32_t gx0, gy0;
    uint32_t tiles_x, tiles_y;
    uint32_t edge_offset, edge_count;
    uint32_t window_offset;  // into the window table (uint2 lo, hi per band)
    uint32_t rule, aa;
    uint32_t table_offset;   // first u16 entry of the job's tile table (tiles_x * tiles_y, row-major)
    uint64_t cov_offset;     // byte offset of the job's slot region (tiles_x * tiles_y * 64 bytes)
};

// An anti-aliased hairline segment set up for per-pixel evaluation: the state of
// tiny-skia's do_anti_hairline (scan/hairline_aa.rs) after its setup and clipping.
// Along the major axis (x, or y with kHairYMajor), pixel i in [istart, istop) sits at
// the 16.16 minor position f = fstart + (i - istart) * slope + 1/2; it blits coverage
// a = (f >> 8) & 255 to minor pixel f >> 16 and 255 - a to the pixel before, both
// scaled by scale_start / 64 at istart and by scale_stop / 64 at istop - 1 when
// scale_stop > 0.  With kHairClipped, blits outside [cl, cr) x [ct, cb) are dropped.
struct HairLine {
    int32_t istart, istop, fstart, slope;
    int32_t scale_start, scale_stop;
    uint32_t flags;
    int32_t cl, ct, cr, cb;
    int32_t pad_;
};
constexpr uint32_t kHairYMajor = 1, kHairClipped = 2;
constexpr int kHairTile = 32;  // pixels per side of a hairline bin tile

// A hairline job: the lines of one hairline stroke binned, in drawing order, into
// kHairTile x kHairTile tiles of the rect [x0, x0 + w) x [y0, y0 + h) (edge space).
// Tile t's line indices are bins[tiles[tile_offset + t] .. tiles[tile_offset + t + 1]).
struct HairJob {
    int32_t x0, y0, w, h;
    uint32_t tiles_x, tiles_y;
    uint32_t tile_offset;
    uint32_t pad_;
};

// Shader parameters (tiny-skia lowp stages for solid, linear and concentric radial).
constexpr int kMaxStops = 16;
// The scalar part of a shader, read by the compositor in one piece.
struct ShaderHead {
    // the fields below that the compositor branches on, packed into one word (kMode*
    // bits, set by seal_shader once the shader is complete), so that one scalar load
    // decides every branch; then the solid colour and the gradient transform, adjacent
    // so that one load brings them
    uint32_t mode;
    uint32_t rgba;       // solid premultiplied colour
    float sx, ky, kx, sy, tx, ty;  // device -> unit transform
    float r_scale, r_bias;  // focal radius: t = (t - r_scale) / r_bias
    uint32_t kind;       // 0 none, 1 solid, 2 gradient, 3 pattern
    uint32_t radial, spread, pad_clamp, two_stop, opaque, r_twist, nstops;
    // Opaque shader (solid alpha 1, linear gradient with opaque stops): tiny-skia
    // reduces SourceOver to Source, so partial coverage lerps towards the colour.
    uint32_t as_source;
    // BlendMode::Clear (clip path children): partial coverage lerps every channel
    // toward zero, div255(d * (255 - c)); the colour is not used
    uint32_t clear;
    // two-point conical gradient with the focal point off the centre (tiny-skia's
    // FocalData, Skia's kFocal): kFocal* below, 0 for none; p0 = 1 / r1.  Such a shader
    // runs in tiny-skia's f32 pipeline (its stages have no u8 form), so the colour,
    // the coverage scale and SourceOver stay in f32 until the store
    uint32_t focal;
    float p0;
    // 1: the shader runs in tiny-skia's f32 pipeline (a focal gradient or a pattern)
    uint32_t f32_pipeline;
    // pattern (tiny-skia's Pattern shader: a pattern paint as path.rs:58-68 sets it
    // up, spread Repeat; a raster image as image.rs:185-191 does, spread Pad): the
    // pixmap's device address and size, 1 / size, nearest (0), bilinear (1) or
    // bicubic (2) sampling, and the opacity applied by scale_1_float; sx..ty map
    // device pixels to pixmap pixels; `spread` holds the spread mode
    uint64_t tile;
    uint32_t tile_w, tile_h;
    float tile_iw, tile_ih;
    uint32_t quality;
    float pattern_opacity;
};
constexpr uint32_t kFocalOnCircle = 1, kFocalWellBehaved = 2, kFocalGreater = 3;
// ShaderHead::mode: kind in bits 0-1, spread in bits 3-4, nstops from bit 16
constexpr uint32_t kModeKind = 3u, kModeRadial = 1u << 2, kModeSpreadShift = 3, kModePadClamp = 1u << 5,
                   kModeTwoStop = 1u << 6, kModeOpaque = 1u << 7, kModeRTwist = 1u << 8, kModeAsSource = 1u << 9,
                   kModeClear = 1u << 10, kModeF32 = 1u << 11, kModeStopsShift = 16;
struct ShaderParams : ShaderHead {
    float t[kMaxStops + 2];
    float f[kMaxStops + 2][4];
    float b[kMaxStops + 2][4];
};

// A draw item of a draw list: fill and/or stroke coverage shaded and composited
// into the target, either directly (lowp SourceOver) or through an elided layer
// (the layer pixel is built in registers, then composited with opacity and blend
// mode in f32 like draw_pixmap).
// What the compositor needs of a job, copied into the draw item so that one load of
// the item brings it.  Coverage job: its tile grid in target tiles (x0, y0 = origin,
// w x h tiles; w = h = 0 for no job), its tile table and slot region.  Hairline job:
// its rect in layer pixels and its bin tile grid.
struct JobRef {
    int32_t x0, y0;
    uint32_t w, h;
    uint32_t table;        // coverage: first tile-table entry; hairline: first entry of its tile offsets
    uint32_t stride;       // hairline: bin tiles per row
    uint64_t cov_offset;   // coverage: byte offset of the job's slot region in the arena
};

struct DrawItem {
    int32_t x0, y0, x1, y1;     // target-pixel rect touched (clipped to the target)
    int32_t fill_job, stroke_job;  // coverage jobs (-1: none)
    int32_t lx, ly;             // layer origin in target pixels (0,0 for direct items)
    int32_t fill_shader, stroke_shader;
    uint32_t stroke_first;
    uint32_t layer;             // 0 direct, 1 elided layer
    float opacity;
    uint32_t blend_mode;
    uint32_t stroke_hair;       // 1: stroke_job is a hairline job
    uint32_t pad_;
    JobRef fj, sj;              // the fill and stroke jobs
};

// What the compositor's culling pass reads of an item, one lane per item: its target
// rect in tiles and the two jobs' tile grids and tables.
struct DrawCull {
    int32_t tx0, ty0, tx1, ty1;  // target tiles touched: [tx0, tx1) x [ty0, ty1)
    int32_t fx0, fy0;            // fill job grid origin (target tiles)
    uint32_t fw, fh, ftable;     // its size (0 x 0 for none) and table
    int32_t sx0, sy0;
    uint32_t sw, sh, stable;     // the stroke job (a hairline: sw = sh = 0)
    uint32_t hair;               // 1: a hairline stroke (drawn wherever the rect meets the tile)
};

struct Pt { float x, y; };
struct Seg { int n; Pt p[4]; };

// Path segments in device space with every contour closed (fill semantics).
void fill_segments(const rvk_scene &sc, const rvk_path &path, rvkh::Ts ts, std::vector<Seg> &out);
// Stroke outline (tiny-skia's stroker, stroker.cpp) in device space, contours closed.
void stroke_segments(const rvk_scene &sc, const rvk_path &path, rvkh::Ts ts, std::vector<Seg> &out);
// Skia edges of `segs` on the supersampled grid (lines combined like tiny-skia's
// edge builder, curves chopped at Y extrema unless `monotonic` says they already
// are, as clip_segments leaves them), each set up to its first piece, in
// tiny-skia's edge-list order (ids 0..n-1).  Returns the maximum number of pieces
// each edge can produce in `caps`.
void build_edges(const std::vector<Seg> &segs, int shift, std::vector<EdgeRec> &out, std::vector<uint32_t> &caps,
                 bool monotonic = false);
// Whether the rounded-out control-point bounds of `segs` lie inside a w x h pixmap
// (tiny-skia's path_contained_in_clip: floor/ceil for anti-aliased fills, the
// conservative rounding of scan/path.rs otherwise); when not, tiny-skia builds the
// edges from the clipped segments.
bool segs_inside_pixmap(const std::vector<Seg> &segs, bool aa, uint32_t w, uint32_t h);
// tiny-skia's EdgeClipper against [0, w] x [0, h] (edge_clip.cpp): lines, and
// quadratics and cubics monotonic in x and y.
void clip_segments(const std::vector<Seg> &segs, uint32_t w, uint32_t h, std::vector<Seg> &out);
// tiny-skia's hairline stroke of a path drawn with `ts` into a w x h pixmap
// (Pixmap::stroke_path's hairline branch): the prepared lines in drawing order.
void hairline_lines(const rvk_scene &sc, const rvk_path &path, rvkh::Ts ts, uint32_t w, uint32_t h,
                    std::vector<HairLine> &out);
// tiny-skia shader setup; false when the paint draws nothing.  `extra_opacity`
// scales the colour's or every stop's alpha (Shader::apply_opacity).
// Packs the branch fields of a complete shader into its mode word.
void seal_shader(ShaderHead &s);
bool make_shader(const rvk_scene &sc, const rvk_paint &paint, rvkh::Ts ts, ShaderParams &out,
                 float extra_opacity = 1.0f);
// render_pattern_pixmap's tile for a pattern paint drawn with `ts` (path.rs:179-205):
// the scale (sx, sy) of ts * Pattern::transform and the size round(rect * scale);
// false when the size is empty or too large for a pixmap.
bool pattern_tile_size(const rvk_paint &paint, rvkh::Ts ts, float &sx, float &sy, uint32_t &w, uint32_t &h);
// tiny-skia's Pattern shader over a tile rendered at scale (sx, sy) and stored at
// device address `tile` (Pattern::new, Shader::transform, Pattern::push_stages);
// false when the transform cannot be inverted.
bool make_pattern_shader(const rvk_paint &paint, rvkh::Ts ts, float sx, float sy, const void *tile, uint32_t w,
                         uint32_t h, ShaderParams &out, float extra_opacity = 1.0f);
// Device-space segments tiny-skia can fill: every point finite (Path::transform's
// Rect::from_points) and within f32::MAX / 4 (is_too_big_for_math).
bool segs_drawable(const std::vector<Seg> &segs);
// image::render_raster's Pattern shader (Pad, the image's filter quality, opacity 1,
// identity pattern transform) for the image drawn with `ts`, its pixmap at device
// address `tile`; false when `ts` cannot be inverted.
bool make_image_shader(const rvk_image &image, rvkh::Ts ts, const void *tile, ShaderParams &out);
// Control-point bounds of device-space segments nearly zero (tiny-skia skips such fills).
bool bounds_nearly_zero(const std::vector<Seg> &segs);

// Edges of one fill or stroke, computed off the walk (the expensive host part).
struct EdgeSet {
    bool ok = false;
    std::vector<EdgeRec> edges;
    std::vector<uint32_t> caps;
    float mnx = 0.0f, mxx = 0.0f;  // control-point x range of the device-space segments
};

// A coverage job prepared off the walk: the CovJob (offsets relative to the job),
// its edges sorted by first row with piece slots relative to the job, the coverage
// kernel's view of them, and per band the edge window and the coverage kernel
// variant.  add_job places it in a draw list.
struct JobPrep {
    bool ok = false;
    bool too_wide = false;           // the tile grid is wider than kMaxTilesX
    CovJob j{};
    std::vector<EdgeRec> edges;
    std::vector<EdgeCov> ecov;
    std::vector<uint2> windows;
    std::vector<uint8_t> tier;       // per band: index into kCoverageKeys
    uint32_t tier_count[3] = {0, 0, 0};
    uint64_t pieces = 0;             // piece slots of all edges
    uint64_t piece_base = 0;         // first piece slot in the draw list (add_job)
};

// A hairline job prepared off the walk: per tile the offset of its bin (relative to
// the job's bins, tiles_x * tiles_y + 1 entries), bins of line indices relative to the
// job's lines, each tile's in drawing order.
struct HairPrep {
    bool ok = false;
    HairJob j{};
    std::vector<uint32_t> tiles;
    std::vector<uint32_t> bins;
    std::vector<HairLine> lines;
    uint32_t bin_base = 0, line_base = 0;  // positions in the draw list (add_hair_job)
};

// The coverage job of an edge set drawn into the pixels `clip` of its pixmap, whose
// origin sits at target pixel (lx, ly): the tile grid aligned to the target, the
// edges sorted by first row (stable, so equal rows keep tiny-skia's edge order), the
// edge window of each band and the coverage kernel variant that holds its
// crossings.  Leaves P.ok false when nothing is covered, and also sets P.too_wide
// when the job's tile grid is wider than a table entry addresses.
void prep_job(const EdgeSet &es, uint32_t rule, bool aa, rvkh::IRect clip, int32_t lx, int32_t ly, JobPrep &P);
// The hairline job of prepared lines, binned into tiles of the part of their pixel
// bounds inside `clip`.  Leaves P.ok false when nothing is drawn.
void prep_hair(std::vector<HairLine> &&lines, rvkh::IRect clip, HairPrep &P);

}  // namespace rvkr
