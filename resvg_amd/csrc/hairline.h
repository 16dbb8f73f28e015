// Anti-aliased hairlines: the geometry half of tiny-skia 0.12.0's hairline stroking
// (scan/hairline.rs stroke_path_impl + scan/hairline_aa.rs anti_hair_line_rgn, ports
// of Skia's SkScan_Hairline.cpp / SkScan_Antihair.cpp).  tiny-skia strokes a path
// whose device width is at most one pixel (treat_as_hairline) by walking its
// segments, subdividing quads and cubics into lines, clipping every line and calling
// do_anti_hairline on it, which blits pixel pairs along the line one after another.
//
// This file produces the ordered list of do_anti_hairline calls (FDot6 end points
// and the clip each runs under).  Plain host C++ compiled into librvk.so and into the
// CPU oracle; the pixel stepping is written separately on each side (the oracle
// blits sequentially, the device evaluates each pixel's blits in closed form).
//
// Restates tiny-skia (BSD-3-Clause, Copyright 2020 Yevhenii Reizner) and Skia
// (BSD-3-Clause, Copyright Google Inc.); see THIRD_PARTY_NOTICES.md.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "stroker.h"

namespace rvks {

// One do_anti_hairline call: end points in FDot6 (26.6 fixed point, device pixels)
// and, when `clipped`, the integer clip rect [cl, cr) x [ct, cb) it runs under.
struct HairSeg {
    int32_t x0, y0, x1, y1;
    int32_t cl, ct, cr, cb;
    bool clipped;
};

// tiny-skia hairline_aa::stroke_path for a device-space path (verbs kMove..kClose,
// points as x, y pairs) drawn into a w x h pixmap with line cap `cap` (0 butt,
// 1 round, 2 square): appends the do_anti_hairline calls in drawing order.
void anti_hairline_segments(const uint8_t *verbs, size_t nverbs, const Pt *pts, size_t npts, uint32_t cap,
                            uint32_t w, uint32_t h, std::vector<HairSeg> &out);

// The opacity Pixmap::stroke_path folds into the paint of a hairline whose coverage
// is below one (the SourceOver pre-scale branch): (255 * (int)(coverage * 256)) >> 8
// over 255; 1 for full coverage.
float hairline_paint_opacity(float coverage);

}  // namespace rvks
