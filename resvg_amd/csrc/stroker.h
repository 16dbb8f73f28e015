// Path stroking: a restatement of tiny-skia 0.12.0's PathStroker (a port of Skia's
// SkStroke / SkStrokerPriv), used by resvg's stroke_path (crates/resvg/src/path.rs:
// 77-116 -> tiny_skia::Pixmap::stroke_path).  The outline is built in user space
// from quadratic approximations of the offset curves (ray intersection), joins
// (miter, miter-clip, round, bevel) and caps (butt, round, square); the caller maps
// it with the path transform and fills it with the nonzero rule.
//
// Plain host C++ with no device code; the same file is compiled into librvk.so and
// into the CPU oracle, so stroke geometry is pinned by the reference's stroke
// fixtures rather than by the GPU parity tests.
//
// Restates tiny-skia (BSD-3-Clause, Copyright 2020 Yevhenii Reizner) and Skia
// (BSD-3-Clause, Copyright Google Inc.); see THIRD_PARTY_NOTICES.md.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace rvks {

struct Pt {
    float x, y;
};

enum Verb : uint8_t { kMove = 0, kLine, kQuad, kCubic, kClose };

// tiny-skia's PathBuilder: consecutive move_to's collapse, close() sets a pending
// move_to that the next drawing verb injects.
struct PathBuilder {
    std::vector<uint8_t> verbs;
    std::vector<Pt> pts;
    size_t last_move_to_index = 0;
    bool move_to_required = true;

    bool empty() const { return verbs.empty(); }
    void clear() {
        verbs.clear();
        pts.clear();
        last_move_to_index = 0;
        move_to_required = true;
    }
    void move_to(float x, float y);
    void line_to(float x, float y);
    void quad_to(float x1, float y1, float x2, float y2);
    void cubic_to(float x1, float y1, float x2, float y2, float x3, float y3);
    void conic_points_to(Pt p1, Pt p2, float weight);
    void close();
    void set_last_point(Pt p);
    bool last_point(Pt &out) const;
    bool is_zero_length_since_point(size_t start) const;
    void reverse_path_to(const PathBuilder &other);
    void push_path(const PathBuilder &other);
    void push_circle(float x, float y, float r);

  private:
    void inject_move_to_if_needed();
};

struct StrokeParams {
    float width;
    float miter_limit;
    uint32_t cap;   // 0 butt, 1 round, 2 square
    uint32_t join;  // 0 miter, 1 miter-clip, 2 round, 3 bevel
};

// tiny-skia PathStroker::compute_resolution_scale for a transform (sx, ky, kx, sy, tx, ty).
float resolution_scale(float sx, float ky, float kx, float sy);

// Path::stroke: strokes the path given as verbs (kMove..kClose) and points (x, y
// pairs) into `out`.  Returns false when the stroke produces no path.
bool stroke(const uint8_t *verbs, size_t nverbs, const float *points, size_t npoints, const StrokeParams &sp,
            float res_scale, PathBuilder &out);

// tiny-skia's treat_as_hairline: true (with the coverage) when an anti-aliased stroke
// is at most one device pixel wide under the 2x2 part of the transform.
bool treat_as_hairline(float width, bool anti_alias, float sx, float ky, float kx, float sy, float &coverage);

// tiny-skia's StrokeDash::new: the dash array (even length, no negative entry,
// positive finite sum), the offset reduced into [0, interval length), and the
// first interval the offset falls in.  False when tiny-skia rejects the dash.
struct DashParams {
    std::vector<float> array;
    float offset = 0.0f, interval_len = 0.0f, first_len = 0.0f;
    size_t first_index = 0;
};
bool dash_params(const float *array, size_t n, float offset, DashParams &out);

// Path::dash (dash.rs): the dashed path of verbs/points at resolution scale
// `res_scale` into `out`.  False when no dash is produced (tiny-skia returns None and
// stroke_path draws nothing).
bool dash(const uint8_t *verbs, size_t nverbs, const float *points, size_t npoints, const DashParams &d,
          float res_scale, PathBuilder &out);

// path_geometry helpers shared with the hairline code: the t values (clamped to
// [0, 1], sorted, distinct) where F'(t) . F''(t) = 0; de Casteljau split at t;
// valid_unit_divide; Transform::map_points.
int cubic_max_curvature(const Pt src[4], float t[3]);
void chop_cubic_at(const Pt src[4], float t, Pt dst[7]);
bool unit_divide(float numer, float denom, float &r);
void transform_points(float sx, float ky, float kx, float sy, float tx, float ty, Pt *p, size_t n);

}  // namespace rvks
