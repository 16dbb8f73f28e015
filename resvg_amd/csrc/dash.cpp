// Path dashing: a restatement of tiny-skia 0.12.0's dash.rs (StrokeDash,
// ContourMeasureIter, ContourMeasure::push_segment; ports of Skia's SkDashPath and
// SkContourMeasure), used by Pixmap::stroke_path before stroking when the stroke
// has a dash array.  Float expressions keep the reference's operation order; build
// without FMA contraction.
//
// Restates tiny-skia (BSD-3-Clause, Copyright 2020 Yevhenii Reizner) and Skia
// (BSD-3-Clause, Copyright Google Inc.); see THIRD_PARTY_NOTICES.md.
#include "stroker.h"

#include <cmath>
#include <cstring>

namespace rvks {

namespace {

constexpr int32_t kMaxTValue = 0x3FFFFFFF;
constexpr float kCheapDistLimit = 0.5f;

inline Pt sub(Pt a, Pt b) { return Pt{a.x - b.x, a.y - b.y}; }
inline float interp(float a, float b, float t) { return a + (b - a) * t; }

// Point::length (f64 fallback when the squared length overflows)
float length_of(Pt v) {
    const float m2 = v.x * v.x + v.y * v.y;
    if (std::isfinite(m2)) return std::sqrt(m2);
    return (float)std::sqrt((double)v.x * v.x + (double)v.y * v.y);
}
inline float distance(Pt a, Pt b) { return length_of(sub(b, a)); }

enum SegKind : uint8_t { kSegLine, kSegQuad, kSegCubic };

struct Segment {
    float distance;     // accumulated length at the end of the segment
    uint32_t point_index;
    int32_t t_value;    // end t of the segment within its curve, in 1/kMaxTValue
    SegKind kind;
    float scalar_t() const {
        const float r = 1.0f / (float)kMaxTValue;
        const float t = (float)t_value * r;
        return t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    }
};

void chop_quad_at(const Pt *s, float t, Pt d[5]) {
    const Pt p01{interp(s[0].x, s[1].x, t), interp(s[0].y, s[1].y, t)};
    const Pt p12{interp(s[1].x, s[2].x, t), interp(s[1].y, s[2].y, t)};
    d[0] = s[0];
    d[1] = p01;
    d[2] = Pt{interp(p01.x, p12.x, t), interp(p01.y, p12.y, t)};
    d[3] = p12;
    d[4] = s[2];
}

bool quad_too_curvy(Pt p0, Pt p1, Pt p2, float tol) {
    // diff = -a/4 + b/2 - c/4
    const float dx = p1.x * 0.5f - ((p0.x + p2.x) * 0.5f) * 0.5f;
    const float dy = p1.y * 0.5f - ((p0.y + p2.y) * 0.5f) * 0.5f;
    return std::max(std::fabs(dx), std::fabs(dy)) > tol;
}

bool cheap_dist_exceeds_limit(Pt pt, float x, float y, float tol) {
    return std::max(std::fabs(x - pt.x), std::fabs(y - pt.y)) > tol;
}

bool cubic_too_curvy(Pt p0, Pt p1, Pt p2, Pt p3, float tol) {
    const bool n0 = cheap_dist_exceeds_limit(p1, interp(p0.x, p3.x, 1.0f / 3.0f), interp(p0.y, p3.y, 1.0f / 3.0f), tol);
    const bool n1 = cheap_dist_exceeds_limit(p2, interp(p0.x, p3.x, 2.0f / 3.0f), interp(p0.y, p3.y, 2.0f / 3.0f), tol);
    return n0 || n1;
}

struct Contour {
    std::vector<Segment> segs;
    std::vector<Pt> pts;
    float length = 0.0f;
    bool closed = false;

    float line_seg(Pt p0, Pt p1, float dist, uint32_t pi) {
        const float d = distance(p0, p1);
        const float prev = dist;
        dist += d;
        if (dist > prev) segs.push_back(Segment{dist, pi, kMaxTValue, kSegLine});
        return dist;
    }
    float quad_segs(Pt p0, Pt p1, Pt p2, float dist, int32_t mint, int32_t maxt, uint32_t pi, float tol) {
        if (((maxt - mint) >> 10) != 0 && quad_too_curvy(p0, p1, p2, tol)) {
            Pt tmp[5];
            const Pt src[3] = {p0, p1, p2};
            const int32_t half = (mint + maxt) >> 1;
            chop_quad_at(src, 0.5f, tmp);
            dist = quad_segs(tmp[0], tmp[1], tmp[2], dist, mint, half, pi, tol);
            dist = quad_segs(tmp[2], tmp[3], tmp[4], dist, half, maxt, pi, tol);
        } else {
            const float d = distance(p0, p2);
            const float prev = dist;
            dist += d;
            if (dist > prev) segs.push_back(Segment{dist, pi, maxt, kSegQuad});
        }
        return dist;
    }
    float cubic_segs(Pt p0, Pt p1, Pt p2, Pt p3, float dist, int32_t mint, int32_t maxt, uint32_t pi, float tol) {
        if (((maxt - mint) >> 10) != 0 && cubic_too_curvy(p0, p1, p2, p3, tol)) {
            Pt tmp[7];
            const Pt src[4] = {p0, p1, p2, p3};
            const int32_t half = (mint + maxt) >> 1;
            chop_cubic_at(src, 0.5f, tmp);
            dist = cubic_segs(tmp[0], tmp[1], tmp[2], tmp[3], dist, mint, half, pi, tol);
            dist = cubic_segs(tmp[3], tmp[4], tmp[5], tmp[6], dist, half, maxt, pi, tol);
        } else {
            const float d = distance(p0, p3);
            const float prev = dist;
            dist += d;
            if (dist > prev) segs.push_back(Segment{dist, pi, maxt, kSegCubic});
        }
        return dist;
    }

    // find_segment: index of the first segment whose distance >= key, or the
    // bit-inverted insertion index
    int32_t find_segment(float key) const {
        uint32_t lo = 0, hi = (uint32_t)segs.size() - 1;
        while (lo < hi) {
            const uint32_t mid = (hi + lo) >> 1;
            if (segs[mid].distance < key) lo = mid + 1;
            else hi = mid;
        }
        if (segs[hi].distance < key) {
            hi += 1;
            hi = ~hi;
        } else if (key < segs[hi].distance) {
            hi = ~hi;
        }
        return (int32_t)hi;
    }

    // distance_to_segment: the segment holding `d` and the curve t there (false when
    // t is not in [0, 1], NaN included)
    bool distance_to_segment(float d, size_t &index, float &t) const {
        int32_t i = find_segment(d);
        i ^= i >> 31;
        index = (size_t)i;
        const Segment &seg = segs[index];
        float start_t = 0.0f, start_d = 0.0f;
        if (index > 0) {
            start_d = segs[index - 1].distance;
            if (segs[index - 1].point_index == seg.point_index) start_t = segs[index - 1].scalar_t();
        }
        t = start_t + (seg.scalar_t() - start_t) * (d - start_d) / (seg.distance - start_d);
        return t >= 0.0f && t <= 1.0f;
    }

    static Pt pos_at(const Pt *p, SegKind kind, float t) {
        if (kind == kSegLine) return Pt{interp(p[0].x, p[1].x, t), interp(p[0].y, p[1].y, t)};
        if (kind == kSegQuad) {  // QuadCoeff: (a t + b) t + c
            const Pt b{(p[1].x - p[0].x) + (p[1].x - p[0].x), (p[1].y - p[0].y) + (p[1].y - p[0].y)};
            const Pt a{(p[2].x - (p[1].x + p[1].x)) + p[0].x, (p[2].y - (p[1].y + p[1].y)) + p[0].y};
            return Pt{(a.x * t + b.x) * t + p[0].x, (a.y * t + b.y) * t + p[0].y};
        }
        // CubicCoeff: ((a t + b) t + c) t + d
        const Pt a{(p[3].x + 3.0f * (p[1].x - p[2].x)) - p[0].x, (p[3].y + 3.0f * (p[1].y - p[2].y)) - p[0].y};
        const Pt b{3.0f * ((p[2].x - (p[1].x + p[1].x)) + p[0].x), 3.0f * ((p[2].y - (p[1].y + p[1].y)) + p[0].y)};
        const Pt c{3.0f * (p[1].x - p[0].x), 3.0f * (p[1].y - p[0].y)};
        return Pt{((a.x * t + b.x) * t + c.x) * t + p[0].x, ((a.y * t + b.y) * t + c.y) * t + p[0].y};
    }

    // NormalizedF32Exclusive::new_bounded
    static float exclusive(float t) {
        const float eps = 1.1920929e-7f;
        return t < eps ? eps : (t > 1.0f - eps ? 1.0f - eps : t);
    }

    static void segment_to(const Pt *p, SegKind kind, float start_t, float stop_t, PathBuilder &pb) {
        if (start_t == stop_t) {
            // a zero-length dash: a zero-length line, which the stroker caps
            Pt last;
            if (pb.last_point(last)) pb.line_to(last.x, last.y);
            return;
        }
        if (kind == kSegLine) {
            if (stop_t == 1.0f) pb.line_to(p[1].x, p[1].y);
            else pb.line_to(interp(p[0].x, p[1].x, stop_t), interp(p[0].y, p[1].y, stop_t));
            return;
        }
        if (kind == kSegQuad) {
            Pt t0[5], t1[5];
            if (start_t == 0.0f) {
                if (stop_t == 1.0f) {
                    pb.quad_to(p[1].x, p[1].y, p[2].x, p[2].y);
                } else {
                    chop_quad_at(p, exclusive(stop_t), t0);
                    pb.quad_to(t0[1].x, t0[1].y, t0[2].x, t0[2].y);
                }
            } else {
                chop_quad_at(p, exclusive(start_t), t0);
                if (stop_t == 1.0f) {
                    pb.quad_to(t0[3].x, t0[3].y, t0[4].x, t0[4].y);
                } else {
                    chop_quad_at(t0 + 2, exclusive((stop_t - start_t) / (1.0f - start_t)), t1);
                    pb.quad_to(t1[1].x, t1[1].y, t1[2].x, t1[2].y);
                }
            }
            return;
        }
        Pt t0[7], t1[7];
        if (start_t == 0.0f) {
            if (stop_t == 1.0f) {
                pb.cubic_to(p[1].x, p[1].y, p[2].x, p[2].y, p[3].x, p[3].y);
            } else {
                chop_cubic_at(p, exclusive(stop_t), t0);
                pb.cubic_to(t0[1].x, t0[1].y, t0[2].x, t0[2].y, t0[3].x, t0[3].y);
            }
        } else {
            chop_cubic_at(p, exclusive(start_t), t0);
            if (stop_t == 1.0f) {
                pb.cubic_to(t0[4].x, t0[4].y, t0[5].x, t0[5].y, t0[6].x, t0[6].y);
            } else {
                chop_cubic_at(t0 + 3, exclusive((stop_t - start_t) / (1.0f - start_t)), t1);
                pb.cubic_to(t1[1].x, t1[1].y, t1[2].x, t1[2].y, t1[3].x, t1[3].y);
            }
        }
    }

    // push_segment (SkContourMeasure::getSegment)
    void push_segment(float start_d, float stop_d, bool move_to, PathBuilder &pb) const {
        if (start_d < 0.0f) start_d = 0.0f;
        if (stop_d > length) stop_d = length;
        if (!(start_d <= stop_d)) return;
        if (segs.empty()) return;
        size_t si, ei;
        float start_t, stop_t;
        if (!distance_to_segment(start_d, si, start_t)) return;
        if (!distance_to_segment(stop_d, ei, stop_t)) return;
        Segment seg = segs[si];
        const Segment &stop_seg = segs[ei];
        if (move_to) {
            const Pt p = pos_at(&pts[seg.point_index], seg.kind, start_t);
            pb.move_to(p.x, p.y);
        }
        if (seg.point_index == stop_seg.point_index) {
            segment_to(&pts[seg.point_index], seg.kind, start_t, stop_t, pb);
            return;
        }
        size_t k = si;
        for (;;) {
            segment_to(&pts[seg.point_index], seg.kind, start_t, 1.0f, pb);
            const uint32_t old = seg.point_index;
            do {
                k++;
            } while (segs[k].point_index == old);
            seg = segs[k];
            start_t = 0.0f;
            if (seg.point_index >= stop_seg.point_index) break;
        }
        segment_to(&pts[seg.point_index], seg.kind, 0.0f, stop_t, pb);
    }
};

// ContourMeasureIter::next over a verb/point stream from `vi`, `pi`; false at the end
bool next_contour(const uint8_t *verbs, size_t nverbs, const Pt *pts, size_t &vi, size_t &pi, float tol,
                  Contour &c) {
    c = Contour{};
    if (vi >= nverbs) return false;
    uint32_t point_index = 0;
    float dist = 0.0f;
    bool seen_close = false, seen_move = false;
    while (vi < nverbs) {
        const uint8_t v = verbs[vi++];
        switch (v) {
            case kMove:
                c.pts.push_back(pts[pi++]);
                seen_move = true;
                break;
            case kLine: {
                if (c.pts.empty()) return false;
                const Pt p1 = pts[pi++];
                const float prev = dist;
                dist = c.line_seg(c.pts.back(), p1, dist, point_index);
                if (dist > prev) {
                    c.pts.push_back(p1);
                    point_index += 1;
                }
                break;
            }
            case kQuad: {
                if (c.pts.empty()) return false;
                const Pt p1 = pts[pi], p2 = pts[pi + 1];
                pi += 2;
                const float prev = dist;
                dist = c.quad_segs(c.pts.back(), p1, p2, dist, 0, kMaxTValue, point_index, tol);
                if (dist > prev) {
                    c.pts.push_back(p1);
                    c.pts.push_back(p2);
                    point_index += 2;
                }
                break;
            }
            case kCubic: {
                if (c.pts.empty()) return false;
                const Pt p1 = pts[pi], p2 = pts[pi + 1], p3 = pts[pi + 2];
                pi += 3;
                const float prev = dist;
                dist = c.cubic_segs(c.pts.back(), p1, p2, p3, dist, 0, kMaxTValue, point_index, tol);
                if (dist > prev) {
                    c.pts.push_back(p1);
                    c.pts.push_back(p2);
                    c.pts.push_back(p3);
                    point_index += 3;
                }
                break;
            }
            case kClose:
                seen_close = true;
                break;
            default:
                return false;
        }
        if (vi < nverbs && verbs[vi] == kMove) break;  // the next contour starts
    }
    (void)seen_move;
    if (!std::isfinite(dist)) return false;
    if (seen_close) {
        const float prev = dist;
        const Pt first = c.pts[0];
        dist = c.line_seg(c.pts[point_index], first, dist, point_index);
        if (dist > prev) c.pts.push_back(first);
    }
    c.length = dist;
    c.closed = seen_close;
    return !c.pts.empty();
}

}  // namespace

bool dash_params(const float *array, size_t n, float offset, DashParams &out) {
    if (!std::isfinite(offset)) return false;
    if (n < 2 || n % 2 != 0) return false;
    float len = 0.0f;
    for (size_t i = 0; i < n; i++) {
        if (array[i] < 0.0f) return false;
        len += array[i];
    }
    if (!(std::isfinite(len) && len > 0.0f && std::fpclassify(len) == FP_NORMAL)) return false;
    // adjust_dash_offset
    if (offset < 0.0f) {
        offset = -offset;
        if (offset > len) offset = std::fmod(offset, len);
        offset = len - offset;
        if (offset == len) offset = 0.0f;
    } else if (offset >= len) {
        offset = std::fmod(offset, len);
    }
    // find_first_interval
    float first_len = array[0];
    size_t first_index = 0;
    float ph = offset;
    bool found = false;
    for (size_t i = 0; i < n; i++) {
        const float gap = array[i];
        if (ph > gap || (ph == gap && gap != 0.0f)) {
            ph -= gap;
        } else {
            first_len = gap - ph;
            first_index = i;
            found = true;
            break;
        }
    }
    if (!found) {
        first_len = array[0];
        first_index = 0;
    }
    out.array.assign(array, array + n);
    out.offset = offset;
    out.interval_len = len;
    out.first_len = first_len;
    out.first_index = first_index;
    return true;
}

bool dash(const uint8_t *verbs, size_t nverbs, const float *points, size_t npoints, const DashParams &d,
          float res_scale, PathBuilder &out) {
    out.clear();
    std::vector<Pt> pts(npoints);
    for (size_t i = 0; i < npoints; i++) pts[i] = Pt{points[2 * i], points[2 * i + 1]};
    const float tol = kCheapDistLimit * (1.0f / res_scale);
    size_t vi = 0, pi = 0;
    float dash_count = 0.0f;
    Contour c;
    while (next_contour(verbs, nverbs, pts.data(), vi, pi, tol, c)) {
        bool skip_first = c.closed;
        bool added = false;
        const float length = c.length;
        size_t index = d.first_index;
        // more than a million dashes: give up (Skia's memory guard)
        dash_count += length * (float)(d.array.size() >> 1) / d.interval_len;
        if (dash_count > 1000000.0f) {
            out.clear();
            return false;
        }
        float distance = 0.0f;
        float dlen = d.first_len;
        while (distance < length) {
            added = false;
            if (index % 2 == 0 && !skip_first) {
                added = true;
                c.push_segment(distance, distance + dlen, true, out);
            }
            distance += dlen;
            skip_first = false;
            index += 1;
            if (index == d.array.size()) index = 0;
            dlen = d.array[index];
        }
        // join the last dash of a closed contour to the skipped first one
        if (c.closed && d.first_index % 2 == 0 && d.first_len >= 0.0f)
            c.push_segment(0.0f, d.first_len, !added, out);
    }
    // PathBuilder::finish: nothing, or a lone move, is no path
    return out.verbs.size() > 1;
}

}  // namespace rvks
