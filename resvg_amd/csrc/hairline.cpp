// Hairline geometry (tiny-skia 0.12.0 scan/hairline.rs, scan/hairline_aa.rs,
// line_clipper.rs, path_geometry.rs; Skia's SkScan_Hairline.cpp, SkScan_Antihair.cpp,
// SkLineClipper.cpp).  Float expressions keep the reference's operation order; build
// without FMA contraction.
//
// Restates tiny-skia (BSD-3-Clause, Copyright 2020 Yevhenii Reizner) and Skia
// (BSD-3-Clause, Copyright Google Inc.); see THIRD_PARTY_NOTICES.md.
#include "hairline.h"

#include <algorithm>
#include <cmath>
#include <utility>

namespace rvks {

namespace {

constexpr float kNearlyZero = 1.0f / 4096.0f;
constexpr int kMaxCubicLevel = 9;  // MAX_CUBIC_SUBDIVIDE_LEVEL
constexpr int kMaxQuadLevel = 5;   // MAX_QUAD_SUBDIVIDE_LEVEL

struct FRect {
    float l, t, r, b;
};

struct IRect {
    int32_t l, t, r, b;  // [l, r) x [t, b)
};

inline Pt sub(Pt a, Pt b) { return Pt{a.x - b.x, a.y - b.y}; }

// fdot6::from_f32: (v * 64) as i32 (saturating, NaN -> 0)
inline int32_t fdot6_from(float v) {
    const float f = v * 64.0f;
    if (!(f == f)) return 0;
    if (f >= 2147483647.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}
inline int32_t fdot6_floor(int32_t x) { return x >> 6; }
inline int32_t fdot6_ceil(int32_t x) { return (int32_t)(((int64_t)x + 63) >> 6); }

// ---- line_clipper::intersect (SkLineClipper::IntersectLine) ----
inline bool nested_lt(float a, float b, float dim) { return a <= b && (a < b || dim > 0.0f); }

double pin_unsorted(double v, double lo, double hi) {
    if (hi < lo) std::swap(lo, hi);
    if (v < lo) v = lo;
    else if (v > hi) v = hi;
    return v;
}

float sect_with_horizontal(const Pt s[2], float y) {
    const float dy = s[1].y - s[0].y;
    if (std::fabs(dy) <= kNearlyZero) return (s[0].x + s[1].x) * 0.5f;
    const double x0 = s[0].x, y0 = s[0].y, x1 = s[1].x, y1 = s[1].y;
    const double r = x0 + ((double)y - y0) * (x1 - x0) / (y1 - y0);
    return (float)pin_unsorted(r, x0, x1);
}

float sect_with_vertical(const Pt s[2], float x) {
    const float dx = s[1].x - s[0].x;
    if (std::fabs(dx) <= kNearlyZero) return (s[0].y + s[1].y) * 0.5f;
    const double x0 = s[0].x, y0 = s[0].y, x1 = s[1].x, y1 = s[1].y;
    return (float)(y0 + ((double)x - x0) * (y1 - y0) / (x1 - x0));
}

bool intersect_line(const Pt src[2], const FRect &clip, Pt dst[2]) {
    const float bl = std::min(src[0].x, src[1].x), bt = std::min(src[0].y, src[1].y);
    const float br = std::max(src[0].x, src[1].x), bb = std::max(src[0].y, src[1].y);
    if (clip.l <= bl && clip.t <= bt && clip.r >= br && clip.b >= bb) {
        dst[0] = src[0];
        dst[1] = src[1];
        return true;
    }
    const float bw = br - bl, bh = bb - bt;
    if (nested_lt(br, clip.l, bw) || nested_lt(clip.r, bl, bw) || nested_lt(bb, clip.t, bh) ||
        nested_lt(clip.b, bt, bh))
        return false;
    int i0 = src[0].y < src[1].y ? 0 : 1, i1 = 1 - i0;
    Pt tmp[2] = {src[0], src[1]};
    if (tmp[i0].y < clip.t) tmp[i0] = Pt{sect_with_horizontal(src, clip.t), clip.t};
    if (tmp[i1].y > clip.b) tmp[i1] = Pt{sect_with_horizontal(src, clip.b), clip.b};
    i0 = tmp[0].x < tmp[1].x ? 0 : 1;
    i1 = 1 - i0;
    if (tmp[i1].x <= clip.l || tmp[i0].x >= clip.r) {
        // a vertical line lying on the clip's edge is kept
        if (tmp[0].x != tmp[1].x || tmp[0].x < clip.l || tmp[0].x > clip.r) return false;
    }
    if (tmp[i0].x < clip.l) tmp[i0] = Pt{clip.l, sect_with_vertical(src, clip.l)};
    if (tmp[i1].x > clip.r) tmp[i1] = Pt{clip.r, sect_with_vertical(src, clip.r)};
    dst[0] = tmp[0];
    dst[1] = tmp[1];
    return true;
}

// ---- anti_hair_line_rgn: every line of a polyline to do_anti_hairline calls ----
struct Walk {
    IRect pixmap;
    std::vector<HairSeg> &out;

    void lines(const Pt *p, size_t n, bool clip) {
        const FRect fixed_bounds{-32767.0f, -32767.0f, 32767.0f, 32767.0f};
        // antialiased hairlines can draw up to half a pixel outside their bounds: the
        // scalar clip is outset by a full pixel
        const FRect clip_bounds{(float)pixmap.l - 1.0f, (float)pixmap.t - 1.0f, (float)pixmap.r + 1.0f,
                                (float)pixmap.b + 1.0f};
        for (size_t i = 0; i + 1 < n; i++) {
            Pt pts[2];
            if (!intersect_line(p + i, fixed_bounds, pts)) continue;
            if (clip) {
                const Pt tmp[2] = {pts[0], pts[1]};
                if (!intersect_line(tmp, clip_bounds, pts)) continue;
            }
            HairSeg s{};
            s.x0 = fdot6_from(pts[0].x);
            s.y0 = fdot6_from(pts[0].y);
            s.x1 = fdot6_from(pts[1].x);
            s.y1 = fdot6_from(pts[1].y);
            if (clip) {
                const IRect ir{fdot6_floor(std::min(s.x0, s.x1)) - 1, fdot6_floor(std::min(s.y0, s.y1)) - 1,
                               fdot6_ceil(std::max(s.x0, s.x1)) + 1, fdot6_ceil(std::max(s.y0, s.y1)) + 1};
                const IRect c{std::max(ir.l, pixmap.l), std::max(ir.t, pixmap.t), std::min(ir.r, pixmap.r),
                              std::min(ir.b, pixmap.b)};
                if (c.l >= c.r || c.t >= c.b) continue;  // quick reject
                if (!(pixmap.l <= ir.l && pixmap.t <= ir.t && pixmap.r >= ir.r && pixmap.b >= ir.b)) {
                    s.clipped = true;
                    s.cl = c.l; s.ct = c.t; s.cr = c.r; s.cb = c.b;
                }
            }
            out.push_back(s);
        }
    }
};

// ---- curves ----
struct Clip {
    bool on = false;        // the pixmap clip applies to the whole path
    FRect inset{}, outset{};
    bool inset_empty = false;
};

// compute_nocheck_quad/cubic_bounds + geometric_overlap / geometric_contains: false
// when the curve is culled; `clip` is cleared when the curve is well inside.
bool cull(const Pt *p, int n, const Clip &c, bool &clip) {
    if (!c.on) return true;
    float l = p[0].x, t = p[0].y, r = l, b = t;
    for (int i = 1; i < n; i++) {
        l = std::min(l, p[i].x); t = std::min(t, p[i].y);
        r = std::max(r, p[i].x); b = std::max(b, p[i].y);
    }
    if (!(std::isfinite(l) && std::isfinite(t) && std::isfinite(r) && std::isfinite(b))) return false;
    const FRect &o = c.outset;
    if (!(o.l < r && l < o.r && o.t < b && t < o.b)) return false;
    const FRect &in = c.inset;
    if (!c.inset_empty && r <= in.r && l >= in.l && b <= in.b && t >= in.t) clip = false;
    return true;
}

uint32_t compute_int_quad_dist(const Pt p[3]) {
    const float dx = std::fabs((p[0].x + p[2].x) * 0.5f - p[1].x);
    const float dy = std::fabs((p[0].y + p[2].y) * 0.5f - p[1].y);
    auto ceil_i = [](float v) -> uint32_t {  // SkScalarCeilToInt, non-negative here
        const float c = std::ceil(v);
        if (!(c == c)) return 0;
        if (c >= 2147483647.0f) return 2147483647u;
        return (uint32_t)(int32_t)c;
    };
    const uint32_t idx = ceil_i(dx), idy = ceil_i(dy);
    return idx > idy ? idx + (idy >> 1) : idy + (idx >> 1);
}

int compute_quad_level(const Pt p[3]) {
    const uint32_t d = compute_int_quad_dist(p);
    const int lz = d == 0 ? 32 : __builtin_clz(d);
    int level = (33 - lz) >> 1;
    return std::min(level, kMaxQuadLevel);
}

void hair_quad(const Pt p[3], bool clip, const Clip &c, int level, Walk &w) {
    if (!cull(p, 3, c, clip)) return;
    // QuadCoeff: a = p2 - 2 p1 + p0, b = 2 (p1 - p0), c = p0; evaluated (a t + b) t + c
    const Pt a{(p[2].x - (p[1].x + p[1].x)) + p[0].x, (p[2].y - (p[1].y + p[1].y)) + p[0].y};
    const Pt b{(p[1].x - p[0].x) + (p[1].x - p[0].x), (p[1].y - p[0].y) + (p[1].y - p[0].y)};
    const int lines = 1 << level;
    Pt tmp[(1 << kMaxQuadLevel) + 1];
    tmp[0] = p[0];
    const float dt = 1.0f / (float)lines;
    float t = 0.0f;
    for (int i = 1; i < lines; i++) {
        t = t + dt;
        tmp[i] = Pt{(a.x * t + b.x) * t + p[0].x, (a.y * t + b.y) * t + p[0].y};
    }
    tmp[lines] = p[2];
    w.lines(tmp, (size_t)lines + 1, clip);
}

int compute_cubic_segments(const Pt p[4]) {
    const float one_third = 1.0f / 3.0f, two_third = 2.0f / 3.0f;
    const Pt p13{one_third * p[3].x + two_third * p[0].x, one_third * p[3].y + two_third * p[0].y};
    const Pt p23{one_third * p[0].x + two_third * p[3].x, one_third * p[0].y + two_third * p[3].y};
    const float mx = std::max(std::fabs(p[1].x - p13.x), std::fabs(p[2].x - p23.x));
    const float my = std::max(std::fabs(p[1].y - p13.y), std::fabs(p[2].y - p23.y));
    const float diff = std::max(mx, my);
    float tol = 1.0f / 8.0f;
    for (int i = 0; i < kMaxCubicLevel; i++) {
        if (diff < tol) return 1 << i;
        tol *= 4.0f;
    }
    return 1 << kMaxCubicLevel;
}

inline bool lt_90(Pt p0, Pt pivot, Pt p2) {
    const Pt u = sub(p0, pivot), v = sub(p2, pivot);
    return u.x * v.x + u.y * v.y >= 0.0f;
}

void hair_cubic2(const Pt p[4], bool clip, Walk &w) {
    const int lines = compute_cubic_segments(p);
    if (lines == 1) {
        const Pt l[2] = {p[0], p[3]};
        w.lines(l, 2, clip);
        return;
    }
    // CubicCoeff: a = p3 + 3 (p1 - p2) - p0, b = 3 (p2 - 2 p1 + p0), c = 3 (p1 - p0), d = p0
    const Pt a{(p[3].x + 3.0f * (p[1].x - p[2].x)) - p[0].x, (p[3].y + 3.0f * (p[1].y - p[2].y)) - p[0].y};
    const Pt b{3.0f * ((p[2].x - (p[1].x + p[1].x)) + p[0].x), 3.0f * ((p[2].y - (p[1].y + p[1].y)) + p[0].y)};
    const Pt cc{3.0f * (p[1].x - p[0].x), 3.0f * (p[1].y - p[0].y)};
    std::vector<Pt> tmp((size_t)lines + 1);
    tmp[0] = p[0];
    const float dt = 1.0f / (float)lines;
    float t = 0.0f;
    for (int i = 1; i < lines; i++) {
        t = t + dt;
        tmp[i] = Pt{((a.x * t + b.x) * t + cc.x) * t + p[0].x, ((a.y * t + b.y) * t + cc.y) * t + p[0].y};
    }
    for (int i = 0; i < lines; i++)
        if (!std::isfinite(tmp[i].x) || !std::isfinite(tmp[i].y)) return;
    tmp[lines] = p[3];
    w.lines(tmp.data(), tmp.size(), clip);
}

// chop_cubic_at (several t values): the pieces share end points; a t that cannot be
// renormalised ends the chain with a degenerate cubic at the last point
void chop_cubic_at_values(const Pt src[4], const float *t, int n, Pt *dst) {
    Pt tmp[4];
    const Pt *s = src;
    float tt = t[0];
    size_t off = 0;
    for (int i = 0; i < n; i++) {
        chop_cubic_at(s, tt, dst + off);
        if (i == n - 1) break;
        off += 3;
        for (int k = 0; k < 4; k++) tmp[k] = dst[off + k];
        s = tmp;
        float r;
        if (unit_divide(t[i + 1] - t[i], 1.0f - t[i], r)) {
            tt = r;
        } else {
            dst[off + 4] = s[3];
            dst[off + 5] = s[3];
            dst[off + 6] = s[3];
            break;
        }
    }
}

void hair_cubic(const Pt p[4], bool clip, const Clip &c, Walk &w) {
    if (!cull(p, 4, c, clip)) return;
    const bool nice = lt_90(p[1], p[0], p[3]) && lt_90(p[2], p[0], p[3]) && lt_90(p[1], p[3], p[0]) &&
                      lt_90(p[2], p[3], p[0]);
    if (nice) {
        hair_cubic2(p, clip, w);
        return;
    }
    // chop_cubic_at_max_curvature
    float roots[3], ts[3];
    const int nr = cubic_max_curvature(p, roots);
    int n = 0;
    for (int i = 0; i < nr; i++)
        if (0.0f < roots[i] && roots[i] < 1.0f) ts[n++] = roots[i];
    Pt tmp[13];
    if (n == 0) {
        for (int k = 0; k < 4; k++) tmp[k] = p[k];
    } else {
        chop_cubic_at_values(p, ts, n, tmp);
    }
    for (int i = 0; i <= n; i++) hair_cubic2(tmp + 3 * i, clip, w);
}

// extend_pts: moves the end points of a segment that starts or ends a contour along
// its tangent by the cap's outset (square 1/2, round pi/8)
void extend_pts(uint32_t cap, uint8_t prev_verb, int next_verb, Pt *pts, int n) {
    const float outset = cap == 2 ? 0.5f : 3.14159265f / 8.0f;
    if (prev_verb == kMove) {
        int ctrl = 0, controls = n - 1;
        Pt tangent;
        do {
            tangent = sub(pts[0], pts[++ctrl]);
        } while (tangent.x == 0.0f && tangent.y == 0.0f && --controls > 0);
        if (tangent.x == 0.0f && tangent.y == 0.0f) {
            tangent = Pt{1.0f, 0.0f};
            controls = n - 1;
        } else {
            const double m = std::sqrt((double)tangent.x * tangent.x + (double)tangent.y * tangent.y);
            const double sc = 1.0 / m;
            tangent = Pt{tangent.x * (float)sc, tangent.y * (float)sc};
        }
        int k = 0;
        do {
            pts[k].x += tangent.x * outset;
            pts[k].y += tangent.y * outset;
            k++;
        } while (++controls < n);
    }
    if (next_verb < 0 || next_verb == kMove || next_verb == kClose) {
        int ctrl = n - 1, controls = n - 1;
        Pt tangent;
        do {
            tangent = sub(pts[n - 1], pts[--ctrl]);
        } while (tangent.x == 0.0f && tangent.y == 0.0f && --controls > 0);
        if (tangent.x == 0.0f && tangent.y == 0.0f) {
            tangent = Pt{-1.0f, 0.0f};
            controls = n - 1;
        } else {
            const double m = std::sqrt((double)tangent.x * tangent.x + (double)tangent.y * tangent.y);
            const double sc = 1.0 / m;
            tangent = Pt{tangent.x * (float)sc, tangent.y * (float)sc};
        }
        int k = n - 1;
        do {
            pts[k].x += tangent.x * outset;
            pts[k].y += tangent.y * outset;
            k--;
        } while (++controls < n);
    }
}

}  // namespace

float hairline_paint_opacity(float coverage) {
    if (coverage == 1.0f) return 1.0f;
    const int32_t scale = (int32_t)(coverage * 256.0f);
    const int32_t new_alpha = (255 * scale) >> 8;
    return (float)new_alpha / 255.0f;
}

void anti_hairline_segments(const uint8_t *verbs, size_t nverbs, const Pt *pts, size_t npts, uint32_t cap,
                            uint32_t w, uint32_t h, std::vector<HairSeg> &out) {
    if (nverbs == 0 || npts == 0) return;
    // Path::bounds of the device-space points; a non-finite path draws nothing
    float l = pts[0].x, t = pts[0].y, r = l, b = t;
    for (size_t i = 1; i < npts; i++) {
        l = std::min(l, pts[i].x); t = std::min(t, pts[i].y);
        r = std::max(r, pts[i].x); b = std::max(b, pts[i].y);
    }
    if (!(std::isfinite(l) && std::isfinite(t) && std::isfinite(r) && std::isfinite(b))) return;
    const IRect pix{0, 0, (int32_t)w, (int32_t)h};
    Clip c;
    {
        const float cap_out = cap == 0 ? 1.0f : 2.0f;
        auto fl = [](float v) { return (int64_t)std::floor(v); };
        auto ce = [](float v) { return (int64_t)std::ceil(v); };
        const int64_t il = fl(l - cap_out), it = fl(t - cap_out), ir = ce(r + cap_out), ib = ce(b + cap_out);
        if (std::max<int64_t>(il, pix.l) >= std::min<int64_t>(ir, pix.r) ||
            std::max<int64_t>(it, pix.t) >= std::min<int64_t>(ib, pix.b))
            return;  // quick reject
        if (!(pix.l <= il && pix.t <= it && pix.r >= ir && pix.b >= ib)) {
            c.on = true;
            c.outset = FRect{(float)pix.l - 1.0f, (float)pix.t - 1.0f, (float)pix.r + 1.0f, (float)pix.b + 1.0f};
            c.inset = FRect{(float)pix.l + 1.0f, (float)pix.t + 1.0f, (float)pix.r - 1.0f, (float)pix.b - 1.0f};
            c.inset_empty = !(c.inset.l < c.inset.r && c.inset.t < c.inset.b);
        }
    }
    Walk wk{pix, out};
    uint8_t prev_verb = kMove;
    Pt first{0, 0}, last{0, 0};
    size_t pi = 0;
    for (size_t vi = 0; vi < nverbs; vi++) {
        const uint8_t verb = verbs[vi];
        const int next_verb = vi + 1 < nverbs ? verbs[vi + 1] : -1;
        Pt lead{0, 0};  // the segment's (possibly extended) first point
        switch (verb) {
            case kMove:
                if (pi >= npts) return;
                first = last = pts[pi++];
                lead = first;
                break;
            case kLine: {
                if (pi + 1 > npts) return;
                Pt p[2] = {last, pts[pi++]};
                if (cap != 0) extend_pts(cap, prev_verb, next_verb, p, 2);
                wk.lines(p, 2, c.on);
                last = pts[pi - 1];
                lead = p[0];
                break;
            }
            case kQuad: {
                if (pi + 2 > npts) return;
                Pt p[3] = {last, pts[pi], pts[pi + 1]};
                pi += 2;
                if (cap != 0) extend_pts(cap, prev_verb, next_verb, p, 3);
                hair_quad(p, c.on, c, compute_quad_level(p), wk);
                last = pts[pi - 1];
                lead = p[0];
                break;
            }
            case kCubic: {
                if (pi + 3 > npts) return;
                Pt p[4] = {last, pts[pi], pts[pi + 1], pts[pi + 2]};
                pi += 3;
                if (cap != 0) extend_pts(cap, prev_verb, next_verb, p, 4);
                hair_cubic(p, c.on, c, wk);
                last = pts[pi - 1];
                lead = p[0];
                break;
            }
            case kClose: {
                Pt p[2] = {last, first};
                if (cap != 0 && prev_verb == kMove) {  // a lone point: a horizontal dash
                    p[0].x += 0.5f;
                    p[1].x += -0.5f;
                }
                wk.lines(p, 2, c.on);
                break;
            }
            default:
                return;
        }
        if (cap != 0) {
            // the curve moved the initial point, so close to it instead
            if (prev_verb == kMove && (verb == kLine || verb == kQuad || verb == kCubic)) first = lead;
            prev_verb = verb;
        }
    }
}

}  // namespace rvks
