// Edge clipping of a fill whose bounds leave the pixmap: tiny-skia's EdgeClipper
// (tiny-skia 0.11 scan/edge_clipper.rs, scan/line_clipper.rs; Skia's SkEdgeClipper
// and SkLineClipper) with the double-precision cubic/axis intersection it takes from
// Skia's path ops (tiny-skia-path path64/: SkDCubic::RootsValidT, RootsReal,
// FindExtrema, searchRoots, binarySearch, chopAt; SkDQuad::RootsReal, AddValidTs;
// the approximately_* tolerances of SkPathOpsTypes).
//
// Every segment is clipped on its own against [0, w] x [0, h] in device pixels:
// parts above or below are dropped, parts left or right become vertical lines on
// the clip edge (nothing is culled to the right), curves are first chopped at their
// Y and X extrema.  Float expressions keep the reference's operation order.
//
// Restates tiny-skia (BSD-3-Clause, Copyright 2020 Yevhenii Reizner) and Skia
// (BSD-3-Clause, Copyright Google Inc.); see THIRD_PARTY_NOTICES.md.
#include "raster.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace rvkr {
namespace {

constexpr float kNearlyZero = 1.0f / 4096.0f;  // SCALAR_NEARLY_ZERO

// ---- path ops tolerances (SkPathOpsTypes.h) ----
constexpr double kFltEps = (double)FLT_EPSILON;
constexpr double kFltEpsCubed = kFltEps * kFltEps * kFltEps;
constexpr double kFltEpsHalf = kFltEps / 2;
constexpr double kFltEpsInverse = 1 / kFltEps;

inline bool approx_zero(double x) { return std::fabs(x) < kFltEps; }
inline bool approx_zero_cubed(double x) { return std::fabs(x) < kFltEpsCubed; }
inline bool approx_zero_inverse(double x) { return std::fabs(x) > kFltEpsInverse; }
inline bool approx_zero_half(double x) { return x < kFltEpsHalf; }  // no fabs, as in Skia
inline bool approx_zero_when_compared_to(double x, double y) { return x == 0 || std::fabs(x) < std::fabs(y * kFltEps); }
inline bool approx_equal(double x, double y) { return approx_zero(x - y); }
inline bool approx_equal_half(double x, double y) { return approx_zero_half(x - y); }
inline bool approx_one_or_less(double x) { return x < 1 + kFltEps; }
inline bool approx_zero_or_more(double x) { return x > -kFltEps; }
inline bool between(double a, double b, double c) { return (a - b) * (c - b) <= 0; }
inline double bound(double v, double lo, double hi) { return std::max(lo, std::min(hi, v)); }

// float as a two's-complement-ordered integer (SkFloatAs2sCompliment)
inline int32_t float_as_2s(float f) {
    int32_t bits;
    std::memcpy(&bits, &f, 4);
    if (bits < 0) {
        bits &= 0x7FFFFFFF;
        bits = -bits;
    }
    return bits;
}
// AlmostDequalUlps (16 ulps of float, relative error for values beyond float range)
inline bool almost_dequal_ulps(double a, double b) {
    if (std::fabs(a) < FLT_MAX && std::fabs(b) < FLT_MAX) {
        const int32_t ab = float_as_2s((float)a), bb = float_as_2s((float)b);
        return ab < bb + 16 && bb < ab + 16;
    }
    return std::fabs(a - b) / std::max(std::fabs(a), std::fabs(b)) < kFltEps * 16;
}

// SkDQuad::RootsReal: real roots of a t^2 + b t + c
int quad_roots_real(double a, double b, double c, double s[3]) {
    auto handle_zero = [&]() {
        if (approx_zero(b)) {
            s[0] = 0;
            return c == 0 ? 1 : 0;
        }
        s[0] = -c / b;
        return 1;
    };
    if (a == 0) return handle_zero();
    const double p = b / (2 * a);
    const double q = c / a;
    if (approx_zero(a) && (approx_zero_inverse(p) || approx_zero_inverse(q))) return handle_zero();
    const double p2 = p * p;
    if (!almost_dequal_ulps(p2, q) && p2 < q) return 0;
    double sqrt_d = 0;
    if (p2 > q) sqrt_d = std::sqrt(p2 - q);
    s[0] = sqrt_d - p;
    s[1] = -sqrt_d - p;
    return 1 + (almost_dequal_ulps(s[0], s[1]) ? 0 : 1);
}

// SkDQuad::AddValidTs: roots within [0, 1] (pinned), without near duplicates
int push_valid_ts(const double *s, int real_roots, double *t) {
    int found = 0;
    for (int i = 0; i < real_roots; i++) {
        double tv = s[i];
        if (!(approx_zero_or_more(tv) && approx_one_or_less(tv))) continue;
        tv = bound(tv, 0, 1);
        bool dup = false;
        for (int k = 0; k < found; k++)
            if (approx_equal(t[k], tv)) { dup = true; break; }
        if (!dup) t[found++] = tv;
    }
    return found;
}

int quad_roots_valid_t(double a, double b, double c, double t[2]) {
    double s[3];
    const int n = quad_roots_real(a, b, c, s);
    return push_valid_ts(s, n, t);
}

// SkDCubeRoot: Kahan's bit-hack estimate refined by three Halley steps
double cube_root(double x) {
    if (approx_zero_cubed(x)) return 0;
    const double d = std::fabs(x);
    uint64_t bits;
    std::memcpy(&bits, &d, 8);
    const uint32_t hi = (uint32_t)(bits >> 32) / 3 + 715094163u;
    const uint64_t tb = (uint64_t)hi << 32;
    double a;
    std::memcpy(&a, &tb, 8);
    for (int i = 0; i < 3; i++) {
        const double a3 = a * a * a;
        a = a * (a3 + d + d) / (a3 + a3 + d);
    }
    return x < 0 ? -a : a;
}

// SkDCubic::RootsReal: real roots of A t^3 + B t^2 + C t + D
int cubic_roots_real(double A, double B, double C, double D, double s[3]) {
    if (approx_zero(A) && approx_zero_when_compared_to(A, B) && approx_zero_when_compared_to(A, C) &&
        approx_zero_when_compared_to(A, D))
        return quad_roots_real(B, C, D, s);
    if (approx_zero_when_compared_to(D, A) && approx_zero_when_compared_to(D, B) && approx_zero_when_compared_to(D, C)) {
        int num = quad_roots_real(A, B, C, s);
        for (int i = 0; i < num; i++)
            if (approx_zero(s[i])) return num;
        s[num++] = 0;
        return num;
    }
    if (approx_zero(A + B + C + D)) {
        int num = quad_roots_real(A, A + B, -D, s);
        for (int i = 0; i < num; i++)
            if (almost_dequal_ulps(s[i], 1)) return num;
        s[num++] = 1;
        return num;
    }
    const double inv_a = 1 / A;
    const double a = B * inv_a, b = C * inv_a, c = D * inv_a;
    const double a2 = a * a;
    const double Q = (a2 - b * 3) / 9;
    const double R = (2 * a2 * a - 9 * a * b + 27 * c) / 54;
    const double R2 = R * R;
    const double Q3 = Q * Q * Q;
    const double R2MinusQ3 = R2 - Q3;
    const double adiv3 = a / 3;
    int n = 0;
    if (R2MinusQ3 < 0) {  // three real roots
        const double theta = std::acos(bound(R / std::sqrt(Q3), -1, 1));
        const double neg2RootQ = -2 * std::sqrt(Q);
        const double pi = 3.141592653589793;
        s[n++] = neg2RootQ * std::cos(theta / 3) - adiv3;
        double rr = neg2RootQ * std::cos((theta + 2 * pi) / 3) - adiv3;
        if (!almost_dequal_ulps(s[0], rr)) s[n++] = rr;
        rr = neg2RootQ * std::cos((theta - 2 * pi) / 3) - adiv3;
        if (!almost_dequal_ulps(s[0], rr) && (n == 1 || !almost_dequal_ulps(s[1], rr))) s[n++] = rr;
    } else {  // one real root
        const double sqrtR2MinusQ3 = std::sqrt(R2MinusQ3);
        double A2 = std::fabs(R) + sqrtR2MinusQ3;
        A2 = cube_root(A2);
        if (R > 0) A2 = -A2;
        if (A2 != 0) A2 += Q / A2;
        s[n++] = A2 - adiv3;
        if (almost_dequal_ulps(R2, Q3)) {
            const double r = -A2 / 2 - adiv3;
            if (!almost_dequal_ulps(s[0], r)) s[n++] = r;
        }
    }
    return n;
}

// SkDCubic::RootsValidT: roots in [0, 1], with roots just outside snapped to 0 or 1
int cubic_roots_valid_t(double A, double B, double C, double D, double t[3]) {
    double s[3];
    const int real = cubic_roots_real(A, B, C, D, s);
    int found = push_valid_ts(s, real, t);
    for (int i = 0; i < real; i++) {
        const double tv = s[i];
        if (!approx_one_or_less(tv) && between(1, tv, 1.00005)) {
            bool dup = false;
            for (int k = 0; k < found; k++)
                if (approx_equal(t[k], 1)) { dup = true; break; }
            if (!dup) t[found++] = 1;
        } else if (!approx_zero_or_more(tv) && between(-0.00005, tv, 0)) {
            bool dup = false;
            for (int k = 0; k < found; k++)
                if (approx_equal(t[k], 0)) { dup = true; break; }
            if (!dup) t[found++] = 0;
        }
    }
    return found;
}

struct DPt { double x, y; };
struct DCubic {
    DPt p[4];
    double c(int i, int axis) const { return axis ? p[i].y : p[i].x; }
    // SkDCubic::ptAtT
    DPt at(double t) const {
        if (t == 0) return p[0];
        if (t == 1) return p[3];
        const double one_t = 1 - t;
        const double one_t2 = one_t * one_t;
        const double a = one_t2 * one_t;
        const double b = 3 * one_t2 * t;
        const double t2 = t * t;
        const double c = 3 * one_t * t2;
        const double d = t2 * t;
        return DPt{a * p[0].x + b * p[1].x + c * p[2].x + d * p[3].x, a * p[0].y + b * p[1].y + c * p[2].y + d * p[3].y};
    }
    // SkDCubic::findInflections
    int inflections(double *t) const {
        const double Ax = p[1].x - p[0].x, Ay = p[1].y - p[0].y;
        const double Bx = p[2].x - 2 * p[1].x + p[0].x, By = p[2].y - 2 * p[1].y + p[0].y;
        const double Cx = p[3].x + 3 * (p[1].x - p[2].x) - p[0].x, Cy = p[3].y + 3 * (p[1].y - p[2].y) - p[0].y;
        return quad_roots_valid_t(Bx * Cy - By * Cx, Ax * Cy - Ay * Cx, Ax * By - Ay * Bx, t);
    }
    // SkDCubic::binarySearch along `axis` (0 x, 1 y); -1 when no point meets it
    double binary_search(double mn, double mx, double axis_intercept, int axis) const {
        double t = (mn + mx) / 2;
        double step = (t - mn) / 2;
        DPt at_t = at(t);
        double calc_pos = axis ? at_t.y : at_t.x;
        double calc_dist = calc_pos - axis_intercept;
        do {
            const double prior_t = std::max(mn, t - step);
            const DPt less = at(prior_t);
            if (approx_equal_half(less.x, at_t.x) && approx_equal_half(less.y, at_t.y)) return -1;
            const double less_dist = (axis ? less.y : less.x) - axis_intercept;
            const double last_step = step;
            step /= 2;
            if (calc_dist > 0 ? calc_dist > less_dist : calc_dist < less_dist) {
                t = prior_t;
            } else {
                const double next_t = t + last_step;
                if (next_t > mx) return -1;
                const DPt more = at(next_t);
                if (approx_equal_half(more.x, at_t.x) && approx_equal_half(more.y, at_t.y)) return -1;
                const double more_dist = (axis ? more.y : more.x) - axis_intercept;
                if (calc_dist > 0 ? calc_dist <= more_dist : calc_dist >= more_dist) continue;
                t = next_t;
            }
            at_t = at(t);
            calc_pos = axis ? at_t.y : at_t.x;
            calc_dist = calc_pos - axis_intercept;
        } while (!approx_equal(calc_pos, axis_intercept));
        return t;
    }
    // SkDCubic::searchRoots
    int search_roots(double *ext, int extrema, double axis_intercept, int axis, double *roots) const {
        extrema += inflections(ext + extrema);
        ext[extrema++] = 0;
        ext[extrema] = 1;
        std::sort(ext, ext + extrema + 1);
        int valid = 0;
        for (int i = 0; i < extrema;) {
            const double mn = ext[i];
            const double mx = ext[++i];
            if (mn == mx) continue;
            const double nt = binary_search(mn, mx, axis_intercept, axis);
            if (nt >= 0) {
                if (valid >= 3) return 0;
                roots[valid++] = nt;
            }
        }
        return valid;
    }
    // LineCubicIntersections::HorizontalIntersect / VerticalIntersect
    int axis_intersect(double axis_intercept, int axis, double roots[3]) const {
        // SkDCubic::Coefficients
        double A = c(3, axis), B = c(2, axis) * 3, C = c(1, axis) * 3;
        const double D0 = c(0, axis);
        A -= D0 - C + B;
        B += 3 * D0 - 2 * C;
        C -= 3 * D0;
        const double D = D0 - axis_intercept;
        int count = cubic_roots_valid_t(A, B, C, D, roots);
        for (int i = 0; i < count; i++) {
            const DPt pt = at(roots[i]);
            if (!approx_equal(axis ? pt.y : pt.x, axis_intercept)) {
                // SkDCubic::FindExtrema, then the search between extrema and inflections
                double ext[6];
                const double a = c(0, axis), b = c(1, axis), cc = c(2, axis), d = c(3, axis);
                const int extrema = quad_roots_valid_t(d - a + 3 * (b - cc), 2 * (a - b - b + cc), b - a, ext);
                count = search_roots(ext, extrema, axis_intercept, axis, roots);
                break;
            }
        }
        return count;
    }
    // SkDCubic::chopAt
    void chop_at(double t, DPt dst[7]) const {
        if (t == 0.5) {
            dst[0] = p[0];
            dst[1] = DPt{(p[0].x + p[1].x) / 2, (p[0].y + p[1].y) / 2};
            dst[2] = DPt{(p[0].x + 2 * p[1].x + p[2].x) / 4, (p[0].y + 2 * p[1].y + p[2].y) / 4};
            dst[3] = DPt{(p[0].x + 3 * (p[1].x + p[2].x) + p[3].x) / 8, (p[0].y + 3 * (p[1].y + p[2].y) + p[3].y) / 8};
            dst[4] = DPt{(p[1].x + 2 * p[2].x + p[3].x) / 4, (p[1].y + 2 * p[2].y + p[3].y) / 4};
            dst[5] = DPt{(p[2].x + p[3].x) / 2, (p[2].y + p[3].y) / 2};
            dst[6] = p[3];
            return;
        }
        auto interp = [](double a, double b, double t) { return a + (b - a) * t; };
        for (int axis = 0; axis < 2; axis++) {
            const double s0 = c(0, axis), s1 = c(1, axis), s2 = c(2, axis), s3 = c(3, axis);
            const double ab = interp(s0, s1, t), bc = interp(s1, s2, t), cd = interp(s2, s3, t);
            const double abc = interp(ab, bc, t), bcd = interp(bc, cd, t), abcd = interp(abc, bcd, t);
            const double v[7] = {s0, ab, abc, abcd, bcd, cd, s3};
            for (int i = 0; i < 7; i++) (axis ? dst[i].y : dst[i].x) = v[i];
        }
    }
};

// ---- tiny-skia-path path_geometry (f32) ----
bool unit_divide(float numer, float denom, float &r) {
    if (numer < 0.0f) { numer = -numer; denom = -denom; }
    if (denom == 0.0f || numer == 0.0f || numer >= denom) return false;
    r = numer / denom;
    return std::isfinite(r) && r > 0.0f && r < 1.0f;
}

int unit_quad_roots(float a, float b, float c, float *roots) {
    if (a == 0.0f) {
        float r;
        if (unit_divide(-c, b, r)) { roots[0] = r; return 1; }
        return 0;
    }
    double dr = (double)b * (double)b - 4.0 * (double)a * (double)c;
    if (dr < 0.0) return 0;
    const float r = (float)std::sqrt(dr);
    if (!std::isfinite(r)) return 0;
    const float q = b < 0.0f ? -(b - r) / 2.0f : -(b + r) / 2.0f;
    int n = 0;
    float t;
    if (unit_divide(q, a, t)) roots[n++] = t;
    if (unit_divide(c, q, t)) roots[n++] = t;
    if (n == 2) {
        if (roots[0] > roots[1]) std::swap(roots[0], roots[1]);
        else if (roots[0] == roots[1]) n = 1;
    }
    return n;
}

inline float& ax(Pt &p, int axis) { return axis ? p.y : p.x; }
inline float ax(const Pt &p, int axis) { return axis ? p.y : p.x; }
inline Pt lerp(Pt a, Pt b, float t) { return Pt{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t}; }

void chop_quad(const Pt *s, float t, Pt *d) {
    const Pt p01 = lerp(s[0], s[1], t), p12 = lerp(s[1], s[2], t);
    d[0] = s[0]; d[1] = p01; d[2] = lerp(p01, p12, t); d[3] = p12; d[4] = s[2];
}

void chop_cubic(const Pt *s, float t, Pt *d) {
    const Pt ab = lerp(s[0], s[1], t), bc = lerp(s[1], s[2], t), cd = lerp(s[2], s[3], t);
    const Pt abc = lerp(ab, bc, t), bcd = lerp(bc, cd, t);
    d[0] = s[0]; d[1] = ab; d[2] = abc; d[3] = lerp(abc, bcd, t); d[4] = bcd; d[5] = cd; d[6] = s[3];
}

// chop_quad_at_{x,y}_extrema
int quad_monotonic(const Pt *s, Pt *d, int axis) {
    const float a = ax(s[0], axis), c = ax(s[2], axis);
    float b = ax(s[1], axis);
    float ab = a - b, bc = b - c;
    if (ab < 0.0f) bc = -bc;
    if (ab == 0.0f || bc < 0.0f) {
        float t;
        if (unit_divide(a - b, a - b - b + c, t)) {
            chop_quad(s, t, d);
            ax(d[1], axis) = ax(d[2], axis);
            ax(d[3], axis) = ax(d[2], axis);
            return 1;
        }
        b = std::fabs(a - b) < std::fabs(b - c) ? a : c;
    }
    d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
    ax(d[0], axis) = a; ax(d[1], axis) = b; ax(d[2], axis) = c;
    return 0;
}

// chop_cubic_at_{x,y}_extrema
int cubic_monotonic(const Pt *src, Pt *dst, int axis) {
    const float a = ax(src[0], axis), b = ax(src[1], axis), c = ax(src[2], axis), d = ax(src[3], axis);
    float ts[3];
    const int n = unit_quad_roots(d - a + 3.0f * (b - c), 2.0f * (a - b - b + c), b - a, ts);
    if (n == 0) {
        for (int i = 0; i < 4; i++) dst[i] = src[i];
        return 0;
    }
    Pt tmp[4];
    const Pt *s = src;
    float t = ts[0];
    int off = 0;
    for (int i = 0; i < n; i++) {
        chop_cubic(s, t, dst + off);
        if (i == n - 1) break;
        off += 3;
        for (int k = 0; k < 4; k++) tmp[k] = dst[off + k];
        s = tmp;
        float nt;
        if (unit_divide(ts[i + 1] - ts[i], 1.0f - ts[i], nt)) t = nt;
        else {
            dst[off + 4] = s[3]; dst[off + 5] = s[3]; dst[off + 6] = s[3];
            break;
        }
    }
    ax(dst[2], axis) = ax(dst[3], axis);
    ax(dst[4], axis) = ax(dst[3], axis);
    if (n == 2) { ax(dst[5], axis) = ax(dst[6], axis); ax(dst[7], axis) = ax(dst[6], axis); }
    return n;
}

// chop_mono_cubic_at_{x,y}: the double-precision intercept, else the closest t of
// a bisection on the f32 polynomial (mono_cubic_closest_t)
void chop_mono_cubic_at(const Pt *src, float v, int axis, Pt dst[7]) {
    DCubic dc;
    for (int i = 0; i < 4; i++) dc.p[i] = DPt{(double)src[i].x, (double)src[i].y};
    double roots[3];
    if (dc.axis_intersect((double)v, axis, roots) > 0) {
        DPt d[7];
        dc.chop_at(roots[0], d);
        for (int i = 0; i < 7; i++) dst[i] = Pt{(float)d[i].x, (float)d[i].y};
        return;
    }
    float t = 0.5f, last_t, best_t = t, step = 0.25f;
    const float d = ax(src[0], axis);
    const float a = ax(src[3], axis) + 3.0f * (ax(src[1], axis) - ax(src[2], axis)) - d;
    const float b = 3.0f * (ax(src[2], axis) - ax(src[1], axis) - ax(src[1], axis) + d);
    const float c = 3.0f * (ax(src[1], axis) - d);
    const float x = v - d;
    float closest = FLT_MAX;
    do {
        const float loc = ((a * t + b) * t + c) * t;
        const float dist = std::fabs(loc - x);
        if (closest > dist) { closest = dist; best_t = t; }
        last_t = t;
        t += loc < x ? step : -step;
        step *= 0.5f;
    } while (closest > 0.25f && last_t != t);
    chop_cubic(src, best_t, dst);
}

// chop_mono_quad_at: the first unit root of the quadratic through c0, c1, c2 = target
bool chop_mono_quad_at(float c0, float c1, float c2, float target, float &t) {
    const float a = c0 - c1 - c1 + c2;
    const float b = 2.0f * (c1 - c0);
    const float c = c0 - target;
    float roots[3];
    if (unit_quad_roots(a, b, c, roots) != 0) { t = roots[0]; return true; }
    return false;
}

struct Clipper {
    float left, top, right, bottom;
    std::vector<Seg> &out;

    void line(Pt a, Pt b) { out.push_back(Seg{2, {a, b}}); }
    void vline(float x, float y0, float y1, bool reverse) {
        if (reverse) std::swap(y0, y1);
        line(Pt{x, y0}, Pt{x, y1});
    }
    void quad(const Pt *p, bool reverse) {
        out.push_back(reverse ? Seg{3, {p[2], p[1], p[0]}} : Seg{3, {p[0], p[1], p[2]}});
    }
    void cubic(const Pt *p, bool reverse) {
        out.push_back(reverse ? Seg{4, {p[3], p[2], p[1], p[0]}} : Seg{4, {p[0], p[1], p[2], p[3]}});
    }

    // ---- line_clipper::clip ----
    static float sect_with_horizontal(const Pt *s, float y) {
        const float dy = s[1].y - s[0].y;
        if (std::fabs(dy) <= kNearlyZero) return (s[0].x + s[1].x) * 0.5f;
        const double x0 = s[0].x, y0 = s[0].y, x1 = s[1].x, y1 = s[1].y;
        const double r = x0 + ((double)y - y0) * (x1 - x0) / (y1 - y0);
        const double lo = std::min(x0, x1), hi = std::max(x0, x1);  // pin_unsorted
        return (float)(r < lo ? lo : (r > hi ? hi : r));
    }
    static float sect_clamp_with_vertical(const Pt *s, float x) {
        const float dx = s[1].x - s[0].x;
        float y;
        if (std::fabs(dx) <= kNearlyZero) y = (s[0].y + s[1].y) * 0.5f;
        else {
            const double x0 = s[0].x, y0 = s[0].y, x1 = s[1].x, y1 = s[1].y;
            y = (float)(y0 + ((double)x - x0) * (y1 - y0) / (x1 - x0));
        }
        const float lo = std::min(s[0].y, s[1].y), hi = std::max(s[0].y, s[1].y);
        return y < lo ? lo : (y > hi ? hi : y);
    }
    void clip_line(Pt p0, Pt p1) {
        const Pt src[2] = {p0, p1};
        int i0 = src[0].y < src[1].y ? 0 : 1, i1 = 1 - i0;
        if (src[i1].y <= top) return;     // above
        if (src[i0].y >= bottom) return;  // below
        Pt tmp[2] = {src[0], src[1]};
        if (src[i0].y < top) tmp[i0] = Pt{sect_with_horizontal(src, top), top};
        if (tmp[i1].y > bottom) tmp[i1] = Pt{sect_with_horizontal(src, bottom), bottom};
        bool reverse;
        if (src[0].x < src[1].x) { i0 = 0; i1 = 1; reverse = false; }
        else { i0 = 1; i1 = 0; reverse = true; }
        Pt res[4];
        int n;
        if (tmp[i1].x <= left) {  // wholly to the left
            tmp[0].x = tmp[1].x = left;
            res[0] = tmp[0]; res[1] = tmp[1];
            n = 1;
            reverse = false;
        } else if (tmp[i0].x >= right) {  // wholly to the right (not culled)
            tmp[0].x = tmp[1].x = right;
            res[0] = tmp[0]; res[1] = tmp[1];
            n = 1;
            reverse = false;
        } else {
            int k = 0;
            if (tmp[i0].x < left) {
                res[k++] = Pt{left, tmp[i0].y};
                res[k] = Pt{left, sect_clamp_with_vertical(tmp, left)};
            } else {
                res[k] = tmp[i0];
            }
            k++;
            if (tmp[i1].x > right) {
                res[k++] = Pt{right, sect_clamp_with_vertical(tmp, right)};
                res[k] = Pt{right, tmp[i1].y};
            } else {
                res[k] = tmp[i1];
            }
            n = k;
        }
        Pt pts[4];
        if (reverse) for (int i = 0; i <= n; i++) pts[n - i] = res[i];
        else for (int i = 0; i <= n; i++) pts[i] = res[i];
        for (int i = 0; i < n; i++) line(pts[i], pts[i + 1]);
    }

    // ---- quads ----
    static bool sort_increasing_y(const Pt *src, Pt *dst, int n) {
        if (src[0].y > src[n - 1].y) {
            for (int i = 0; i < n; i++) dst[i] = src[n - 1 - i];
            return true;
        }
        for (int i = 0; i < n; i++) dst[i] = src[i];
        return false;
    }
    void chop_quad_in_y(Pt *pts) const {
        float t;
        Pt tmp[5];
        if (pts[0].y < top) {
            if (chop_mono_quad_at(pts[0].y, pts[1].y, pts[2].y, top, t)) {
                chop_quad(pts, t, tmp);
                tmp[2].y = top;
                tmp[3].y = std::max(tmp[3].y, top);
                pts[0] = tmp[2];
                pts[1] = tmp[3];
            } else {
                for (int i = 0; i < 3; i++)
                    if (pts[i].y < top) pts[i].y = top;
            }
        }
        if (pts[2].y > bottom) {
            if (chop_mono_quad_at(pts[0].y, pts[1].y, pts[2].y, bottom, t)) {
                chop_quad(pts, t, tmp);
                tmp[1].y = std::min(tmp[1].y, bottom);
                tmp[2].y = bottom;
                pts[1] = tmp[1];
                pts[2] = tmp[2];
            } else {
                for (int i = 0; i < 3; i++)
                    if (pts[i].y > bottom) pts[i].y = bottom;
            }
        }
    }
    void clip_mono_quad(const Pt *src) {
        Pt pts[3];
        bool reverse = sort_increasing_y(src, pts, 3);
        if (pts[2].y <= top || pts[0].y >= bottom) return;
        chop_quad_in_y(pts);
        if (pts[0].x > pts[2].x) {
            std::swap(pts[0], pts[2]);
            reverse = !reverse;
        }
        if (pts[2].x <= left) { vline(left, pts[0].y, pts[2].y, reverse); return; }
        if (pts[0].x >= right) { vline(right, pts[0].y, pts[2].y, reverse); return; }
        float t;
        Pt tmp[5];
        if (pts[0].x < left) {
            if (chop_mono_quad_at(pts[0].x, pts[1].x, pts[2].x, left, t)) {
                chop_quad(pts, t, tmp);
                vline(left, tmp[0].y, tmp[2].y, reverse);
                tmp[2].x = left;
                tmp[3].x = std::max(tmp[3].x, left);
                pts[0] = tmp[2];
                pts[1] = tmp[3];
            } else {
                vline(left, pts[0].y, pts[2].y, reverse);
                return;
            }
        }
        if (pts[2].x > right) {
            if (chop_mono_quad_at(pts[0].x, pts[1].x, pts[2].x, right, t)) {
                chop_quad(pts, t, tmp);
                tmp[1].x = std::min(tmp[1].x, right);
                tmp[2].x = right;
                quad(tmp, reverse);
                vline(right, tmp[2].y, tmp[4].y, reverse);
            } else {
                pts[1].x = std::min(pts[1].x, right);
                pts[2].x = std::min(pts[2].x, right);
                quad(pts, reverse);
            }
        } else {
            quad(pts, reverse);
        }
    }
    void clip_quad(const Pt *p) {
        const float mn = std::min({p[0].y, p[1].y, p[2].y}), mx = std::max({p[0].y, p[1].y, p[2].y});
        if (mn >= bottom || mx <= top) return;  // quick_reject
        Pt mono_y[5];
        const int ny = quad_monotonic(p, mono_y, 1);
        for (int y = 0; y <= ny; y++) {
            Pt mono_x[5];
            const int nx = quad_monotonic(mono_y + 2 * y, mono_x, 0);
            for (int x = 0; x <= nx; x++) clip_mono_quad(mono_x + 2 * x);
        }
    }

    // ---- cubics ----
    void chop_cubic_in_y(Pt *pts) const {
        if (pts[0].y < top) {
            Pt tmp[7];
            chop_mono_cubic_at(pts, top, 1, tmp);
            // a poor chop can leave the lower cubic's first three Ys above the clip:
            // chop that part again
            if (tmp[3].y < top && tmp[4].y < top && tmp[5].y < top) {
                const Pt tmp2[4] = {tmp[3], tmp[4], tmp[5], tmp[6]};
                chop_mono_cubic_at(tmp2, top, 1, tmp);
            }
            tmp[3].y = top;
            tmp[4].y = std::max(tmp[4].y, top);
            pts[0] = tmp[3];
            pts[1] = tmp[4];
            pts[2] = tmp[5];
        }
        if (pts[3].y > bottom) {
            Pt tmp[7];
            chop_mono_cubic_at(pts, bottom, 1, tmp);
            tmp[3].y = bottom;
            tmp[2].y = std::min(tmp[2].y, bottom);
            pts[1] = tmp[1];
            pts[2] = tmp[2];
            pts[3] = tmp[3];
        }
    }
    void clip_mono_cubic(const Pt *src) {
        Pt pts[4];
        bool reverse = sort_increasing_y(src, pts, 4);
        if (pts[3].y <= top || pts[0].y >= bottom) return;
        chop_cubic_in_y(pts);
        if (pts[0].x > pts[3].x) {
            std::swap(pts[0], pts[3]);
            std::swap(pts[1], pts[2]);
            reverse = !reverse;
        }
        if (pts[3].x <= left) { vline(left, pts[0].y, pts[3].y, reverse); return; }
        if (pts[0].x >= right) { vline(right, pts[0].y, pts[3].y, reverse); return; }
        if (pts[0].x < left) {
            Pt tmp[7];
            chop_mono_cubic_at(pts, left, 0, tmp);
            vline(left, tmp[0].y, tmp[3].y, reverse);
            tmp[3].x = left;
            tmp[4].x = std::max(tmp[4].x, left);
            pts[0] = tmp[3];
            pts[1] = tmp[4];
            pts[2] = tmp[5];
        }
        if (pts[3].x > right) {
            Pt tmp[7];
            chop_mono_cubic_at(pts, right, 0, tmp);
            tmp[3].x = right;
            tmp[2].x = std::min(tmp[2].x, right);
            cubic(tmp, reverse);
            vline(right, tmp[3].y, tmp[6].y, reverse);
        } else {
            cubic(pts, reverse);
        }
    }
    void clip_cubic(const Pt *p) {
        float l = p[0].x, t = p[0].y, r = p[0].x, b = p[0].y;
        for (int i = 1; i < 4; i++) {
            l = std::min(l, p[i].x); r = std::max(r, p[i].x);
            t = std::min(t, p[i].y); b = std::max(b, p[i].y);
        }
        if (!(b > top && t < bottom)) return;
        const float kMaxLimit = (float)(1 << 22);  // too_big_for_reliable_float_math
        if (l < -kMaxLimit || t < -kMaxLimit || r > kMaxLimit || b > kMaxLimit) {
            clip_line(p[0], p[3]);
            return;
        }
        Pt mono_y[10];
        const int ny = cubic_monotonic(p, mono_y, 1);
        for (int y = 0; y <= ny; y++) {
            Pt mono_x[10];
            const int nx = cubic_monotonic(mono_y + 3 * y, mono_x, 0);
            for (int x = 0; x <= nx; x++) clip_mono_cubic(mono_x + 3 * x);
        }
    }
};

}  // namespace

bool segs_inside_pixmap(const std::vector<Seg> &segs, bool aa, uint32_t w, uint32_t h) {
    float l = INFINITY, t = INFINITY, r = -INFINITY, b = -INFINITY;
    for (const Seg &s : segs)
        for (int i = 0; i < s.n; i++) {
            l = std::min(l, s.p[i].x); r = std::max(r, s.p[i].x);
            t = std::min(t, s.p[i].y); b = std::max(b, s.p[i].y);
        }
    double il, it, ir, ib;
    if (aa) {  // scan/path_aa.rs fill_path: floor / ceil, then round_out
        il = std::floor(l); it = std::floor(t); ir = std::ceil(r); ib = std::ceil(b);
    } else {  // scan/path.rs conservative_round_to_int (bias 0.5 + 1.5 / 64)
        const double bias = 0.5 + 1.5 / 64.0;
        il = std::ceil((double)l - bias); it = std::ceil((double)t - bias);
        ir = std::floor((double)r + bias); ib = std::floor((double)b + bias);
    }
    // IntRect::to_screen_int_rect fails for a negative origin; ScreenIntRect::contains
    return il >= 0 && it >= 0 && ir <= (double)w && ib <= (double)h;
}

void clip_segments(const std::vector<Seg> &segs, uint32_t w, uint32_t h, std::vector<Seg> &out) {
    Clipper c{0.0f, 0.0f, (float)w, (float)h, out};
    for (const Seg &s : segs) {
        if (s.n == 2) c.clip_line(s.p[0], s.p[1]);
        else if (s.n == 3) c.clip_quad(s.p);
        else c.clip_cubic(s.p);
    }
}

}  // namespace rvkr
