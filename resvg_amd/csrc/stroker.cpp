// Path stroking (tiny-skia 0.12.0 stroker.rs / path_builder.rs / path_geometry.rs,
// themselves ports of Skia's SkStroke.cpp, SkStrokerPriv.cpp, SkGeometry.cpp).
// Every float expression keeps the reference's operation order; build without FMA
// contraction.
//
// Restates tiny-skia (BSD-3-Clause, Copyright 2020 Yevhenii Reizner) and Skia
// (BSD-3-Clause, Copyright Google Inc.); see THIRD_PARTY_NOTICES.md.
#include "stroker.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>

namespace rvks {

namespace {

constexpr float kNearlyZero = 1.0f / 4096.0f;    // SCALAR_NEARLY_ZERO
constexpr float kRoot2Over2 = 0.707106781f;      // SCALAR_ROOT_2_OVER_2
constexpr int kRecursiveLimits[4] = {5 * 3, 26 * 3, 11 * 3, 11 * 3};  // tangent, cubic, conic, quad
constexpr int kTangentLimit = 0, kQuadLimit = 3;

inline Pt operator+(Pt a, Pt b) { return Pt{a.x + b.x, a.y + b.y}; }
inline Pt operator-(Pt a, Pt b) { return Pt{a.x - b.x, a.y - b.y}; }
inline Pt operator-(Pt a) { return Pt{-a.x, -a.y}; }
inline bool operator==(Pt a, Pt b) { return a.x == b.x && a.y == b.y; }
inline bool operator!=(Pt a, Pt b) { return !(a == b); }
inline Pt scaled(Pt a, float s) { return Pt{a.x * s, a.y * s}; }
inline float dot(Pt a, Pt b) { return a.x * b.x + a.y * b.y; }
inline float cross(Pt a, Pt b) { return a.x * b.y - a.y * b.x; }
inline float length_sqd(Pt a) { return dot(a, a); }
inline float distance_to_sqd(Pt a, Pt b) {
    const float dx = a.x - b.x, dy = a.y - b.y;
    return dx * dx + dy * dy;
}
inline bool is_finite(Pt a) { return std::isfinite(a.x) && std::isfinite(a.y); }
inline bool can_normalize(Pt a) { return std::isfinite(a.x) && std::isfinite(a.y) && (a.x != 0.0f || a.y != 0.0f); }
inline Pt rotate_cw(Pt a) { return Pt{-a.y, a.x}; }
inline Pt rotate_ccw(Pt a) { return Pt{a.y, -a.x}; }
inline bool is_nearly_zero(float v) { return std::fabs(v) <= kNearlyZero; }

// Point::set_length_from (the magnitude in f64, the scale applied in f32)
bool set_length_from(Pt &p, float x, float y, float length) {
    const double xx = x, yy = y;
    const double dmag = std::sqrt(xx * xx + yy * yy);
    const double dscale = (double)length / dmag;
    x *= (float)dscale;
    y *= (float)dscale;
    if (!std::isfinite(x) || !std::isfinite(y) || (x == 0.0f && y == 0.0f)) {
        p = Pt{0.0f, 0.0f};
        return false;
    }
    p = Pt{x, y};
    return true;
}
inline bool set_length(Pt &p, float length) { return set_length_from(p, p.x, p.y, length); }
inline bool normalize(Pt &p) { return set_length_from(p, p.x, p.y, 1.0f); }

// ---------------------------------------------------------------------------
// transforms (tiny-skia-path transform.rs): concat with f64 mul-add in the skew case
// ---------------------------------------------------------------------------
struct Ts {
    float sx, ky, kx, sy, tx, ty;
};
inline bool has_skew(const Ts &t) { return t.kx != 0.0f || t.ky != 0.0f; }
inline bool is_identity(const Ts &t) {
    return t.sx == 1.0f && t.ky == 0.0f && t.kx == 0.0f && t.sy == 1.0f && t.tx == 0.0f && t.ty == 0.0f;
}
inline float mul_add_mul(float a, float b, float c, float d) {
    return (float)((double)a * (double)b + (double)c * (double)d);
}
Ts concat(const Ts &a, const Ts &b) {
    if (is_identity(a)) return b;
    if (is_identity(b)) return a;
    if (!has_skew(a) && !has_skew(b))
        return Ts{a.sx * b.sx, 0.0f, 0.0f, a.sy * b.sy, a.sx * b.tx + a.tx, a.sy * b.ty + a.ty};
    return Ts{mul_add_mul(a.sx, b.sx, a.kx, b.ky),        mul_add_mul(a.ky, b.sx, a.sy, b.ky),
              mul_add_mul(a.sx, b.kx, a.kx, b.sy),        mul_add_mul(a.ky, b.kx, a.sy, b.sy),
              mul_add_mul(a.sx, b.tx, a.kx, b.ty) + a.tx, mul_add_mul(a.ky, b.tx, a.sy, b.ty) + a.ty};
}
void map_points(const Ts &t, Pt *p, size_t n) {
    if (is_identity(t)) return;
    if (!has_skew(t) && t.sx == 1.0f && t.sy == 1.0f) {
        for (size_t i = 0; i < n; i++) { p[i].x += t.tx; p[i].y += t.ty; }
    } else if (!has_skew(t)) {
        for (size_t i = 0; i < n; i++) { p[i].x = p[i].x * t.sx + t.tx; p[i].y = p[i].y * t.sy + t.ty; }
    } else {
        for (size_t i = 0; i < n; i++) {
            const float x = p[i].x * t.sx + p[i].y * t.kx + t.tx;
            const float y = p[i].x * t.ky + p[i].y * t.sy + t.ty;
            p[i] = Pt{x, y};
        }
    }
}

// ---------------------------------------------------------------------------
// geometry (path_geometry.rs)
// ---------------------------------------------------------------------------
bool valid_unit_divide(float numer, float denom, float &r) {
    if (numer < 0.0f) { numer = -numer; denom = -denom; }
    if (denom == 0.0f || numer == 0.0f || numer >= denom) return false;
    r = numer / denom;
    return std::isfinite(r) && r > 0.0f && r < 1.0f;
}

int find_unit_quad_roots(float a, float b, float c, float roots[2]) {
    if (a == 0.0f) {
        float r;
        if (valid_unit_divide(-c, b, r)) { roots[0] = r; return 1; }
        return 0;
    }
    double dr = (double)b * (double)b - 4.0 * (double)a * (double)c;
    if (dr < 0.0) return 0;
    dr = std::sqrt(dr);
    const float r = (float)dr;
    if (!std::isfinite(r)) return 0;
    const float q = b < 0.0f ? -(b - r) / 2.0f : -(b + r) / 2.0f;
    int n = 0;
    float t;
    if (valid_unit_divide(q, a, t)) roots[n++] = t;
    if (valid_unit_divide(c, q, t)) roots[n++] = t;
    if (n == 2) {
        if (roots[0] > roots[1]) std::swap(roots[0], roots[1]);
        else if (roots[0] == roots[1]) n = 1;
    }
    return n;
}

Pt eval_quad_at(const Pt q[3], float t) {  // QuadCoeff: (a t + b) t + c
    const Pt b{(q[1].x - q[0].x) + (q[1].x - q[0].x), (q[1].y - q[0].y) + (q[1].y - q[0].y)};
    const Pt a{(q[2].x - (q[1].x + q[1].x)) + q[0].x, (q[2].y - (q[1].y + q[1].y)) + q[0].y};
    return Pt{(a.x * t + b.x) * t + q[0].x, (a.y * t + b.y) * t + q[0].y};
}

Pt eval_quad_tangent_at(const Pt q[3], float t) {
    if ((t == 0.0f && q[0] == q[1]) || (t == 1.0f && q[1] == q[2])) return q[2] - q[0];
    const Pt b = q[1] - q[0];
    const Pt a = (q[2] - q[1]) - b;
    const Pt r{a.x * t + b.x, a.y * t + b.y};
    return Pt{r.x + r.x, r.y + r.y};
}

float find_quad_max_curvature(const Pt q[3]) {
    const float ax = q[1].x - q[0].x, ay = q[1].y - q[0].y;
    const float bx = q[0].x - q[1].x - q[1].x + q[2].x, by = q[0].y - q[1].y - q[1].y + q[2].y;
    float numer = -(ax * bx + ay * by);
    float denom = bx * bx + by * by;
    if (denom < 0.0f) { numer = -numer; denom = -denom; }
    if (numer <= 0.0f) return 0.0f;
    if (numer >= denom) return 1.0f;
    return numer / denom;
}

Pt eval_cubic_pos_at(const Pt c[4], float t) {  // CubicCoeff: ((a t + b) t + c) t + d
    if (t == 0.0f) return c[0];
    const Pt a{(c[3].x + 3.0f * (c[1].x - c[2].x)) - c[0].x, (c[3].y + 3.0f * (c[1].y - c[2].y)) - c[0].y};
    const Pt b{3.0f * ((c[2].x - (c[1].x + c[1].x)) + c[0].x), 3.0f * ((c[2].y - (c[1].y + c[1].y)) + c[0].y)};
    const Pt cc{3.0f * (c[1].x - c[0].x), 3.0f * (c[1].y - c[0].y)};
    return Pt{((a.x * t + b.x) * t + cc.x) * t + c[0].x, ((a.y * t + b.y) * t + cc.y) * t + c[0].y};
}

Pt eval_cubic_derivative(const Pt c[4], float t) {
    const Pt a{(c[3].x + 3.0f * (c[1].x - c[2].x)) - c[0].x, (c[3].y + 3.0f * (c[1].y - c[2].y)) - c[0].y};
    const Pt bh{(c[2].x - (c[1].x + c[1].x)) + c[0].x, (c[2].y - (c[1].y + c[1].y)) + c[0].y};
    const Pt b{bh.x + bh.x, bh.y + bh.y};
    const Pt cc = c[1] - c[0];
    return Pt{(a.x * t + b.x) * t + cc.x, (a.y * t + b.y) * t + cc.y};
}

Pt eval_cubic_tangent_at(const Pt c[4], float t) {
    if ((t == 0.0f && c[0] == c[1]) || (t == 1.0f && c[2] == c[3])) {
        Pt tan = t == 0.0f ? c[2] - c[0] : c[3] - c[1];
        if (tan.x == 0.0f && tan.y == 0.0f) tan = c[3] - c[0];
        return tan;
    }
    return eval_cubic_derivative(c, t);
}

inline Pt interp(Pt a, Pt b, float t) { return Pt{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t}; }

void chop_cubic_at2(const Pt s[4], float t, Pt d[7]) {
    const Pt ab = interp(s[0], s[1], t), bc = interp(s[1], s[2], t), cd = interp(s[2], s[3], t);
    const Pt abc = interp(ab, bc, t), bcd = interp(bc, cd, t), abcd = interp(abc, bcd, t);
    d[0] = s[0]; d[1] = ab; d[2] = abc; d[3] = abcd; d[4] = bcd; d[5] = cd; d[6] = s[3];
}

int find_cubic_inflections(const Pt s[4], float t[2]) {
    const float ax = s[1].x - s[0].x, ay = s[1].y - s[0].y;
    const float bx = s[2].x - 2.0f * s[1].x + s[0].x, by = s[2].y - 2.0f * s[1].y + s[0].y;
    const float cx = s[3].x + 3.0f * (s[1].x - s[2].x) - s[0].x, cy = s[3].y + 3.0f * (s[1].y - s[2].y) - s[0].y;
    return find_unit_quad_roots(bx * cy - by * cx, ax * cy - ay * cx, ax * by - ay * bx, t);
}

// SkFindCubicMaxCurvature: roots of F'(t) . F''(t)
void formulate_f1_dot_f2(const float src[7], float coeff[4]) {  // src: stride 2
    const float a = src[2] - src[0];
    const float b = src[4] - 2.0f * src[2] + src[0];
    const float c = src[6] + 3.0f * (src[2] - src[4]) - src[0];
    coeff[0] = c * c;
    coeff[1] = 3.0f * b * c;
    coeff[2] = 2.0f * b * b + c * a;
    coeff[3] = a * b;
}

int collapse_duplicates(float *a, int n) {
    int count = n;
    for (int i = count - 1; i > 0; i--)
        if (a[i] == a[i - 1]) {
            count--;
            for (int j = i; j < count; j++) a[j] = a[j + 1];
        }
    return count;
}

int solve_cubic_poly(const float coeff[4], float t[3]) {
    if (is_nearly_zero(coeff[0])) {
        float r[2];
        const int n = find_unit_quad_roots(coeff[1], coeff[2], coeff[3], r);
        for (int i = 0; i < n; i++) t[i] = r[i];
        return n;
    }
    const float inva = 1.0f / coeff[0];
    const float a = coeff[1] * inva, b = coeff[2] * inva, c = coeff[3] * inva;
    const float q = (a * a - b * 3.0f) / 9.0f;
    const float r = (2.0f * a * a * a - 9.0f * a * b + 27.0f * c) / 54.0f;
    const float q3 = q * q * q;
    const float r2_minus_q3 = r * r - q3;
    const float adiv3 = a / 3.0f;
    auto pin01 = [](float v) { return std::min(std::max(v, 0.0f), 1.0f); };
    if (r2_minus_q3 < 0.0f) {
        const float theta = std::acos(std::min(std::max(r / std::sqrt(q3), -1.0f), 1.0f));
        const float neg2_root_q = -2.0f * std::sqrt(q);
        const float pi = 3.14159265f;
        t[0] = pin01(neg2_root_q * std::cos(theta / 3.0f) - adiv3);
        t[1] = pin01(neg2_root_q * std::cos((theta + 2.0f * pi) / 3.0f) - adiv3);
        t[2] = pin01(neg2_root_q * std::cos((theta - 2.0f * pi) / 3.0f) - adiv3);
        std::sort(t, t + 3);
        return collapse_duplicates(t, 3);
    }
    float aa = std::fabs(r) + std::sqrt(r2_minus_q3);
    aa = std::pow(aa, 0.3333333f);  // scalar_cube_root
    if (r > 0.0f) aa = -aa;
    if (aa != 0.0f) aa += q / aa;
    t[0] = pin01(aa - adiv3);
    return 1;
}

int find_cubic_max_curvature(const Pt s[4], float t[3]) {
    float xs[7], ys[7], cx[4], cy[4];
    for (int i = 0; i < 4; i++) { xs[2 * i] = s[i].x; ys[2 * i] = s[i].y; }
    formulate_f1_dot_f2(xs, cx);
    formulate_f1_dot_f2(ys, cy);
    for (int i = 0; i < 4; i++) cx[i] += cy[i];
    return solve_cubic_poly(cx, t);
}

bool on_same_side(const Pt s[4], int test_index, int line_index) {
    const Pt origin = s[line_index];
    const Pt line = s[line_index + 1] - origin;
    float crosses[2];
    for (int i = 0; i < 2; i++) crosses[i] = cross(line, s[test_index + i] - origin);
    return crosses[0] * crosses[1] >= 0.0f;
}

float calc_cubic_precision(const Pt s[4]) {
    return (distance_to_sqd(s[1], s[0]) + distance_to_sqd(s[2], s[1]) + distance_to_sqd(s[3], s[2])) * 1e-8f;
}

// SkFindCubicCusp: t of a cusp in (0, 1), or -1
float find_cubic_cusp(const Pt s[4]) {
    if (s[0] == s[1] || s[2] == s[3]) return -1.0f;
    if (on_same_side(s, 0, 2) || on_same_side(s, 2, 0)) return -1.0f;
    float mc[3];
    const int n = find_cubic_max_curvature(s, mc);
    for (int i = 0; i < n; i++) {
        const float t = mc[i];
        if (0.0f >= t || t >= 1.0f) continue;
        const Pt d = eval_cubic_derivative(s, t);
        if (length_sqd(d) < calc_cubic_precision(s)) return t;
    }
    return -1.0f;
}

// ---------------------------------------------------------------------------
// conics (path_geometry.rs Conic, AutoConicToQuads)
// ---------------------------------------------------------------------------
struct Conic {
    Pt p[3];
    float w;
};

inline bool between(float a, float b, float c) { return (a - b) * (c - b) <= 0.0f; }

void conic_chop(const Conic &c, Conic out[2]) {
    const float scale = 1.0f / (1.0f + c.w);
    const float new_w = std::sqrt(0.5f + c.w * 0.5f);
    const Pt wp1 = scaled(c.p[1], c.w);
    Pt m{((c.p[0].x + (wp1.x + wp1.x)) + c.p[2].x) * scale * 0.5f,
         ((c.p[0].y + (wp1.y + wp1.y)) + c.p[2].y) * scale * 0.5f};
    if (!is_finite(m)) {
        const double wd = c.w, w2 = wd * 2.0, sh = 1.0 / (1.0 + wd) * 0.5;
        m.x = (float)(((double)c.p[0].x + w2 * (double)c.p[1].x + (double)c.p[2].x) * sh);
        m.y = (float)(((double)c.p[0].y + w2 * (double)c.p[1].y + (double)c.p[2].y) * sh);
    }
    out[0] = Conic{{c.p[0], Pt{(c.p[0].x + wp1.x) * scale, (c.p[0].y + wp1.y) * scale}, m}, new_w};
    out[1] = Conic{{m, Pt{(wp1.x + c.p[2].x) * scale, (wp1.y + c.p[2].y) * scale}, c.p[2]}, new_w};
}

Pt *conic_subdivide(const Conic &src, Pt *pts, int level) {
    if (level == 0) {
        pts[0] = src.p[1];
        pts[1] = src.p[2];
        return pts + 2;
    }
    Conic dst[2];
    conic_chop(src, dst);
    const float start_y = src.p[0].y, end_y = src.p[2].y;
    if (between(start_y, src.p[1].y, end_y)) {
        const float mid_y = dst[0].p[2].y;
        if (!between(start_y, mid_y, end_y)) {
            const float closer = std::fabs(mid_y - start_y) < std::fabs(mid_y - end_y) ? start_y : end_y;
            dst[0].p[2].y = closer;
            dst[1].p[0].y = closer;
        }
        if (!between(start_y, dst[0].p[1].y, dst[0].p[2].y)) dst[0].p[1].y = start_y;
        if (!between(dst[1].p[0].y, dst[1].p[1].y, end_y)) dst[1].p[1].y = end_y;
    }
    pts = conic_subdivide(dst[0], pts, level - 1);
    return conic_subdivide(dst[1], pts, level - 1);
}

// AutoConicToQuads::compute (tolerance 0.25, at least one subdivision); returns the
// number of quads, points in out[0 .. 2n].
int conic_to_quads(Pt p0, Pt p1, Pt p2, float w, Pt out[33]) {
    const Conic c{{p0, p1, p2}, w};
    if (!is_finite(p0) || !is_finite(p1) || !is_finite(p2)) return 0;
    const float a = w - 1.0f;
    const float k = a / (4.0f * (2.0f + a));
    const float x = k * (p0.x - 2.0f * p1.x + p2.x);
    const float y = k * (p0.y - 2.0f * p1.y + p2.y);
    float error = std::sqrt(x * x + y * y);
    int pow2 = 0;
    for (int i = 0; i < 4; i++) {
        if (error <= 0.25f) break;
        error *= 0.25f;
        pow2++;
    }
    pow2 = std::max(pow2, 1);
    out[0] = p0;
    conic_subdivide(c, out + 1, pow2);
    const int quads = 1 << pow2;
    const int npts = 2 * quads + 1;
    bool finite = true;
    for (int i = 0; i < npts; i++) finite = finite && is_finite(out[i]);
    if (!finite)
        for (int i = 1; i < npts - 1; i++) out[i] = p1;
    return quads;
}

// Conic::build_unit_arc (SkConic::BuildUnitArc); returns the number of conics.
int build_unit_arc(Pt u_start, Pt u_stop, bool ccw, const Ts &user, Conic dst[5]) {
    const float x = dot(u_start, u_stop);
    float y = cross(u_start, u_stop);
    const float abs_y = std::fabs(y);
    if (abs_y <= kNearlyZero && x > 0.0f && ((y >= 0.0f && !ccw) || (y <= 0.0f && ccw))) return 0;
    if (ccw) y = -y;
    int quadrant = 0;
    if (y == 0.0f) {
        quadrant = 2;
    } else if (x == 0.0f) {
        quadrant = y > 0.0f ? 1 : 3;
    } else {
        if (y < 0.0f) quadrant += 2;
        if ((x < 0.0f) != (y < 0.0f)) quadrant += 1;
    }
    static const Pt qp[8] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};
    int count = quadrant;
    for (int i = 0; i < count; i++) dst[i] = Conic{{qp[i * 2], qp[i * 2 + 1], qp[(i * 2 + 2) % 8]}, kRoot2Over2};
    const Pt final_pt{x, y};
    const Pt last_q = qp[(quadrant * 2) % 8];
    const float d = dot(last_q, final_pt);
    if (d < 1.0f) {
        Pt off{last_q.x + x, last_q.y + y};
        const float cos_half = std::sqrt((1.0f + d) / 2.0f);
        set_length(off, 1.0f / cos_half);
        if (can_normalize(last_q - off)) dst[count++] = Conic{{last_q, off, final_pt}, cos_half};
    }
    Ts t{u_start.x, u_start.y, -u_start.y, u_start.x, 0.0f, 0.0f};  // from_sin_cos(sin, cos)
    if (ccw) t = concat(t, Ts{1.0f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f});
    t = concat(user, t);
    for (int i = 0; i < count; i++) map_points(t, dst[i].p, 3);
    return count;
}

// ---------------------------------------------------------------------------
// caps and joins (SkStrokerPriv)
// ---------------------------------------------------------------------------
enum class Angle { NearlyLine, Shallow, Sharp, Nearly180 };

Angle dot_to_angle_type(float d) {
    if (d >= 0.0f) return is_nearly_zero(1.0f - d) ? Angle::NearlyLine : Angle::Shallow;
    return is_nearly_zero(1.0f + d) ? Angle::Nearly180 : Angle::Sharp;
}

inline bool is_clockwise(Pt before, Pt after) { return before.x * after.y > before.y * after.x; }

void handle_inner_join(Pt pivot, Pt after, PathBuilder &inner) {
    inner.line_to(pivot.x, pivot.y);
    inner.line_to(pivot.x - after.x, pivot.y - after.y);
}

void cap(uint32_t kind, Pt pivot, Pt normal, Pt stop, const PathBuilder *other, PathBuilder &path) {
    if (kind == 1) {  // round
        const Pt parallel = rotate_cw(normal);
        const Pt projected = pivot + parallel;
        path.conic_points_to(projected + normal, projected, kRoot2Over2);
        path.conic_points_to(projected - normal, stop, kRoot2Over2);
    } else if (kind == 2) {  // square
        const Pt parallel = rotate_cw(normal);
        if (other) {
            path.set_last_point(Pt{pivot.x + normal.x + parallel.x, pivot.y + normal.y + parallel.y});
            path.line_to(pivot.x - normal.x + parallel.x, pivot.y - normal.y + parallel.y);
        } else {
            path.line_to(pivot.x + normal.x + parallel.x, pivot.y + normal.y + parallel.y);
            path.line_to(pivot.x - normal.x + parallel.x, pivot.y - normal.y + parallel.y);
            path.line_to(stop.x, stop.y);
        }
    } else {  // butt
        path.line_to(stop.x, stop.y);
    }
}

void bevel_join(Pt before_unit, Pt pivot, Pt after_unit, float radius, PathBuilder *outer, PathBuilder *inner) {
    Pt after = scaled(after_unit, radius);
    if (!is_clockwise(before_unit, after_unit)) {
        std::swap(outer, inner);
        after = -after;
    }
    outer->line_to(pivot.x + after.x, pivot.y + after.y);
    handle_inner_join(pivot, after, *inner);
}

void round_join(Pt before_unit, Pt pivot, Pt after_unit, float radius, PathBuilder *outer, PathBuilder *inner) {
    const float d = dot(before_unit, after_unit);
    if (dot_to_angle_type(d) == Angle::NearlyLine) return;
    Pt before = before_unit, after = after_unit;
    bool ccw = false;
    if (!is_clockwise(before, after)) {
        std::swap(outer, inner);
        before = -before;
        ccw = true;
    }
    const Ts ts{radius, 0.0f, 0.0f, radius, pivot.x, pivot.y};
    Conic conics[5];
    const int n = build_unit_arc(before, after, ccw, ts, conics);
    if (n > 0) {
        for (int i = 0; i < n; i++) outer->conic_points_to(conics[i].p[1], conics[i].p[2], conics[i].w);
        after = scaled(after, radius);
        handle_inner_join(pivot, after, *inner);
    }
}

void miter_join(Pt before_unit, Pt pivot, Pt after_unit, float radius, float inv_miter_limit, bool miter_clip,
                bool prev_is_line, bool curr_is_line, PathBuilder *outer, PathBuilder *inner) {
    auto blunt_or_clipped = [&](Pt before, Pt mid, Pt after) {
        after = scaled(after, radius);
        if (miter_clip) {
            normalize(mid);
            const float cos_beta = before.x * mid.x + before.y * mid.y;
            const float sin_beta = cross(before, mid);
            const float x = std::fabs(sin_beta) <= kNearlyZero ? 1.0f / inv_miter_limit
                                                                : ((1.0f / inv_miter_limit) - cos_beta) / sin_beta;
            before = scaled(before, radius);
            const Pt before_tangent = rotate_cw(before);
            const Pt after_tangent = rotate_ccw(after);
            const Pt c1 = (pivot + before) + scaled(before_tangent, x);
            const Pt c2 = (pivot + after) + scaled(after_tangent, x);
            if (prev_is_line) outer->set_last_point(c1);
            else outer->line_to(c1.x, c1.y);
            outer->line_to(c2.x, c2.y);
        }
        if (!curr_is_line) outer->line_to(pivot.x + after.x, pivot.y + after.y);
        handle_inner_join(pivot, after, *inner);
    };
    auto miter = [&](Pt mid, Pt after) {
        after = scaled(after, radius);
        if (prev_is_line) outer->set_last_point(Pt{pivot.x + mid.x, pivot.y + mid.y});
        else outer->line_to(pivot.x + mid.x, pivot.y + mid.y);
        if (!curr_is_line) outer->line_to(pivot.x + after.x, pivot.y + after.y);
        handle_inner_join(pivot, after, *inner);
    };
    const float d = dot(before_unit, after_unit);
    const Angle angle = dot_to_angle_type(d);
    Pt before = before_unit, after = after_unit;
    if (angle == Angle::NearlyLine) return;
    if (angle == Angle::Nearly180) {
        curr_is_line = false;
        blunt_or_clipped(before, scaled(after - before, radius / 2.0f), after);
        return;
    }
    const bool ccw = !is_clockwise(before, after);
    if (ccw) {
        std::swap(outer, inner);
        before = -before;
        after = -after;
    }
    if (d == 0.0f && inv_miter_limit <= kRoot2Over2) {
        miter(scaled(before + after, radius), after);
        return;
    }
    Pt mid;
    if (angle == Angle::Sharp) {
        mid = Pt{after.y - before.y, before.x - after.x};
        if (ccw) mid = -mid;
    } else {
        mid = Pt{before.x + after.x, before.y + after.y};
    }
    const float sin_half = std::sqrt((1.0f + d) * 0.5f);
    if (sin_half < inv_miter_limit) {
        curr_is_line = false;
        blunt_or_clipped(before, mid, after);
        return;
    }
    set_length(mid, radius / sin_half);
    miter(mid, after);
}

// ---------------------------------------------------------------------------
// the stroker
// ---------------------------------------------------------------------------
enum class Reduction { Point, Line, Quad, Degenerate, Degenerate2, Degenerate3 };
enum class Result { Split, Degenerate, Quad };

struct QuadConstruct {
    Pt quad[3]{};
    Pt tangent_start{}, tangent_end{};
    float start_t = 0.0f, mid_t = 0.0f, end_t = 0.0f;
    bool start_set = false, end_set = false, opposite_tangents = false;

    bool init(float start, float end) {
        start_t = start;
        mid_t = std::min(std::max((start + end) * 0.5f, 0.0f), 1.0f);
        end_t = end;
        start_set = end_set = false;
        return start_t < mid_t && mid_t < end_t;
    }
    bool init_with_start(const QuadConstruct &parent) {
        if (!init(parent.start_t, parent.mid_t)) return false;
        quad[0] = parent.quad[0];
        tangent_start = parent.tangent_start;
        start_set = true;
        return true;
    }
    bool init_with_end(const QuadConstruct &parent) {
        if (!init(parent.mid_t, parent.end_t)) return false;
        quad[2] = parent.quad[2];
        tangent_end = parent.tangent_end;
        end_set = true;
        return true;
    }
};

bool degenerate_vector(Pt v) { return !can_normalize(v); }

float pt_to_line(Pt pt, Pt line_start, Pt line_end) {
    const Pt dxy = line_end - line_start;
    const Pt ab0 = pt - line_start;
    const float numer = dot(dxy, ab0);
    const float denom = dot(dxy, dxy);
    const float t = numer / denom;
    if (t >= 0.0f && t <= 1.0f) {
        const Pt hit{line_start.x * (1.0f - t) + line_end.x * t, line_start.y * (1.0f - t) + line_end.y * t};
        return distance_to_sqd(hit, pt);
    }
    return distance_to_sqd(pt, line_start);
}

bool quad_in_line(const Pt q[3]) {
    float pt_max = -1.0f;
    int outer1 = 0, outer2 = 0;
    for (int i = 0; i < 2; i++)
        for (int j = i + 1; j < 3; j++) {
            const Pt d = q[j] - q[i];
            const float m = std::max(std::fabs(d.x), std::fabs(d.y));
            if (pt_max < m) { outer1 = i; outer2 = j; pt_max = m; }
        }
    const int mid = outer1 ^ outer2 ^ 3;
    const float slop = pt_max * pt_max * 0.000005f;
    return pt_to_line(q[mid], q[outer1], q[outer2]) <= slop;
}

bool cubic_in_line(const Pt c[4]) {
    float pt_max = -1.0f;
    int outer1 = 0, outer2 = 0;
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 4; j++) {
            const Pt d = c[j] - c[i];
            const float m = std::max(std::fabs(d.x), std::fabs(d.y));
            if (pt_max < m) { outer1 = i; outer2 = j; pt_max = m; }
        }
    const int mid1 = (1 + (2 >> outer2)) >> outer1;
    const int mid2 = outer1 ^ outer2 ^ mid1;
    const float slop = pt_max * pt_max * 0.00001f;
    return pt_to_line(c[mid1], c[outer1], c[outer2]) <= slop && pt_to_line(c[mid2], c[outer1], c[outer2]) <= slop;
}

Reduction check_quad_linear(const Pt q[3], Pt &reduction) {
    const bool ab = degenerate_vector(q[1] - q[0]), bc = degenerate_vector(q[2] - q[1]);
    if (ab && bc) return Reduction::Point;
    if (ab || bc) return Reduction::Line;
    if (!quad_in_line(q)) return Reduction::Quad;
    const float t = find_quad_max_curvature(q);
    if (t == 0.0f || t == 1.0f) return Reduction::Line;
    reduction = eval_quad_at(q, t);
    return Reduction::Degenerate;
}

Reduction check_cubic_linear(const Pt c[4], Pt reduction[3], Pt &tangent_pt) {
    const bool ab = degenerate_vector(c[1] - c[0]), bc = degenerate_vector(c[2] - c[1]),
               cd = degenerate_vector(c[3] - c[2]);
    if (ab && bc && cd) return Reduction::Point;
    if ((int)ab + (int)bc + (int)cd == 2) return Reduction::Line;
    if (!cubic_in_line(c)) {
        tangent_pt = ab ? c[2] : c[1];
        return Reduction::Quad;
    }
    float t[3];
    const int n = find_cubic_max_curvature(c, t);
    int r = 0;
    for (int i = 0; i < n; i++) {
        if (0.0f >= t[i] || t[i] >= 1.0f) continue;
        reduction[r] = eval_cubic_pos_at(c, t[i]);
        if (reduction[r] != c[0] && reduction[r] != c[3]) r++;
    }
    if (r == 0) return Reduction::Line;
    return r == 1 ? Reduction::Degenerate : (r == 2 ? Reduction::Degenerate2 : Reduction::Degenerate3);
}

bool sharp_angle(const Pt q[3]) {
    Pt smaller = q[1] - q[0], larger = q[1] - q[2];
    const float smaller_len = length_sqd(smaller);
    float larger_len = length_sqd(larger);
    if (smaller_len > larger_len) {
        std::swap(smaller, larger);
        larger_len = smaller_len;
    }
    if (!set_length(smaller, larger_len)) return false;
    return dot(smaller, larger) > 0.0f;
}

bool points_within_dist(Pt near_pt, Pt far_pt, float limit) { return distance_to_sqd(near_pt, far_pt) <= limit * limit; }

int intersect_quad_ray(const Pt line[2], const Pt q[3], float roots[2]) {
    const Pt vec = line[1] - line[0];
    float r[3];
    for (int n = 0; n < 3; n++) r[n] = (q[n].y - line[0].y) * vec.x - (q[n].x - line[0].x) * vec.y;
    float a = r[2], b = r[1];
    const float c = r[0];
    a += c - 2.0f * b;
    b -= c;
    return find_unit_quad_roots(a, 2.0f * b, c, roots);
}

struct Stroker {
    float radius = 0, inv_miter_limit = 0, res_scale = 1, inv_res_scale = 1, inv_res_scale_sq = 1;
    Pt first_normal{}, prev_normal{}, first_unit_normal{}, prev_unit_normal{};
    Pt first_pt{}, prev_pt{}, first_outer_pt{};
    size_t first_outer_pt_index = 0;
    int segment_count = -1;
    bool prev_is_line = false;
    uint32_t cap_kind = 0, join_kind = 0;
    bool round_join_override = false;
    PathBuilder inner, outer, cusper;
    float stroke_type = 1.0f;  // +1 outer, -1 inner
    int recursion_depth = 0;
    bool found_tangents = false, join_completed = false;

    void join(Pt before_unit, Pt pivot, Pt after_unit, bool prev_line, bool curr_line, PathBuilder *o,
              PathBuilder *i) {
        const uint32_t k = round_join_override ? 2u : join_kind;
        if (k == 2) round_join(before_unit, pivot, after_unit, radius, o, i);
        else if (k == 3) bevel_join(before_unit, pivot, after_unit, radius, o, i);
        else miter_join(before_unit, pivot, after_unit, radius, inv_miter_limit, k == 1, prev_line, curr_line, o, i);
    }

    bool set_normal_unit_normal(Pt before, Pt after, Pt &normal, Pt &unit) {
        if (!set_length_from(unit, (after.x - before.x) * res_scale, (after.y - before.y) * res_scale, 1.0f))
            return false;
        unit = rotate_ccw(unit);
        normal = scaled(unit, radius);
        return true;
    }
    bool set_normal_unit_normal2(Pt vec, Pt &normal, Pt &unit) {
        if (!set_length_from(unit, vec.x, vec.y, 1.0f)) return false;
        unit = rotate_ccw(unit);
        normal = scaled(unit, radius);
        return true;
    }

    bool pre_join_to(Pt p, bool curr_is_line, Pt &normal, Pt &unit) {
        const float px = prev_pt.x, py = prev_pt.y;
        if (!set_normal_unit_normal(prev_pt, p, normal, unit)) {
            if (cap_kind == 0) return false;
            normal = Pt{radius, 0.0f};
            unit = Pt{1.0f, 0.0f};
        }
        if (segment_count == 0) {
            first_normal = normal;
            first_unit_normal = unit;
            first_outer_pt = Pt{px + normal.x, py + normal.y};
            outer.move_to(first_outer_pt.x, first_outer_pt.y);
            inner.move_to(px - normal.x, py - normal.y);
        } else {
            join(prev_unit_normal, prev_pt, unit, prev_is_line, curr_is_line, &outer, &inner);
        }
        prev_is_line = curr_is_line;
        return true;
    }

    void post_join_to(Pt p, Pt normal, Pt unit) {
        join_completed = true;
        prev_pt = p;
        prev_unit_normal = unit;
        prev_normal = normal;
        segment_count++;
    }

    void finish_contour(bool close, bool curr_is_line) {
        if (segment_count > 0) {
            if (close) {
                join(prev_unit_normal, prev_pt, first_unit_normal, prev_is_line, curr_is_line, &outer, &inner);
                outer.close();
                Pt pt{0, 0};
                inner.last_point(pt);
                outer.move_to(pt.x, pt.y);
                outer.reverse_path_to(inner);
                outer.close();
            } else {
                Pt pt{0, 0};
                inner.last_point(pt);
                cap(cap_kind, prev_pt, prev_normal, pt, curr_is_line ? &inner : nullptr, outer);
                outer.reverse_path_to(inner);
                cap(cap_kind, first_pt, -first_normal, first_outer_pt, prev_is_line ? &inner : nullptr, outer);
                outer.close();
            }
            if (!cusper.empty()) {
                outer.push_path(cusper);
                cusper.clear();
            }
        }
        inner.clear();
        segment_count = -1;
        first_outer_pt_index = outer.pts.size();
    }

    void move_to(Pt p) {
        if (segment_count > 0) finish_contour(false, false);
        segment_count = 0;
        first_pt = p;
        prev_pt = p;
        join_completed = false;
    }

    void line_to(Pt p, bool has_valid_tangent_after) {
        const bool teeny = std::fabs(prev_pt.x - p.x) <= kNearlyZero * inv_res_scale &&
                           std::fabs(prev_pt.y - p.y) <= kNearlyZero * inv_res_scale;
        if (cap_kind == 0 && teeny) return;
        if (teeny && (join_completed || has_valid_tangent_after)) return;
        Pt normal, unit;
        if (!pre_join_to(p, true, normal, unit)) return;
        outer.line_to(p.x + normal.x, p.y + normal.y);
        inner.line_to(p.x - normal.x, p.y - normal.y);
        post_join_to(p, normal, unit);
    }

    void add_degenerate_line(const QuadConstruct &qp) {
        PathBuilder &path = stroke_type > 0.0f ? outer : inner;
        path.line_to(qp.quad[2].x, qp.quad[2].y);
    }

    void set_ray_points(Pt tp, Pt &dxy, Pt &on_p, Pt *tangent) {
        if (!set_length(dxy, radius)) dxy = Pt{radius, 0.0f};
        const float axis_flip = stroke_type;
        on_p.x = tp.x + axis_flip * dxy.y;
        on_p.y = tp.y - axis_flip * dxy.x;
        if (tangent) {
            tangent->x = on_p.x + dxy.x;
            tangent->y = on_p.y + dxy.y;
        }
    }

    void quad_perp_ray(const Pt q[3], float t, Pt &tp, Pt &on_p, Pt *tangent) {
        tp = eval_quad_at(q, t);
        Pt dxy = eval_quad_tangent_at(q, t);
        if (dxy.x == 0.0f && dxy.y == 0.0f) dxy = q[2] - q[0];
        set_ray_points(tp, dxy, on_p, tangent);
    }

    void cubic_perp_ray(const Pt c[4], float t, Pt &tp, Pt &on_p, Pt *tangent) {
        tp = eval_cubic_pos_at(c, t);
        Pt dxy = eval_cubic_tangent_at(c, t);
        Pt chopped[7];
        const Pt *cp = c;
        if (dxy.x == 0.0f && dxy.y == 0.0f) {
            if (is_nearly_zero(t)) {
                dxy = c[2] - c[0];
            } else if (is_nearly_zero(1.0f - t)) {
                dxy = c[3] - c[1];
            } else {
                chop_cubic_at2(c, t, chopped);
                dxy = chopped[3] - chopped[2];
                if (dxy.x == 0.0f && dxy.y == 0.0f) {
                    dxy = chopped[3] - chopped[1];
                    cp = chopped;
                }
            }
            if (dxy.x == 0.0f && dxy.y == 0.0f) dxy = cp[3] - cp[0];
        }
        set_ray_points(tp, dxy, on_p, tangent);
    }

    Result intersect_ray(bool ctrl_pt, QuadConstruct &qp) {
        const Pt start = qp.quad[0], end = qp.quad[2];
        const Pt a_len = qp.tangent_start - start, b_len = qp.tangent_end - end;
        const float denom = cross(a_len, b_len);
        if (denom == 0.0f || !std::isfinite(denom)) {
            qp.opposite_tangents = dot(a_len, b_len) < 0.0f;
            return Result::Degenerate;
        }
        qp.opposite_tangents = false;
        const Pt ab0 = start - end;
        float numer_a = cross(b_len, ab0);
        const float numer_b = cross(a_len, ab0);
        if ((numer_a >= 0.0f) == (numer_b >= 0.0f)) {
            const float dist1 = pt_to_line(start, end, qp.tangent_end);
            const float dist2 = pt_to_line(end, start, qp.tangent_start);
            if (std::max(dist1, dist2) <= inv_res_scale_sq) return Result::Degenerate;
            return Result::Split;
        }
        numer_a /= denom;
        const bool valid_divide = numer_a > numer_a - 1.0f;
        if (valid_divide) {
            if (ctrl_pt) {
                qp.quad[1].x = start.x * (1.0f - numer_a) + qp.tangent_start.x * numer_a;
                qp.quad[1].y = start.y * (1.0f - numer_a) + qp.tangent_start.y * numer_a;
            }
            return Result::Quad;
        }
        qp.opposite_tangents = dot(a_len, b_len) < 0.0f;
        return Result::Degenerate;
    }

    bool pt_in_quad_bounds(const Pt q[3], Pt pt) const {
        const float x_min = std::min(std::min(q[0].x, q[1].x), q[2].x);
        if (pt.x + inv_res_scale < x_min) return false;
        const float x_max = std::max(std::max(q[0].x, q[1].x), q[2].x);
        if (pt.x - inv_res_scale > x_max) return false;
        const float y_min = std::min(std::min(q[0].y, q[1].y), q[2].y);
        if (pt.y + inv_res_scale < y_min) return false;
        const float y_max = std::max(std::max(q[0].y, q[1].y), q[2].y);
        if (pt.y - inv_res_scale > y_max) return false;
        return true;
    }

    Result stroke_close_enough(const Pt stroke[3], const Pt ray[2], QuadConstruct &qp) {
        const Pt mid = eval_quad_at(stroke, 0.5f);
        if (points_within_dist(ray[0], mid, inv_res_scale)) return sharp_angle(qp.quad) ? Result::Split : Result::Quad;
        if (!pt_in_quad_bounds(stroke, ray[0])) return Result::Split;
        float roots[2];
        if (intersect_quad_ray(ray, stroke, roots) != 1) return Result::Split;
        const Pt quad_pt = eval_quad_at(stroke, roots[0]);
        const float error = inv_res_scale * (1.0f - std::fabs(roots[0] - 0.5f) * 2.0f);
        if (points_within_dist(ray[0], quad_pt, error)) return sharp_angle(qp.quad) ? Result::Split : Result::Quad;
        return Result::Split;
    }

    Result compare_quad_quad(const Pt q[3], QuadConstruct &qp) {
        if (!qp.start_set) {
            Pt tp;
            quad_perp_ray(q, qp.start_t, tp, qp.quad[0], &qp.tangent_start);
            qp.start_set = true;
        }
        if (!qp.end_set) {
            Pt tp;
            quad_perp_ray(q, qp.end_t, tp, qp.quad[2], &qp.tangent_end);
            qp.end_set = true;
        }
        const Result r = intersect_ray(true, qp);
        if (r != Result::Quad) return r;
        Pt ray[2];
        quad_perp_ray(q, qp.mid_t, ray[1], ray[0], nullptr);
        const Pt stroke[3] = {qp.quad[0], qp.quad[1], qp.quad[2]};
        return stroke_close_enough(stroke, ray, qp);
    }

    void cubic_quad_ends(const Pt c[4], QuadConstruct &qp) {
        if (!qp.start_set) {
            Pt tp;
            cubic_perp_ray(c, qp.start_t, tp, qp.quad[0], &qp.tangent_start);
            qp.start_set = true;
        }
        if (!qp.end_set) {
            Pt tp;
            cubic_perp_ray(c, qp.end_t, tp, qp.quad[2], &qp.tangent_end);
            qp.end_set = true;
        }
    }

    Result compare_quad_cubic(const Pt c[4], QuadConstruct &qp) {
        cubic_quad_ends(c, qp);
        const Result r = intersect_ray(true, qp);
        if (r != Result::Quad) return r;
        Pt ray[2];
        cubic_perp_ray(c, qp.mid_t, ray[1], ray[0], nullptr);
        const Pt stroke[3] = {qp.quad[0], qp.quad[1], qp.quad[2]};
        return stroke_close_enough(stroke, ray, qp);
    }

    bool cubic_mid_on_line(const Pt c[4], const QuadConstruct &qp) {
        Pt tp, mid;
        cubic_perp_ray(c, qp.mid_t, tp, mid, nullptr);
        return pt_to_line(mid, qp.quad[0], qp.quad[2]) < inv_res_scale_sq;
    }

    bool quad_stroke(const Pt q[3], QuadConstruct &qp) {
        const Result r = compare_quad_quad(q, qp);
        if (r == Result::Quad) {
            PathBuilder &path = stroke_type > 0.0f ? outer : inner;
            path.quad_to(qp.quad[1].x, qp.quad[1].y, qp.quad[2].x, qp.quad[2].y);
            return true;
        }
        if (r == Result::Degenerate) {
            add_degenerate_line(qp);
            return true;
        }
        if (++recursion_depth > kRecursiveLimits[kQuadLimit]) return false;
        QuadConstruct half;
        half.init_with_start(qp);
        if (!quad_stroke(q, half)) return false;
        half.init_with_end(qp);
        if (!quad_stroke(q, half)) return false;
        recursion_depth--;
        return true;
    }

    bool cubic_stroke(const Pt c[4], QuadConstruct &qp) {
        if (!found_tangents) {
            cubic_quad_ends(c, qp);
            const Result r = intersect_ray(false, qp);
            if (r != Result::Quad) {
                const bool close = points_within_dist(qp.quad[0], qp.quad[2], inv_res_scale);
                if ((r == Result::Degenerate || close) && cubic_mid_on_line(c, qp)) {
                    add_degenerate_line(qp);
                    return true;
                }
            } else {
                found_tangents = true;
            }
        }
        if (found_tangents) {
            const Result r = compare_quad_cubic(c, qp);
            if (r == Result::Quad) {
                PathBuilder &path = stroke_type > 0.0f ? outer : inner;
                path.quad_to(qp.quad[1].x, qp.quad[1].y, qp.quad[2].x, qp.quad[2].y);
                return true;
            }
            if (r == Result::Degenerate && !qp.opposite_tangents) {
                add_degenerate_line(qp);
                return true;
            }
        }
        if (!std::isfinite(qp.quad[2].x) || !std::isfinite(qp.quad[2].x)) return false;
        if (++recursion_depth > kRecursiveLimits[found_tangents ? 1 : kTangentLimit]) return false;
        QuadConstruct half;
        if (!half.init_with_start(qp)) {
            add_degenerate_line(qp);
            recursion_depth--;
            return true;
        }
        if (!cubic_stroke(c, half)) return false;
        if (!half.init_with_end(qp)) {
            add_degenerate_line(qp);
            recursion_depth--;
            return true;
        }
        if (!cubic_stroke(c, half)) return false;
        recursion_depth--;
        return true;
    }

    void init_quad(float type, float start, float end, QuadConstruct &qp) {
        stroke_type = type;
        found_tangents = false;
        qp.init(start, end);
    }

    void quad_to(Pt p1, Pt p2) {
        const Pt q[3] = {prev_pt, p1, p2};
        Pt reduction{};
        const Reduction rt = check_quad_linear(q, reduction);
        if (rt == Reduction::Point || rt == Reduction::Line) {
            line_to(p2, false);
            return;
        }
        if (rt == Reduction::Degenerate) {
            line_to(reduction, false);
            const bool saved = round_join_override;
            round_join_override = true;
            line_to(p2, false);
            round_join_override = saved;
            return;
        }
        Pt normal_ab, unit_ab, normal_bc{}, unit_bc{};
        if (!pre_join_to(p1, false, normal_ab, unit_ab)) {
            line_to(p2, false);
            return;
        }
        QuadConstruct qp;
        init_quad(1.0f, 0.0f, 1.0f, qp);
        quad_stroke(q, qp);
        init_quad(-1.0f, 0.0f, 1.0f, qp);
        quad_stroke(q, qp);
        set_normal_unit_normal(q[1], q[2], normal_bc, unit_bc);
        post_join_to(p2, normal_bc, unit_bc);
    }

    void set_cubic_end_normal(const Pt c[4], Pt normal_ab, Pt unit_ab, Pt &normal_cd, Pt &unit_cd) {
        Pt ab = c[1] - c[0], cd = c[3] - c[2];
        bool dab = degenerate_vector(ab), dcd = degenerate_vector(cd);
        if (dab && dcd) {
            normal_cd = normal_ab;
            unit_cd = unit_ab;
            return;
        }
        if (dab) {
            ab = c[2] - c[0];
            dab = degenerate_vector(ab);
        }
        if (dcd) {
            cd = c[3] - c[1];
            dcd = degenerate_vector(cd);
        }
        if (dab || dcd) {
            normal_cd = normal_ab;
            unit_cd = unit_ab;
            return;
        }
        set_normal_unit_normal2(cd, normal_cd, unit_cd);
    }

    void cubic_to(Pt p1, Pt p2, Pt p3) {
        const Pt c[4] = {prev_pt, p1, p2, p3};
        Pt reduction[3]{}, tangent_pt{};
        const Reduction rt = check_cubic_linear(c, reduction, tangent_pt);
        if (rt == Reduction::Point || rt == Reduction::Line) {
            line_to(p3, false);
            return;
        }
        if (rt == Reduction::Degenerate || rt == Reduction::Degenerate2 || rt == Reduction::Degenerate3) {
            line_to(reduction[0], false);
            const bool saved = round_join_override;
            round_join_override = true;
            if (rt == Reduction::Degenerate2 || rt == Reduction::Degenerate3) line_to(reduction[1], false);
            if (rt == Reduction::Degenerate3) line_to(reduction[2], false);
            line_to(p3, false);
            round_join_override = saved;
            return;
        }
        Pt normal_ab, unit_ab, normal_cd{}, unit_cd{};
        if (!pre_join_to(tangent_pt, false, normal_ab, unit_ab)) {
            line_to(p3, false);
            return;
        }
        float tv[2];
        const int n = find_cubic_inflections(c, tv);
        float last_t = 0.0f;
        for (int i = 0; i <= n; i++) {
            const float next_t = i < n ? tv[i] : 1.0f;
            QuadConstruct qp;
            init_quad(1.0f, last_t, next_t, qp);
            cubic_stroke(c, qp);
            init_quad(-1.0f, last_t, next_t, qp);
            cubic_stroke(c, qp);
            last_t = next_t;
        }
        const float cusp = find_cubic_cusp(c);
        if (cusp > 0.0f) {
            const Pt loc = eval_cubic_pos_at(c, cusp);
            cusper.push_circle(loc.x, loc.y, radius);
        }
        set_cubic_end_normal(c, normal_ab, unit_ab, normal_cd, unit_cd);
        post_join_to(p3, normal_cd, unit_cd);
    }

    bool is_current_contour_empty() const {
        return inner.is_zero_length_since_point(0) && outer.is_zero_length_since_point(first_outer_pt_index);
    }
};

// Does the rest of the contour after verb index `vi` (with the auto-closing line
// already accounted for by the caller) contain a segment of non-zero length?
// (SkStroke's has_valid_tangent: looks ahead to the next move or close.)
bool has_valid_tangent(const uint8_t *verbs, size_t nverbs, const Pt *pts, size_t vi, size_t pi, Pt last,
                       Pt move_pt) {
    for (size_t i = vi; i < nverbs; i++) {
        switch (verbs[i]) {
            case kMove: return false;
            case kLine:
                if (pts[pi] == last) { pi += 1; continue; }
                return true;
            case kQuad:
                if (pts[pi] == last && pts[pi + 1] == last) { last = pts[pi + 1]; pi += 2; continue; }
                return true;
            case kCubic:
                if (pts[pi] == last && pts[pi + 1] == last && pts[pi + 2] == last) { pi += 3; continue; }
                return true;
            case kClose: return last != move_pt;  // the auto-closing line
            default: return false;
        }
    }
    return false;
}

}  // namespace

// ---------------------------------------------------------------------------
// PathBuilder
// ---------------------------------------------------------------------------
void PathBuilder::move_to(float x, float y) {
    if (!verbs.empty() && verbs.back() == kMove) {
        pts.back() = Pt{x, y};
    } else {
        last_move_to_index = pts.size();
        move_to_required = false;
        verbs.push_back(kMove);
        pts.push_back(Pt{x, y});
    }
}

void PathBuilder::inject_move_to_if_needed() {
    if (!move_to_required) return;
    if (last_move_to_index < pts.size()) {
        const Pt p = pts[last_move_to_index];
        move_to(p.x, p.y);
    } else {
        move_to(0.0f, 0.0f);
    }
}

void PathBuilder::line_to(float x, float y) {
    inject_move_to_if_needed();
    verbs.push_back(kLine);
    pts.push_back(Pt{x, y});
}

void PathBuilder::quad_to(float x1, float y1, float x2, float y2) {
    inject_move_to_if_needed();
    verbs.push_back(kQuad);
    pts.push_back(Pt{x1, y1});
    pts.push_back(Pt{x2, y2});
}

void PathBuilder::cubic_to(float x1, float y1, float x2, float y2, float x3, float y3) {
    inject_move_to_if_needed();
    verbs.push_back(kCubic);
    pts.push_back(Pt{x1, y1});
    pts.push_back(Pt{x2, y2});
    pts.push_back(Pt{x3, y3});
}

void PathBuilder::conic_points_to(Pt p1, Pt p2, float weight) {
    if (!(weight > 0.0f)) {
        line_to(p2.x, p2.y);
    } else if (!std::isfinite(weight)) {
        line_to(p1.x, p1.y);
        line_to(p2.x, p2.y);
    } else if (weight == 1.0f) {
        quad_to(p1.x, p1.y, p2.x, p2.y);
    } else {
        inject_move_to_if_needed();
        Pt last{0, 0};
        last_point(last);
        Pt q[33];
        const int n = conic_to_quads(last, p1, p2, weight, q);
        for (int i = 0; i < n; i++) quad_to(q[1 + 2 * i].x, q[1 + 2 * i].y, q[2 + 2 * i].x, q[2 + 2 * i].y);
    }
}

void PathBuilder::close() {
    if (!verbs.empty() && verbs.back() != kClose) verbs.push_back(kClose);
    move_to_required = true;
}

void PathBuilder::set_last_point(Pt p) {
    if (pts.empty()) move_to(p.x, p.y);
    else pts.back() = p;
}

bool PathBuilder::last_point(Pt &out) const {
    if (pts.empty()) return false;
    out = pts.back();
    return true;
}

bool PathBuilder::is_zero_length_since_point(size_t start) const {
    if (pts.size() < start + 2) return true;
    const Pt first = pts[start];
    for (size_t i = start + 1; i < pts.size(); i++)
        if (first != pts[i]) return false;
    return true;
}

void PathBuilder::reverse_path_to(const PathBuilder &other) {
    if (other.verbs.empty()) return;
    size_t off = other.pts.size() - 1;
    for (size_t vi = other.verbs.size(); vi-- > 0;) {
        switch (other.verbs[vi]) {
            case kMove: return;
            case kLine: {
                const Pt p = other.pts[off - 1];
                off -= 1;
                line_to(p.x, p.y);
                break;
            }
            case kQuad: {
                const Pt p1 = other.pts[off - 1], p2 = other.pts[off - 2];
                off -= 2;
                quad_to(p1.x, p1.y, p2.x, p2.y);
                break;
            }
            case kCubic: {
                const Pt p1 = other.pts[off - 1], p2 = other.pts[off - 2], p3 = other.pts[off - 3];
                off -= 3;
                cubic_to(p1.x, p1.y, p2.x, p2.y, p3.x, p3.y);
                break;
            }
            default: break;
        }
    }
}

void PathBuilder::push_path(const PathBuilder &other) {
    last_move_to_index = pts.size() + other.last_move_to_index;
    verbs.insert(verbs.end(), other.verbs.begin(), other.verbs.end());
    pts.insert(pts.end(), other.pts.begin(), other.pts.end());
}

void PathBuilder::push_circle(float x, float y, float r) {
    const float l = x - r, t = y - r, w = r + r, h = r + r;
    if (!(w > 0.0f) || !(h > 0.0f) || !std::isfinite(l + w) || !std::isfinite(t + h)) return;
    const float rr = l + w, bb = t + h;  // Rect::from_xywh -> right, bottom
    const float cx = l * 0.5f + rr * 0.5f, cy = t * 0.5f + bb * 0.5f;
    const Pt oval[4] = {{cx, bb}, {l, cy}, {cx, t}, {rr, cy}};
    const Pt rect[4] = {{rr, bb}, {l, bb}, {l, t}, {rr, t}};
    move_to(oval[3].x, oval[3].y);
    for (int i = 0; i < 4; i++) conic_points_to(rect[i], oval[i], kRoot2Over2);
    close();
}

// ---------------------------------------------------------------------------
int cubic_max_curvature(const Pt src[4], float t[3]) { return find_cubic_max_curvature(src, t); }
void chop_cubic_at(const Pt src[4], float t, Pt dst[7]) { chop_cubic_at2(src, t, dst); }
bool unit_divide(float numer, float denom, float &r) { return valid_unit_divide(numer, denom, r); }
void transform_points(float sx, float ky, float kx, float sy, float tx, float ty, Pt *p, size_t n) {
    map_points(Ts{sx, ky, kx, sy, tx, ty}, p, n);
}

float resolution_scale(float sx, float ky, float kx, float sy) {
    auto length = [](float x, float y) {
        const float m2 = x * x + y * y;
        if (std::isfinite(m2)) return std::sqrt(m2);
        return (float)std::sqrt((double)x * x + (double)y * y);
    };
    const float a = length(sx, kx), b = length(ky, sy);
    if (std::isfinite(a) && std::isfinite(b)) {
        const float s = std::max(a, b);
        if (s > 0.0f) return s;
    }
    return 1.0f;
}

bool treat_as_hairline(float width, bool anti_alias, float sx, float ky, float kx, float sy, float &coverage) {
    if (width == 0.0f) {
        coverage = 1.0f;
        return true;
    }
    if (!anti_alias) return false;
    auto fast_len = [](float x, float y) {
        x = std::fabs(x);
        y = std::fabs(y);
        if (x < y) std::swap(x, y);
        return x + y * 0.5f;
    };
    Pt p0{width, 0.0f}, p1{0.0f, width};
    Ts t{sx, ky, kx, sy, 0.0f, 0.0f};
    map_points(t, &p0, 1);
    map_points(t, &p1, 1);
    const float l0 = fast_len(p0.x, p0.y), l1 = fast_len(p1.x, p1.y);
    if (l0 <= 1.0f && l1 <= 1.0f) {
        coverage = (l0 + l1) * 0.5f;
        return true;
    }
    return false;
}

bool stroke(const uint8_t *verbs, size_t nverbs, const float *points, size_t npoints, const StrokeParams &sp,
            float res_scale, PathBuilder &out) {
    out.clear();
    if (!(sp.width > 0.0f) || !std::isfinite(sp.width)) return false;
    Stroker s;
    uint32_t join = sp.join;
    float inv_miter = 0.0f;
    if (join == 0) {
        if (sp.miter_limit <= 1.0f) join = 3;
        else inv_miter = 1.0f / sp.miter_limit;
    }
    if (join == 1) inv_miter = 1.0f / sp.miter_limit;
    s.res_scale = res_scale;
    s.inv_res_scale = 1.0f / (res_scale * 4.0f);
    s.inv_res_scale_sq = s.inv_res_scale * s.inv_res_scale;
    s.radius = sp.width * 0.5f;
    s.inv_miter_limit = inv_miter;
    s.cap_kind = sp.cap;
    s.join_kind = join;
    const Pt *pts = reinterpret_cast<const Pt *>(points);
    // the outline holds a few quadratics per input segment and side
    s.outer.verbs.reserve(nverbs * 8);
    s.outer.pts.reserve(npoints * 16);
    s.inner.verbs.reserve(nverbs * 4);
    s.inner.pts.reserve(npoints * 8);
    bool last_is_line = false;
    size_t pi = 0;
    Pt move_pt{0, 0}, last{0, 0};
    for (size_t vi = 0; vi < nverbs; vi++) {
        switch (verbs[vi]) {
            case kMove:
                move_pt = last = pts[pi++];
                s.move_to(move_pt);
                break;
            case kLine:
                last = pts[pi];
                s.line_to(pts[pi], has_valid_tangent(verbs, nverbs, pts, vi + 1, pi + 1, pts[pi], move_pt));
                pi += 1;
                last_is_line = true;
                break;
            case kQuad:
                s.quad_to(pts[pi], pts[pi + 1]);
                last = pts[pi + 1];
                pi += 2;
                last_is_line = false;
                break;
            case kCubic:
                s.cubic_to(pts[pi], pts[pi + 1], pts[pi + 2]);
                last = pts[pi + 2];
                pi += 3;
                last_is_line = false;
                break;
            case kClose:
                if (last != move_pt) {  // the segment iterator's auto-closing line
                    s.line_to(move_pt, false);
                    last = move_pt;
                    last_is_line = true;
                }
                if (sp.cap != 0) {
                    if (s.segment_count == 0) {  // has_only_move_to
                        s.line_to(s.first_pt, false);
                        last_is_line = true;
                        break;
                    }
                    if (s.is_current_contour_empty()) {
                        last_is_line = true;
                        break;
                    }
                }
                s.finish_contour(true, last_is_line);
                break;
            default: return false;
        }
    }
    s.finish_contour(false, last_is_line);
    out = std::move(s.outer);
    // PathBuilder::finish: nothing, or a lone move_to, is no path
    return out.verbs.size() > 1;
}

}  // namespace rvks
