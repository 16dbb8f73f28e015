// Host front end of the rasterizer: path iteration, Skia edge stepping expanded
// into line pieces, stroke outlines (stroker.cpp) and shader setup.
//
// Edge arithmetic restates tiny-skia 0.12 (Skia's SkEdge: FDot6 inputs on the 4x
// supersampled grid, 16.16 slopes, quadratic and cubic forward differencing with
// MAX_COEFF_SHIFT 6, curves chopped at Y extrema first).  Every update step of a
// curve edge is run here once and each non-empty line piece is kept, which is the
// sequence tiny-skia's scanline walk visits lazily.
//
// Restates tiny-skia (BSD-3-Clause, Copyright 2020 Yevhenii Reizner) and Skia
// (BSD-3-Clause, Copyright Google Inc.); see THIRD_PARTY_NOTICES.md.
#include "raster.h"
#include "hairline.h"
#include "stroker.h"
#include <algorithm>
#include <array>
#include <cmath>

using namespace rvkh;

namespace rvkr {

// ---------------------------------------------------------------------------
// path iteration
// ---------------------------------------------------------------------------
void fill_segments(const rvk_scene &sc, const rvk_path &path, Ts ts, std::vector<Seg> &out) {
    const uint8_t *verbs = sc.verbs + path.verb_offset;
    const float *pts = sc.points + 2 * (size_t)path.point_offset;
    out.reserve(out.size() + path.verb_count);
    size_t pi = 0;
    Pt move{0, 0}, last{0, 0};
    bool open = false;
    auto next = [&]() {
        Pt p{pts[2 * pi], pts[2 * pi + 1]};
        pi++;
        ts_map(ts, p.x, p.y);
        return p;
    };
    auto close = [&]() {
        if (open && (last.x != move.x || last.y != move.y)) out.push_back(Seg{2, {last, move}});
        open = false;
    };
    for (uint32_t v = 0; v < path.verb_count; v++) {
        const uint8_t verb = verbs[v];
        if (verb == RVK_VERB_MOVE) {
            close();
            move = last = next();
            open = true;
        } else if (verb == RVK_VERB_LINE) {
            Pt p = next();
            out.push_back(Seg{2, {last, p}});
            last = p;
        } else if (verb == RVK_VERB_QUAD) {
            Seg s{3, {last}};
            s.p[1] = next();
            s.p[2] = next();
            out.push_back(s);
            last = s.p[2];
        } else if (verb == RVK_VERB_CUBIC) {
            Seg s{4, {last}};
            s.p[1] = next();
            s.p[2] = next();
            s.p[3] = next();
            out.push_back(s);
            last = s.p[3];
        } else {  // close
            close();
            last = move;
        }
    }
    close();
}

bool segs_drawable(const std::vector<Seg> &segs) {
    // Rect::from_points (non-finite -> None, so Path::transform fails) and
    // is_too_big_for_math (bounds beyond f32::MAX / 4)
    const float lim = 3.40282347e+38f * 0.25f;
    for (const Seg &s : segs)
        for (int i = 0; i < s.n; i++)
            if (!(s.p[i].x >= -lim && s.p[i].x <= lim && s.p[i].y >= -lim && s.p[i].y <= lim)) return false;
    return !segs.empty();
}

bool bounds_nearly_zero(const std::vector<Seg> &segs) {
    float mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
    for (const Seg &s : segs)
        for (int i = 0; i < s.n; i++) {
            mnx = std::min(mnx, s.p[i].x); mxx = std::max(mxx, s.p[i].x);
            mny = std::min(mny, s.p[i].y); mxy = std::max(mxy, s.p[i].y);
        }
    const float nz = 1.0f / 4096.0f;
    return !(mxx - mnx > nz) || !(mxy - mny > nz);
}

// ---------------------------------------------------------------------------
// fixed-point edge stepping
// ---------------------------------------------------------------------------
namespace {

inline int32_t fdot6_round(int32_t x) { return (x + 32) >> 6; }
inline int32_t fdot6_to_fixed(int32_t x) { return (int32_t)((uint32_t)x << 10); }
inline int32_t fixed_mul(int32_t a, int32_t b) { return (int32_t)(((int64_t)a * (int64_t)b) >> 16); }
inline int32_t fdot6_div(int32_t a, int32_t b) {
    if ((int32_t)(int16_t)a == a) return (int32_t)((uint32_t)a << 16) / b;
    int64_t v = (int64_t)a * 65536 / (int64_t)b;
    return (int32_t)std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, v));
}
inline int32_t compute_dy(int32_t top, int32_t y0) {  // ((top << 6) + 32) - y0, wrapping
    return (int32_t)(((uint32_t)top << 6) + 32u - (uint32_t)y0);
}
inline int32_t diff_to_shift(int32_t dx, int32_t dy, int32_t shift_aa) {
    dx = std::abs(dx);
    dy = std::abs(dy);
    uint32_t dist = (uint32_t)(dx > dy ? dx + (dy >> 1) : dy + (dx >> 1));
    dist = (dist + (1u << 4)) >> (3 + shift_aa);
    return (32 - (dist ? __builtin_clz(dist) : 32)) >> 1;
}
inline int32_t cubic_delta_from_line(int32_t a, int32_t b, int32_t c, int32_t d) {
    int32_t one_third = ((a * 8 - b * 15 + 6 * c + d) * 19) >> 9;
    int32_t two_third = ((a + 6 * b - c * 15 + d * 8) * 19) >> 9;
    return std::max(std::abs(one_third), std::abs(two_third));
}
inline int32_t fdot6_of(float v, float scale) { return as_i32(v * scale); }
constexpr int kMaxCoeffShift = 6;

// SkEdge::updateLine on 16.16 inputs; emits the piece when it spans a row
bool line_piece(int32_t x0, int32_t y0, int32_t x1, int32_t y1, PieceDev &p) {
    y0 >>= 10;
    y1 >>= 10;
    int32_t top = fdot6_round(y0), bot = fdot6_round(y1);
    if (top == bot) return false;
    x0 >>= 10;
    x1 >>= 10;
    int32_t slope = fdot6_div(x1 - x0, y1 - y0);
    p.x = fdot6_to_fixed(x0 + fixed_mul(slope, compute_dy(top, y0)));
    p.dx = slope;
    p.first_y = top;
    p.last_y = bot - 1;
    return true;
}

// SkQuadraticEdge::updateQuadratic: the next non-empty piece, or false when done
bool quad_update(EdgeRec &e) {
    int32_t count = e.count, oldx = e.cx, oldy = e.cy, dx = e.d1x, dy = e.d1y, newx = oldx, newy = oldy;
    PieceDev pc;
    bool ok;
    do {
        if (--count > 0) {
            newx = oldx + (dx >> e.shift);
            dx += e.d2x;
            newy = oldy + (dy >> e.shift);
            dy += e.d2y;
        } else {
            newx = e.lx;
            newy = e.ly;
        }
        ok = line_piece(oldx, oldy, newx, newy, pc);
        oldx = newx;
        oldy = newy;
    } while (count > 0 && !ok);
    e.cx = newx; e.cy = newy; e.d1x = dx; e.d1y = dy; e.count = count;
    if (ok) { e.x = pc.x; e.dx = pc.dx; e.first_y = pc.first_y; e.last_y = pc.last_y; }
    return ok;
}

// SkCubicEdge::updateCubic
bool cubic_update(EdgeRec &e) {
    int32_t count = e.count, oldx = e.cx, oldy = e.cy, newx = oldx, newy = oldy;
    PieceDev pc;
    bool ok;
    do {
        if (++count < 0) {
            newx = oldx + (e.d1x >> e.dshift);
            e.d1x += e.d2x >> e.shift;
            e.d2x += e.d3x;
            newy = oldy + (e.d1y >> e.dshift);
            e.d1y += e.d2y >> e.shift;
            e.d2y += e.d3y;
        } else {
            newx = e.lx;
            newy = e.ly;
        }
        if (newy < oldy) newy = oldy;
        ok = line_piece(oldx, oldy, newx, newy, pc);
        oldx = newx;
        oldy = newy;
    } while (count < 0 && !ok);
    e.cx = newx; e.cy = newy; e.count = count;
    if (ok) { e.x = pc.x; e.dx = pc.dx; e.first_y = pc.first_y; e.last_y = pc.last_y; }
    return ok;
}

struct Builder {
    int shift;
    std::vector<EdgeRec> &out;
    std::vector<uint32_t> &caps;

    void push(EdgeRec &e, int32_t w, uint32_t cap) {
        e.winding = w;
        out.push_back(e);
        caps.push_back(cap);
    }

    // LineEdge::new + push_line's combine_vertical (tiny-skia edge_builder.rs)
    void line(Pt a, Pt b) {
        float scale = (float)(1 << (shift + 6));
        int32_t x0 = fdot6_of(a.x, scale), y0 = fdot6_of(a.y, scale), x1 = fdot6_of(b.x, scale),
                y1 = fdot6_of(b.y, scale);
        int32_t w = 1;
        if (y0 > y1) { std::swap(x0, x1); std::swap(y0, y1); w = -1; }
        int32_t top = fdot6_round(y0), bot = fdot6_round(y1);
        if (top == bot) return;
        int32_t slope = fdot6_div(x1 - x0, y1 - y0);
        EdgeRec e{};
        e.x = fdot6_to_fixed(x0 + fixed_mul(slope, compute_dy(top, y0)));
        e.dx = slope;
        e.first_y = top;
        e.last_y = bot - 1;
        e.type = 0;
        if (e.dx == 0 && !out.empty() && out.back().type == 0) {
            EdgeRec &l = out.back();
            int32_t &lw = l.winding;
            if (l.dx == 0 && e.x == l.x) {
                if (w == lw) {
                    if (e.last_y + 1 == l.first_y) { l.first_y = e.first_y; return; }
                    if (e.first_y == l.last_y + 1) { l.last_y = e.last_y; return; }
                } else if (e.first_y == l.first_y) {
                    if (e.last_y == l.last_y) {  // the two cancel: drop the previous edge
                        out.pop_back();
                        caps.pop_back();
                        return;
                    }
                    if (e.last_y < l.last_y) { l.first_y = e.last_y + 1; return; }
                    l.first_y = l.last_y + 1;
                    l.last_y = e.last_y;
                    lw = w;
                    return;
                } else if (e.last_y == l.last_y) {
                    if (e.first_y > l.first_y) { l.last_y = e.first_y - 1; return; }
                    l.last_y = l.first_y - 1;
                    l.first_y = e.first_y;
                    lw = w;
                    return;
                }
            }
        }
        push(e, w, 1);
    }

    // QuadraticEdge::new (setup + first update)
    void quad(const Pt *p) {
        float scale = (float)(1 << (shift + 6));
        int32_t x0 = fdot6_of(p[0].x, scale), y0 = fdot6_of(p[0].y, scale), x1 = fdot6_of(p[1].x, scale),
                y1 = fdot6_of(p[1].y, scale), x2 = fdot6_of(p[2].x, scale), y2 = fdot6_of(p[2].y, scale);
        int32_t w = 1;
        if (y0 > y2) { std::swap(x0, x2); std::swap(y0, y2); w = -1; }
        if (fdot6_round(y0) == fdot6_round(y2)) return;
        int32_t sh = diff_to_shift(((int32_t)((uint32_t)x1 << 1) - x0 - x2) >> 2,
                                   ((int32_t)((uint32_t)y1 << 1) - y0 - y2) >> 2, shift);
        if (sh == 0) sh = 1;
        else if (sh > kMaxCoeffShift) sh = kMaxCoeffShift;
        EdgeRec e{};
        e.type = 1;
        e.count = 1 << sh;
        e.shift = sh - 1;
        int32_t A = (int32_t)((uint32_t)(x0 - x1 - x1 + x2) << 9), B = fdot6_to_fixed(x1 - x0);
        e.cx = fdot6_to_fixed(x0);
        e.d1x = B + (A >> sh);
        e.d2x = A >> (sh - 1);
        A = (int32_t)((uint32_t)(y0 - y1 - y1 + y2) << 9);
        B = fdot6_to_fixed(y1 - y0);
        e.cy = fdot6_to_fixed(y0);
        e.d1y = B + (A >> sh);
        e.d2y = A >> (sh - 1);
        e.lx = fdot6_to_fixed(x2);
        e.ly = fdot6_to_fixed(y2);
        if (!quad_update(e)) return;
        push(e, w, 1u + (uint32_t)std::max(e.count, 0));
    }

    // CubicEdge::new (setup + first update)
    void cubic(const Pt *p) {
        float scale = (float)(1 << (shift + 6));
        int32_t x0 = fdot6_of(p[0].x, scale), y0 = fdot6_of(p[0].y, scale), x1 = fdot6_of(p[1].x, scale),
                y1 = fdot6_of(p[1].y, scale), x2 = fdot6_of(p[2].x, scale), y2 = fdot6_of(p[2].y, scale),
                x3 = fdot6_of(p[3].x, scale), y3 = fdot6_of(p[3].y, scale);
        int32_t w = 1;
        if (y0 > y3) {
            std::swap(x0, x3); std::swap(x1, x2); std::swap(y0, y3); std::swap(y1, y2);
            w = -1;
        }
        if (fdot6_round(y0) == fdot6_round(y3)) return;
        int32_t sh = diff_to_shift(cubic_delta_from_line(x0, x1, x2, x3), cubic_delta_from_line(y0, y1, y2, y3),
                                   shift) + 1;
        if (sh > kMaxCoeffShift) sh = kMaxCoeffShift;
        int32_t up = 6, down = sh + up - 10;
        if (down < 0) { down = 0; up = 10 - sh; }
        auto upsh = [&](int32_t v) { return (int32_t)((uint32_t)v << up); };
        EdgeRec e{};
        e.type = 2;
        e.count = (int32_t)((uint32_t)-1 << sh);
        e.shift = sh;
        e.dshift = down;
        int32_t B = upsh(3 * (x1 - x0)), C = upsh(3 * (x0 - x1 - x1 + x2)), D = upsh(x3 + 3 * (x1 - x2) - x0);
        e.cx = fdot6_to_fixed(x0);
        e.d1x = B + (C >> sh) + (D >> (2 * sh));
        e.d2x = 2 * C + ((3 * D) >> (sh - 1));
        e.d3x = (3 * D) >> (sh - 1);
        B = upsh(3 * (y1 - y0)); C = upsh(3 * (y0 - y1 - y1 + y2)); D = upsh(y3 + 3 * (y1 - y2) - y0);
        e.cy = fdot6_to_fixed(y0);
        e.d1y = B + (C >> sh) + (D >> (2 * sh));
        e.d2y = 2 * C + ((3 * D) >> (sh - 1));
        e.d3y = (3 * D) >> (sh - 1);
        e.lx = fdot6_to_fixed(x3);
        e.ly = fdot6_to_fixed(y3);
        if (!cubic_update(e)) return;
        push(e, w, 1u + (uint32_t)std::max(-e.count, 0));
    }
};

// chopping at Y extrema (tiny-skia-path path_geometry)
bool unit_divide(float numer, float denom, float &r) {
    if (numer < 0.0f) { numer = -numer; denom = -denom; }
    if (denom == 0.0f || numer == 0.0f || numer >= denom) return false;
    r = numer / denom;
    return std::isfinite(r) && r > 0.0f && r < 1.0f;
}
inline Pt lerp_pt(Pt a, Pt b, float t) { return Pt{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t}; }

int quad_y_monotonic(const Pt *s, Pt *d) {
    float a = s[0].y, b = s[1].y, c = s[2].y;
    float ab = a - b, bc = b - c;
    if (ab < 0.0f) bc = -bc;
    if (ab == 0.0f || bc < 0.0f) {
        float t;
        if (unit_divide(a - b, a - b - b + c, t)) {
            Pt p01 = lerp_pt(s[0], s[1], t), p12 = lerp_pt(s[1], s[2], t);
            d[0] = s[0]; d[1] = p01; d[2] = lerp_pt(p01, p12, t); d[3] = p12; d[4] = s[2];
            d[1].y = d[2].y;
            d[3].y = d[2].y;
            return 1;
        }
        b = std::fabs(a - b) < std::fabs(b - c) ? a : c;
    }
    d[0] = Pt{s[0].x, a}; d[1] = Pt{s[1].x, b}; d[2] = Pt{s[2].x, c};
    return 0;
}

int unit_quad_roots(float a, float b, float c, float *roots) {
    if (a == 0.0f) {
        float r;
        if (unit_divide(-c, b, r)) { roots[0] = r; return 1; }
        return 0;
    }
    double dr = (double)b * (double)b - 4.0 * (double)a * (double)c;
    if (dr < 0.0) return 0;
    float r = (float)std::sqrt(dr);
    if (!std::isfinite(r)) return 0;
    float q = b < 0.0f ? -(b - r) / 2.0f : -(b + r) / 2.0f;
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

void chop_cubic(const Pt *s, float t, Pt *d) {
    Pt ab = lerp_pt(s[0], s[1], t), bc = lerp_pt(s[1], s[2], t), cd = lerp_pt(s[2], s[3], t);
    Pt abc = lerp_pt(ab, bc, t), bcd = lerp_pt(bc, cd, t);
    d[0] = s[0]; d[1] = ab; d[2] = abc; d[3] = lerp_pt(abc, bcd, t); d[4] = bcd; d[5] = cd; d[6] = s[3];
}

int cubic_y_monotonic(const Pt *src, Pt *dst) {
    float a = src[0].y, b = src[1].y, c = src[2].y, d = src[3].y;
    float ts[3];
    int n = unit_quad_roots(d - a + 3.0f * (b - c), 2.0f * (a - b - b + c), b - a, ts);
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
    dst[2].y = dst[3].y;
    dst[4].y = dst[3].y;
    if (n == 2) { dst[5].y = dst[6].y; dst[7].y = dst[6].y; }
    return n;
}

}  // namespace

void build_edges(const std::vector<Seg> &segs, int shift, std::vector<EdgeRec> &out, std::vector<uint32_t> &caps,
                 bool monotonic) {
    Builder b{shift, out, caps};
    // a line gives at most one edge, a quadratic two, a cubic three
    size_t bound = 0;
    for (const Seg &s : segs) bound += (size_t)(s.n - 1);
    out.reserve(out.size() + bound);
    caps.reserve(caps.size() + bound);
    for (const Seg &s : segs) {
        if (s.n == 2) {
            b.line(s.p[0], s.p[1]);
        } else if (monotonic) {  // clipped pieces are pushed as they are (edge_builder.rs build)
            if (s.n == 3) b.quad(s.p);
            else b.cubic(s.p);
        } else if (s.n == 3) {
            Pt mono[5];
            int n = quad_y_monotonic(s.p, mono);
            for (int i = 0; i <= n; i++) b.quad(mono + 2 * i);
        } else {
            Pt mono[10];
            int n = cubic_y_monotonic(s.p, mono);
            for (int i = 0; i <= n; i++) b.cubic(mono + 3 * i);
        }
    }
}

// ---------------------------------------------------------------------------
// Strokes: the path or its dashes (dash.cpp), tiny-skia's PathStroker (stroker.cpp)
// in user space, the outline mapped with the transform and handed to the fill path
// (Pixmap::stroke_path -> fill_path).
// ---------------------------------------------------------------------------
namespace {

// Pixmap::stroke_path's input: the path, or its dashes when the stroke has a dash
// array that StrokeDash::new accepts.  False when dashing leaves nothing to draw.
bool stroke_input(const rvk_scene &sc, const rvk_path &path, float res, rvks::PathBuilder &dashed,
                  const uint8_t *&verbs, size_t &nverbs, const float *&pts, size_t &npts) {
    verbs = sc.verbs + path.verb_offset;
    nverbs = path.verb_count;
    pts = sc.points + 2 * (size_t)path.point_offset;
    npts = 0;
    for (uint32_t v = 0; v < path.verb_count; v++)
        npts += verbs[v] == RVK_VERB_MOVE || verbs[v] == RVK_VERB_LINE ? 1 : verbs[v] == RVK_VERB_QUAD ? 2
              : verbs[v] == RVK_VERB_CUBIC ? 3 : 0;
    if (path.dash_count == 0 || !sc.dashes) return true;
    rvks::DashParams dp;
    if (!rvks::dash_params(sc.dashes + path.dash_offset, path.dash_count, path.dash_phase, dp)) return true;
    if (!rvks::dash(verbs, nverbs, pts, npts, dp, res, dashed)) return false;
    verbs = dashed.verbs.data();
    nverbs = dashed.verbs.size();
    pts = &dashed.pts[0].x;
    npts = dashed.pts.size();
    return true;
}

}  // namespace

void stroke_segments(const rvk_scene &sc, const rvk_path &path, Ts ts, std::vector<Seg> &out) {
    rvks::StrokeParams sp{path.stroke_width, path.stroke_miter_limit, path.stroke_cap, path.stroke_join};
    const float res = rvks::resolution_scale(ts.sx, ts.ky, ts.kx, ts.sy);
    rvks::PathBuilder dashed;
    const uint8_t *verbs;
    const float *pts;
    size_t nverbs, npts;
    if (!stroke_input(sc, path, res, dashed, verbs, nverbs, pts, npts)) return;
    rvks::PathBuilder outline;
    if (!rvks::stroke(verbs, nverbs, pts, npts, sp, res, outline)) return;
    out.reserve(out.size() + outline.verbs.size());
    size_t pi = 0;
    Pt move{0, 0}, last{0, 0};
    bool open = false;
    auto P = [&](size_t i) {
        float x = outline.pts[i].x, y = outline.pts[i].y;
        ts_map(ts, x, y);
        return Pt{x, y};
    };
    auto close = [&]() {
        if (open && (last.x != move.x || last.y != move.y)) out.push_back(Seg{2, {last, move}});
        open = false;
    };
    for (uint8_t v : outline.verbs) {
        switch (v) {
            case rvks::kMove:
                close();
                move = last = P(pi++);
                open = true;
                break;
            case rvks::kLine: {
                const Pt p = P(pi++);
                out.push_back(Seg{2, {last, p}});
                last = p;
                break;
            }
            case rvks::kQuad: {
                const Pt a = P(pi), b = P(pi + 1);
                pi += 2;
                out.push_back(Seg{3, {last, a, b}});
                last = b;
                break;
            }
            case rvks::kCubic: {
                const Pt a = P(pi), b = P(pi + 1), c = P(pi + 2);
                pi += 3;
                out.push_back(Seg{4, {last, a, b, c}});
                last = c;
                break;
            }
            default:
                close();
                last = move;
                break;
        }
    }
    close();
}

// ---------------------------------------------------------------------------
// Hairlines: the shared segment walk (hairline.cpp), then do_anti_hairline's setup
// (tiny-skia scan/hairline_aa.rs): halving of lines longer than 511 px, the major
// axis, the start position and slope, the cap scales and the clip adjustments.  The
// per-pixel blits are evaluated on the device from the resulting HairLine.
// ---------------------------------------------------------------------------
namespace {

void prepare_anti_hairline(int32_t x0, int32_t y0, int32_t x1, int32_t y1, const rvks::HairSeg &seg,
                           std::vector<HairLine> &out) {
    if (x0 == INT32_MIN || y0 == INT32_MIN || x1 == INT32_MIN || y1 == INT32_MIN) return;
    if (std::abs(x1 - x0) > 511 * 64 || std::abs(y1 - y0) > 511 * 64) {
        const int32_t hx = (x0 >> 1) + (x1 >> 1), hy = (y0 >> 1) + (y1 >> 1);
        prepare_anti_hairline(x0, y0, hx, hy, seg, out);
        prepare_anti_hairline(hx, hy, x1, y1, seg, out);
        return;
    }
    auto fdot6_floor = [](int32_t v) { return v >> 6; };
    auto fdot6_ceil = [](int32_t v) { return (v + 63) >> 6; };
    auto fixed_floor = [](int32_t v) { return v >> 16; };
    auto fixed_ceil = [](int32_t v) { return (int32_t)(((int64_t)v + 0xffff) >> 16); };
    HairLine L{};
    const bool ymajor = !(std::abs(x1 - x0) > std::abs(y1 - y0));
    // work in (major, minor) coordinates
    int32_t a0 = ymajor ? y0 : x0, b0 = ymajor ? x0 : y0, a1 = ymajor ? y1 : x1, b1 = ymajor ? x1 : y1;
    if (a0 > a1) { std::swap(a0, a1); std::swap(b0, b1); }
    if (ymajor && x0 == x1 && y0 == y1) return;  // zero length
    int32_t istart = fdot6_floor(a0), istop = fdot6_ceil(a1);
    int32_t fstart = (int32_t)((uint32_t)b0 << 10), slope = 0;
    if (b0 != b1) {
        slope = (int32_t)((uint32_t)(b1 - b0) << 16) / (a1 - a0);
        fstart += (slope * (32 - (a0 & 63)) + 32) >> 6;
    }
    int32_t scale_start, scale_stop;
    if (istop - istart == 1) {
        scale_start = a1 - a0;
        scale_stop = 0;
    } else {
        scale_start = 64 - (a0 & 63);
        scale_stop = a1 & 63;
    }
    bool clipped = seg.clipped;
    if (clipped) {
        // the clip along (major, minor)
        const int32_t lo = ymajor ? seg.ct : seg.cl, hi = ymajor ? seg.cb : seg.cr;
        const int32_t mlo = ymajor ? seg.cl : seg.ct, mhi = ymajor ? seg.cr : seg.cb;
        if (istart >= hi || istop <= lo) return;
        if (istart < lo) {
            fstart += slope * (lo - istart);
            istart = lo;
            scale_start = 64;
            if (istop - istart == 1) {
                scale_start = ((a1 - 1) & 63) + 1;  // contribution_64
                scale_stop = 0;
            }
        }
        if (istop > hi) {
            istop = hi;
            scale_stop = 0;
        }
        if (istart == istop) return;
        int32_t first, last;  // minor extent estimate
        if (slope >= 0) {
            first = fixed_floor(fstart - 0x8000);
            last = fixed_ceil(fstart + (istop - istart - 1) * slope + 0x8000);
        } else {
            last = fixed_ceil(fstart + 0x8000);
            first = fixed_floor(fstart + (istop - istart - 1) * slope - 0x8000);
        }
        if (first >= mhi || last <= mlo) return;
        if (mlo <= first && mhi >= last) clipped = false;
    }
    L.istart = istart;
    L.istop = istop;
    L.fstart = fstart;
    L.slope = slope;
    L.scale_start = scale_start;
    L.scale_stop = scale_stop;
    L.flags = (ymajor ? kHairYMajor : 0u) | (clipped ? kHairClipped : 0u);
    if (clipped) { L.cl = seg.cl; L.ct = seg.ct; L.cr = seg.cr; L.cb = seg.cb; }
    out.push_back(L);
}

}  // namespace

void hairline_lines(const rvk_scene &sc, const rvk_path &path, Ts ts, uint32_t w, uint32_t h,
                    std::vector<HairLine> &out) {
    rvks::PathBuilder dashed;
    const uint8_t *verbs;
    const float *pts;
    size_t nverbs, npts;
    if (!stroke_input(sc, path, rvks::resolution_scale(ts.sx, ts.ky, ts.kx, ts.sy), dashed, verbs, nverbs, pts,
                      npts))
        return;
    std::vector<rvks::Pt> dev(npts);
    for (size_t i = 0; i < npts; i++) {
        float x = pts[2 * i], y = pts[2 * i + 1];
        ts_map(ts, x, y);  // Path::transform
        dev[i] = rvks::Pt{x, y};
    }
    std::vector<rvks::HairSeg> segs;
    rvks::anti_hairline_segments(verbs, nverbs, dev.data(), npts, path.stroke_cap, w, h, segs);
    for (const rvks::HairSeg &g : segs) prepare_anti_hairline(g.x0, g.y0, g.x1, g.y1, g, out);
}

// ---------------------------------------------------------------------------
// Shader setup (tiny-skia LinearGradient / RadialGradient / Gradient::new and the
// gradient stage's factor/bias tables)
// ---------------------------------------------------------------------------
namespace {

bool invert(Ts t, Ts &o) {
    if (ts_is_identity(t)) { o = t; return true; }
    if (!ts_has_skew(t)) {
        if (ts_has_scale(t)) {
            float ix = 1.0f / t.sx, iy = 1.0f / t.sy;
            o = Ts{ix, 0.0f, 0.0f, iy, -t.tx * ix, -t.ty * iy};
        } else {
            o = Ts{1.0f, 0.0f, 0.0f, 1.0f, -t.tx, -t.ty};
        }
        return std::isfinite(o.sx) && std::isfinite(o.sy) && std::isfinite(o.tx) && std::isfinite(o.ty);
    }
    double det = (double)t.sx * (double)t.sy - (double)t.kx * (double)t.ky;
    const float nz = 1.0f / 4096.0f;
    if (std::fabs((float)det) <= nz * nz * nz) return false;
    double inv = 1.0 / det;
    o.sx = (float)((double)t.sy * inv);
    o.ky = (float)((double)-t.ky * inv);
    o.kx = (float)((double)-t.kx * inv);
    o.sy = (float)((double)t.sx * inv);
    o.tx = (float)(((double)t.kx * (double)t.ty - (double)t.sy * (double)t.tx) * inv);
    o.ty = (float)(((double)t.ky * (double)t.tx - (double)t.sx * (double)t.ty) * inv);
    return std::isfinite(o.sx) && std::isfinite(o.ky) && std::isfinite(o.kx) && std::isfinite(o.sy) &&
           std::isfinite(o.tx) && std::isfinite(o.ty);
}

// Shader::SolidColor: premultiplied lowp colour; opaque when alpha is exactly 1.
void set_solid(ShaderParams &s, const float c[4]) {
    float r = c[0], g = c[1], b = c[2], a = c[3];
    if (a != 1.0f) { r *= a; g *= a; b *= a; }
    uint32_t R = (uint32_t)(r * 255.0f + 0.5f), G = (uint32_t)(g * 255.0f + 0.5f), B = (uint32_t)(b * 255.0f + 0.5f),
             A = (uint32_t)(a * 255.0f + 0.5f);
    s.kind = 1;
    s.rgba = R | (G << 8) | (B << 16) | (A << 24);
    s.as_source = a == 1.0f;
}

}  // namespace

bool pattern_tile_size(const rvk_paint &p, Ts ts, float &sx, float &sy, uint32_t &w, uint32_t &h) {
    ts_get_scale(ts_concat(ts, ts_from(p.transform)), sx, sy);
    auto to_u32 = [](float v) -> uint32_t {  // `as u32`: saturating, NaN to 0
        if (!(v > 0.0f)) return 0u;
        return v >= 4294967296.0f ? UINT32_MAX : (uint32_t)v;
    };
    w = to_u32(std::round(p.rect.w * sx));
    h = to_u32(std::round(p.rect.h * sy));
    if (w == 0 || h == 0) return false;  // IntSize::from_wh
    // Pixmap::new: the row bytes and the whole buffer stay within i32
    return (uint64_t)w * 4 <= (uint64_t)INT32_MAX && (uint64_t)w * h * 4 <= (1ull << 31);
}

bool make_pattern_shader(const rvk_paint &p, Ts ts, float sx, float sy, const void *tile, uint32_t w, uint32_t h,
                         ShaderParams &s, float extra_opacity) {
    s = ShaderParams{};
    // pattern.transform * translate(rect.xy) * scale(1 / sx, 1 / sy), post-concatenated
    // with the path transform (Shader::transform), inverted by push_stages
    const Ts pts = ts_concat(ts_concat(ts_from(p.transform), Ts{1.0f, 0.0f, 0.0f, 1.0f, p.rect.x, p.rect.y}),
                             Ts{1.0f / sx, 0.0f, 0.0f, 1.0f / sy, 0.0f, 0.0f});
    Ts inv;
    if (!invert(ts_concat(ts, pts), inv)) return false;
    s.kind = 3;
    s.f32_pipeline = 1;
    s.sx = inv.sx; s.ky = inv.ky; s.kx = inv.kx; s.sy = inv.sy; s.tx = inv.tx; s.ty = inv.ty;
    // FilterQuality::Bicubic, lowered to Nearest when the transform is a translation
    s.quality = (inv.sx == 1.0f && inv.sy == 1.0f && inv.kx == 0.0f && inv.ky == 0.0f) ? 0u : 2u;
    s.spread = RVK_SPREAD_REPEAT;
    s.tile = (uint64_t)(uintptr_t)tile;
    s.tile_w = w;
    s.tile_h = h;
    s.tile_iw = 1.0f / (float)w;
    s.tile_ih = 1.0f / (float)h;
    // NormalizedF32::new_clamped(opacity), then Shader::apply_opacity
    const float o = std::min(std::max(p.opacity, 0.0f), 1.0f);
    s.pattern_opacity = std::min(std::max(o * std::min(std::max(extra_opacity, 0.0f), 1.0f), 0.0f), 1.0f);
    return true;
}

bool make_image_shader(const rvk_image &im, Ts ts, const void *tile, ShaderParams &s) {
    s = ShaderParams{};
    Ts inv;  // Pattern transform identity, post-concatenated with `ts`, inverted
    if (!invert(ts, inv)) return false;
    s.kind = 3;
    s.f32_pipeline = 1;
    s.sx = inv.sx; s.ky = inv.ky; s.kx = inv.kx; s.sy = inv.sy; s.tx = inv.tx; s.ty = inv.ty;
    const bool translate = inv.sx == 1.0f && inv.sy == 1.0f && inv.kx == 0.0f && inv.ky == 0.0f;
    s.quality = translate ? 0u : std::min(im.quality, 2u);
    s.spread = RVK_SPREAD_PAD;
    s.tile = (uint64_t)(uintptr_t)tile;
    s.tile_w = im.width;
    s.tile_h = im.height;
    s.tile_iw = 1.0f / (float)im.width;
    s.tile_ih = 1.0f / (float)im.height;
    s.pattern_opacity = 1.0f;
    return true;
}

// ts_from_poly_to_poly(c0, c1, (0, 0), (1, 0)) (tiny-skia radial_gradient.rs): the
// transform that maps c0 to the origin and c1 to (1, 0); false when c0 and c1 coincide.
static bool poly_to_unit_x(float x0, float y0, float x1, float y1, Ts &out) {
    // from_poly2(p0, p1) = [p1.y - p0.y, p0.x - p1.x, p1.x - p0.x, p1.y - p0.y, p0.x, p0.y]
    const Ts src{y1 - y0, x0 - x1, x1 - x0, y1 - y0, x0, y0};
    const Ts dst{0.0f, -1.0f, 1.0f, 0.0f, 0.0f, 0.0f};  // from_poly2((0, 0), (1, 0))
    Ts inv;
    if (!invert(src, inv)) return false;
    out = ts_concat(dst, inv);
    return true;
}

void seal_shader(ShaderHead &s) {
    s.mode = (s.kind & kModeKind) | (s.radial ? kModeRadial : 0u) | ((s.spread & 3u) << kModeSpreadShift) |
             (s.pad_clamp ? kModePadClamp : 0u) | (s.two_stop ? kModeTwoStop : 0u) | (s.opaque ? kModeOpaque : 0u) |
             (s.r_twist ? kModeRTwist : 0u) | (s.as_source ? kModeAsSource : 0u) | (s.clear ? kModeClear : 0u) |
             (s.f32_pipeline ? kModeF32 : 0u) | (s.nstops << kModeStopsShift);
}

bool make_shader(const rvk_scene &sc, const rvk_paint &p, Ts ts, ShaderParams &s, float extra_opacity) {
    s = ShaderParams{};
    // Shader::apply_opacity (Color::apply_opacity on the colour or on every stop)
    const float xo = std::min(std::max(extra_opacity, 0.0f), 1.0f);
    auto apply = [&](float &a) { a = std::min(std::max(a * xo, 0.0f), 1.0f); };
    if (p.kind == RVK_PAINT_COLOR) {
        float c[4] = {(float)p.color[0] / 255.0f, (float)p.color[1] / 255.0f, (float)p.color[2] / 255.0f,
                      (float)opacity_to_u8(p.opacity) / 255.0f};
        if (xo != 1.0f) apply(c[3]);
        set_solid(s, c);
        return true;
    }
    if (p.kind != RVK_PAINT_LINEAR && p.kind != RVK_PAINT_RADIAL) return false;
    struct GS { float pos; std::array<float, 4> c; };
    std::vector<GS> stops;
    const float fo = std::min(std::max(p.opacity, 0.0f), 1.0f);
    for (uint32_t i = 0; i < p.stop_count; i++) {
        const rvk_stop &st = sc.stops[p.stop_offset + i];
        float a = std::min(std::max(st.opacity, 0.0f), 1.0f) * fo;
        stops.push_back(GS{std::min(std::max(st.offset, 0.0f), 1.0f),
                           {(float)st.color[0] / 255.0f, (float)st.color[1] / 255.0f, (float)st.color[2] / 255.0f,
                            (float)opacity_to_u8(a) / 255.0f}});
    }
    if (stops.empty()) return false;
    if (stops.size() == 1) {
        if (xo != 1.0f) apply(stops[0].c[3]);
        set_solid(s, stops[0].c.data());
        return true;
    }
    if (stops.size() + 2 > (size_t)kMaxStops) return false;
    Ts unit;
    if (p.kind == RVK_PAINT_LINEAR) {
        float dx = p.x2 - p.x1, dy = p.y2 - p.y1, len = std::sqrt(dx * dx + dy * dy);
        if (!std::isfinite(len)) return false;
        if (len <= 1.0f / 32768.0f) {
            float c[4] = {0, 0, 0, 0};
            if (p.spread == RVK_SPREAD_PAD) {
                for (int k = 0; k < 4; k++) c[k] = stops.back().c[k];
            } else {
                for (size_t i = 0; i + 1 < stops.size(); i++) {
                    float wgt = stops[i + 1].pos - stops[i].pos;
                    for (int k = 0; k < 4; k++) c[k] += wgt * (stops[i].c[k] + stops[i + 1].c[k]) * 0.5f;
                }
            }
            if (xo != 1.0f) apply(c[3]);
            set_solid(s, c);
            return true;
        }
        float inv = 1.0f / len, vx = dx * inv, vy = dy * inv;
        float sn = -vy, cs = vx, cinv = 1.0f - cs;
        Ts rot{cs, sn, -sn, cs, sn * p.y1 + cinv * p.x1, -sn * p.x1 + cinv * p.y1};
        unit = ts_concat(Ts{inv, 0, 0, inv, 0, 0}, ts_concat(Ts{1, 0, 0, 1, -p.x1, -p.y1}, rot));
    } else if (float fdx = p.fx - p.cx, fdy = p.fy - p.cy, dc = std::sqrt(fdx * fdx + fdy * fdy);
               !(dc <= 1.0f / 4096.0f)) {
        // the focal point (fx, fy) off the centre: a two-point conical gradient from
        // (fx, fy), radius fr, to (cx, cy), radius r (tiny-skia RadialGradient::new,
        // SkTwoPointConicalGradient kFocal).  A focal radius with a moved focal point is
        // not restated here.
        if (!std::isfinite(dc) || p.fr != 0.0f || !(p.r > 0.0f)) return false;
        Ts unit0;
        if (!poly_to_unit_x(p.fx, p.fy, p.cx, p.cy, unit0)) return false;
        const float r1 = p.r / dc;
        if (std::fabs(1.0f - r1) <= 1.0f / 4096.0f) {
            s.focal = kFocalOnCircle;
            s.f32_pipeline = 1;
            unit = ts_concat(Ts{0.5f, 0, 0, 0.5f, 0, 0}, unit0);
        } else {
            s.focal = r1 > 1.0f ? kFocalWellBehaved : kFocalGreater;
            s.f32_pipeline = 1;
            unit = ts_concat(Ts{r1 / (r1 * r1 - 1.0f), 0, 0, 1.0f / std::sqrt(std::fabs(r1 * r1 - 1.0f)), 0, 0}, unit0);
        }
        s.p0 = 1.0f / r1;
        s.radial = 1;
    } else {
        float r0 = p.fr, r1 = p.r;
        // equal radii make the twist divide by zero; the NaN t clamps to the first stop
        if (!(r1 >= 0.0f) || !(std::max(r0, r1) > 0.0f)) return false;
        float scale = 1.0f / std::max(r0, r1);
        unit = ts_concat(Ts{scale, 0, 0, scale, 0, 0}, Ts{1, 0, 0, 1, -p.cx, -p.cy});
        s.radial = 1;
        if (r0 != 0.0f) {
            // focal radius: t = (t - r0/m) / ((r1 - r0)/m), m = max(r0, r1)
            const float m = std::max(r0, r1);
            s.r_twist = 1;
            s.r_scale = r0 / m;
            s.r_bias = (r1 - r0) / m;
        }
    }
    bool dummy_first = stops.front().pos != 0.0f, dummy_last = stops.back().pos != 1.0f;
    if (dummy_first) stops.insert(stops.begin(), GS{0.0f, stops.front().c});
    if (dummy_last) stops.push_back(GS{1.0f, stops.back().c});
    if (xo != 1.0f)
        for (auto &g : stops) apply(g.c[3]);
    bool opaque = true;
    for (auto &g : stops) opaque = opaque && g.c[3] == 1.0f;
    size_t start = dummy_first ? 0 : 1;
    float prev = 0.0f;
    bool uniform = true;
    const float step = stops[start].pos - prev;
    for (size_t i = start; i < stops.size(); i++) {
        float curr = i + 1 == stops.size() ? 1.0f : std::max(prev, std::min(stops[i].pos, 1.0f));
        uniform = uniform && std::fabs(step - (curr - prev)) <= 1.0f / 4096.0f;
        stops[i].pos = std::min(std::max(curr, 0.0f), 1.0f);
        prev = curr;
    }
    Ts inv;
    if (!invert(ts_concat(ts, ts_from(p.transform)), inv)) return false;
    Ts u = ts_concat(unit, inv);
    s.kind = 2;
    s.sx = u.sx; s.ky = u.ky; s.kx = u.kx; s.sy = u.sy; s.tx = u.tx; s.ty = u.ty;
    s.spread = p.spread;
    s.pad_clamp = p.spread == RVK_SPREAD_PAD && uniform;
    s.opaque = opaque;
    s.as_source = opaque && !s.radial;
    if (stops.size() == 2 && uniform) {
        s.two_stop = 1;
        for (int k = 0; k < 4; k++) {
            s.f[0][k] = stops[1].c[k] - stops[0].c[k];
            s.b[0][k] = stops[0].c[k];
        }
        return true;
    }
    size_t first = 0, last = stops.size() - 1;
    if (stops.size() > 2) {
        if (stops[0].c == stops[1].c) first = 1;
        if (stops[stops.size() - 2].c == stops[stops.size() - 1].c) last = stops.size() - 2;
    }
    uint32_t n = 0;
    float tl = stops[first].pos;
    std::array<float, 4> cl = stops[first].c;
    s.t[n] = 0.0f;
    for (int k = 0; k < 4; k++) { s.f[n][k] = 0.0f; s.b[n][k] = cl[k]; }
    n++;
    for (size_t i = first; i < last; i++) {
        float tr = stops[i + 1].pos;
        std::array<float, 4> cr = stops[i + 1].c;
        if (tl < tr) {
            s.t[n] = tl;
            for (int k = 0; k < 4; k++) {
                s.f[n][k] = (cr[k] - cl[k]) / (tr - tl);
                s.b[n][k] = cl[k] - s.f[n][k] * tl;
            }
            n++;
        }
        tl = tr;
        cl = cr;
    }
    s.t[n] = tl;
    for (int k = 0; k < 4; k++) { s.f[n][k] = 0.0f; s.b[n][k] = cl[k]; }
    n++;
    s.nstops = n;
    return true;
}

// ---------------------------------------------------------------------------
// Coverage and hairline jobs of one path, prepared off the render walk
// ---------------------------------------------------------------------------
void prep_job(const EdgeSet &es, uint32_t rule, bool aa, IRect clip, int32_t lx, int32_t ly, JobPrep &P) {
    if (!es.ok) return;
    const int shift = aa ? 2 : 0;
    const std::vector<EdgeRec> &er = es.edges;
    const std::vector<uint32_t> &caps = es.caps;
    // last covered row of each edge: exact for lines; for curves the row of the
    // end point plus one (the stepped points stay within a fraction of a row of
    // the monotonic curve; the expansion kernel checks the bound)
    std::vector<int32_t> ylast(er.size());
    int32_t ymin = INT32_MAX, ymax = INT32_MIN;
    for (size_t i = 0; i < er.size(); i++) {
        const EdgeRec &e = er[i];
        ylast[i] = e.type == 0 ? e.last_y : (((e.ly >> 10) + 32) >> 6) + 1;
        ymin = std::min(ymin, e.first_y);
        ymax = std::max(ymax, ylast[i]);
    }
    // (clamped in double: the control points may lie far outside the pixmap)
    int64_t x0 = (int64_t)std::max<double>(std::floor((double)es.mnx) - 1.0, (double)clip.x);
    int64_t x1 = (int64_t)std::min<double>(std::ceil((double)es.mxx) + 2.0, (double)clip.x + clip.w);
    int64_t y0 = std::max<int64_t>(ymin >> shift, clip.y);
    int64_t y1 = std::min<int64_t>((ymax >> shift) + 1, (int64_t)clip.y + clip.h);
    if (x1 <= x0 || y1 <= y0) return;
    CovJob &j = P.j;
    j = CovJob{};
    j.x0 = (int32_t)x0; j.y0 = (int32_t)y0; j.w = (int32_t)(x1 - x0); j.h = (int32_t)(y1 - y0);
    j.rule = rule;
    j.aa = aa ? 1 : 0;
    // the tile grid: kTile-aligned in target pixels (edge space + (lx, ly))
    auto floor_tile = [](int64_t v) { return v >= 0 ? v / kTile * kTile : -((-v + kTile - 1) / kTile) * kTile; };
    const int64_t gx0 = floor_tile(x0 + lx) - lx, gy0 = floor_tile(y0 + ly) - ly;
    const int64_t tiles_x = (x1 - gx0 + kTile - 1) / kTile, tiles_y = (y1 - gy0 + kTile - 1) / kTile;
    if (tiles_x > (int64_t)kMaxTilesX) {
        P.too_wide = true;
        return;
    }
    j.gx0 = (int32_t)gx0; j.gy0 = (int32_t)gy0;
    j.tiles_x = (uint32_t)tiles_x; j.tiles_y = (uint32_t)tiles_y;
    // edges sorted by first row, stably (ids keep tiny-skia's edge-list order): a
    // counting sort over the rows when they are few against the edges, else a sort
    // of (row, index) keys
    const size_t n = er.size();
    std::vector<uint32_t> order(n);
    int32_t fmin = INT32_MAX, fmax = INT32_MIN;
    for (const EdgeRec &e : er) { fmin = std::min(fmin, e.first_y); fmax = std::max(fmax, e.first_y); }
    const uint64_t range = (uint64_t)((int64_t)fmax - fmin) + 1;
    if (range <= 8 * (uint64_t)n + 256) {
        std::vector<uint32_t> start((size_t)range + 1, 0);
        for (const EdgeRec &e : er) start[(size_t)(e.first_y - fmin) + 1]++;
        for (size_t r = 1; r <= range; r++) start[r] += start[r - 1];
        for (uint32_t i = 0; i < n; i++) order[start[(size_t)(er[i].first_y - fmin)]++] = i;
    } else {
        std::vector<uint64_t> keys(n);
        for (uint32_t i = 0; i < n; i++) keys[i] = ((uint64_t)(uint32_t)(er[i].first_y - fmin) << 32) | i;
        std::sort(keys.begin(), keys.end());
        for (uint32_t i = 0; i < n; i++) order[i] = (uint32_t)keys[i];
    }
    j.edge_count = (uint32_t)er.size();
    P.edges.resize(er.size());
    P.ecov.resize(er.size());
    std::vector<int32_t> prefmax(er.size());
    int32_t pm = INT32_MIN;
    uint64_t slots = 0;
    for (size_t k = 0; k < order.size(); k++) {
        const uint32_t i = order[k];
        EdgeRec e = er[i];
        e.piece_offset = (uint32_t)slots;
        slots += caps[i];
        P.edges[k] = e;
        P.ecov[k] = EdgeCov{e.first_y, ylast[i], e.winding, (int32_t)i, e.piece_offset, 0};
        pm = std::max(pm, ylast[i]);
        prefmax[k] = pm;
    }
    P.pieces = slots;
    // per band: the edges that start at or before the band's last covered row, minus
    // the leading ones that all end before its first covered row
    const int32_t bands = (int32_t)j.tiles_y;
    P.windows.resize((size_t)bands);
    uint32_t lo = 0, hi = 0;
    for (int32_t g = 0; g < bands; g++) {
        const int32_t r0 = std::max(j.y0, j.gy0 + g * kBandRows), r1 = std::min(j.y0 + j.h, j.gy0 + (g + 1) * kBandRows);
        const int32_t s0 = r0 << shift, s1 = (r1 << shift) - 1;
        while (lo < j.edge_count && prefmax[lo] < s0) lo++;
        while (hi < j.edge_count && P.ecov[hi].first_y <= s1) hi++;
        P.windows[(size_t)g] = uint2{lo, std::max(lo, hi)};
    }
    // edges whose row range meets each band (difference array): an edge crosses a
    // sub-scanline at most once, so this bounds every sub-scanline's crossings and
    // picks the coverage kernel variant that holds them
    std::vector<int32_t> diff((size_t)bands + 1, 0);
    for (size_t i = 0; i < er.size(); i++) {
        const int64_t fr = (int64_t)(er[i].first_y >> shift), lr = (int64_t)(ylast[i] >> shift);
        const int64_t a = std::max<int64_t>(fr, j.y0), b = std::min<int64_t>(lr, (int64_t)j.y0 + j.h - 1);
        if (b < a) continue;
        diff[(size_t)((a - j.gy0) / kBandRows)] += 1;
        diff[(size_t)((b - j.gy0) / kBandRows) + 1] -= 1;
    }
    P.tier.resize((size_t)bands);
    for (int32_t g = 0, run = 0; g < bands; g++) {
        run += diff[g];
        const uint32_t n = (uint32_t)run;
        const uint8_t t = n <= kCoverageKeys[0] ? 0 : n <= kCoverageKeys[1] ? 1 : 2;
        P.tier[(size_t)g] = t;
        P.tier_count[t]++;
    }
    P.ok = true;
}

void prep_hair(std::vector<HairLine> &&lines, IRect clip, HairPrep &P) {
    if (lines.empty()) return;
    struct Box { int32_t x0, y0, x1, y1; };
    std::vector<Box> boxes(lines.size());
    int32_t jx0 = INT32_MAX, jy0 = INT32_MAX, jx1 = INT32_MIN, jy1 = INT32_MIN;
    for (size_t i = 0; i < lines.size(); i++) {
        const HairLine &L = lines[i];
        // minor pixels touched: f >> 16 and the one before, f at both ends
        const int64_t fa = (int64_t)L.fstart + 0x8000;
        const int64_t fb = fa + (int64_t)(L.istop - L.istart - 1) * L.slope;
        const int32_t m0 = (int32_t)(std::min(fa, fb) >> 16) - 1, m1 = (int32_t)(std::max(fa, fb) >> 16) + 1;
        Box b = (L.flags & kHairYMajor) ? Box{m0, L.istart, m1, L.istop} : Box{L.istart, m0, L.istop, m1};
        if (L.flags & kHairClipped) {
            b.x0 = std::max(b.x0, L.cl); b.y0 = std::max(b.y0, L.ct);
            b.x1 = std::min(b.x1, L.cr); b.y1 = std::min(b.y1, L.cb);
        }
        b.x0 = std::max(b.x0, clip.x); b.y0 = std::max(b.y0, clip.y);
        b.x1 = std::min(b.x1, clip.x + clip.w); b.y1 = std::min(b.y1, clip.y + clip.h);
        boxes[i] = b;
        if (b.x0 >= b.x1 || b.y0 >= b.y1) continue;
        jx0 = std::min(jx0, b.x0); jy0 = std::min(jy0, b.y0);
        jx1 = std::max(jx1, b.x1); jy1 = std::max(jy1, b.y1);
    }
    if (jx0 >= jx1 || jy0 >= jy1) return;
    HairJob &j = P.j;
    j = HairJob{};
    j.x0 = jx0; j.y0 = jy0; j.w = jx1 - jx0; j.h = jy1 - jy0;
    j.tiles_x = (uint32_t)((j.w + kHairTile - 1) / kHairTile);
    j.tiles_y = (uint32_t)((j.h + kHairTile - 1) / kHairTile);
    const size_t nt = (size_t)j.tiles_x * j.tiles_y;
    std::vector<uint32_t> count(nt + 1, 0);
    auto tiles_of = [&](const Box &b, auto &&fn) {
        const int32_t tx0 = (b.x0 - j.x0) / kHairTile, tx1 = (b.x1 - 1 - j.x0) / kHairTile;
        const int32_t ty0 = (b.y0 - j.y0) / kHairTile, ty1 = (b.y1 - 1 - j.y0) / kHairTile;
        for (int32_t ty = ty0; ty <= ty1; ty++)
            for (int32_t tx = tx0; tx <= tx1; tx++) fn((size_t)ty * j.tiles_x + tx);
    };
    for (const Box &b : boxes)
        if (b.x0 < b.x1 && b.y0 < b.y1) tiles_of(b, [&](size_t t) { count[t]++; });
    P.tiles.resize(nt + 1);
    uint32_t run = 0;
    for (size_t t = 0; t < nt; t++) {
        P.tiles[t] = run;
        const uint32_t c = count[t];
        count[t] = run;  // becomes the fill cursor
        run += c;
    }
    P.tiles[nt] = run;
    P.bins.resize(run);
    for (size_t i = 0; i < lines.size(); i++) {
        const Box &b = boxes[i];
        if (b.x0 < b.x1 && b.y0 < b.y1) tiles_of(b, [&](size_t t) { P.bins[count[t]++] = (uint32_t)i; });
    }
    P.lines = std::move(lines);
    P.ok = true;
}

}  // namespace rvkr
