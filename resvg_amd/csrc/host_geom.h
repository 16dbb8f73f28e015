// Host-side geometry of the render walk and filter region math: tiny-skia-path's
// Transform / Rect / IntRect arithmetic in f32 as resvg uses it
// (crates/resvg/src/render.rs:49-143, filter/mod.rs:357-521, 1063-1147).
#pragma once
#include <functional>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include "../../include/resvg_amd.h"
#include "rvk_internal.h"

namespace rvkh {

struct IRect { int32_t x, y, w, h; };
static inline int32_t ir_right(IRect r) { return r.x + r.w; }
static inline int32_t ir_bottom(IRect r) { return r.y + r.h; }
static inline bool ir_eq(IRect a, IRect b) { return a.x == b.x && a.y == b.y && a.w == b.w && a.h == b.h; }

static inline uint32_t f32_bits(float v) { uint32_t u; std::memcpy(&u, &v, 4); return u; }

// strict-num ApproxEqUlps / usvg ApproxZeroUlps (crates/usvg/src/tree/geom.rs:9-24)
static inline bool approx_eq_ulps(float a, float b, int32_t ulps) {
    if (a == b) return true;
    if (std::signbit(a) != std::signbit(b)) return false;
    int32_t d = (int32_t)(f32_bits(a) - f32_bits(b));
    return d >= -ulps && d <= ulps;
}
static inline bool approx_zero_ulps(float a, int32_t ulps) { return approx_eq_ulps(a, 0.0f, ulps); }

static inline int32_t as_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483647.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}
static inline uint32_t as_u32(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 4294967295.0f) return UINT32_MAX;
    return (uint32_t)v;
}
static inline uint8_t as_u8(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}

// usvg::Opacity::to_u8 (strict-num NormalizedF32): (v * 255 + 0.5) as u8 on [0, 1]
static inline uint8_t opacity_to_u8(float v) {
    v = std::min(std::max(v, 0.0f), 1.0f);
    return as_u8(v * 255.0f + 0.5f);
}

// Color::from_rgba8(..).premultiply().to_color_u8() (tiny-skia color.rs)
static inline void premultiply_color(const uint8_t rgba[4], uint8_t out[4]) {
    float r = (float)rgba[0] / 255.0f, g = (float)rgba[1] / 255.0f, b = (float)rgba[2] / 255.0f,
          a = (float)rgba[3] / 255.0f;
    if (a != 1.0f) { r *= a; g *= a; b *= a; }
    out[0] = as_u8(r * 255.0f + 0.5f);
    out[1] = as_u8(g * 255.0f + 0.5f);
    out[2] = as_u8(b * 255.0f + 0.5f);
    out[3] = as_u8(a * 255.0f + 0.5f);
}

struct Ts { float sx, ky, kx, sy, tx, ty; };
static inline Ts ts_from(rvk_transform t) { return Ts{t.a, t.b, t.c, t.d, t.e, t.f}; }
static inline Ts ts_identity() { return Ts{1, 0, 0, 1, 0, 0}; }
static inline bool ts_is_identity(Ts t) {
    return t.sx == 1 && t.ky == 0 && t.kx == 0 && t.sy == 1 && t.tx == 0 && t.ty == 0;
}
static inline bool ts_has_skew(Ts t) { return t.kx != 0 || t.ky != 0; }
static inline bool ts_has_scale(Ts t) { return t.sx != 1 || t.sy != 1; }
static inline bool ts_has_translate(Ts t) { return t.tx != 0 || t.ty != 0; }

// a.pre_concat(b) == concat(a, b): maps by b first, then a
static inline Ts ts_concat(Ts a, Ts b) {
    if (ts_is_identity(a)) return b;
    if (ts_is_identity(b)) return a;
    if (!ts_has_skew(a) && !ts_has_skew(b))
        return Ts{a.sx * b.sx, 0.0f, 0.0f, a.sy * b.sy, a.sx * b.tx + a.tx, a.sy * b.ty + a.ty};
    auto mam = [](float p, float q, float r, float s) { return (float)((double)p * (double)q + (double)r * (double)s); };
    return Ts{mam(a.sx, b.sx, a.kx, b.ky), mam(a.ky, b.sx, a.sy, b.ky), mam(a.sx, b.kx, a.kx, b.sy),
              mam(a.ky, b.kx, a.sy, b.sy), mam(a.sx, b.tx, a.kx, b.ty) + a.tx, mam(a.ky, b.tx, a.sy, b.ty) + a.ty};
}

static inline void ts_map(Ts t, float &x, float &y) {
    if (ts_is_identity(t)) return;
    if (!ts_has_skew(t)) {
        if (!ts_has_scale(t)) { x = x + t.tx; y = y + t.ty; return; }
        x = x * t.sx + t.tx;
        y = y * t.sy + t.ty;
        return;
    }
    float nx = x * t.sx + y * t.kx + t.tx;
    float ny = x * t.ky + y * t.sy + t.ty;
    x = nx; y = ny;
}

static inline void ts_get_scale(Ts t, float &sx, float &sy) {
    sx = std::sqrt(t.sx * t.sx + t.kx * t.kx);
    sy = std::sqrt(t.ky * t.ky + t.sy * t.sy);
}

struct FRect { float l, t, r, b; bool ok; };

// NonZeroRect::from_xywh(..).transform(ts): bounds of the four mapped corners
static inline FRect rect_transform(rvk_rect rc, Ts ts) {
    float l = rc.x, t = rc.y, r = rc.x + rc.w, b = rc.y + rc.h;
    bool ok = std::isfinite(l) && std::isfinite(t) && std::isfinite(r) && std::isfinite(b) && r > l && b > t;
    if (!ok || ts_is_identity(ts)) return FRect{l, t, r, b, ok};
    float xs[4] = {l, r, r, l}, ys[4] = {t, t, b, b};
    for (int i = 0; i < 4; i++) ts_map(ts, xs[i], ys[i]);
    float mnx = xs[0], mxx = xs[0], mny = ys[0], mxy = ys[0];
    for (int i = 1; i < 4; i++) {
        mnx = std::min(mnx, xs[i]); mxx = std::max(mxx, xs[i]);
        mny = std::min(mny, ys[i]); mxy = std::max(mxy, ys[i]);
    }
    ok = std::isfinite(mnx) && std::isfinite(mxx) && std::isfinite(mny) && std::isfinite(mxy) && mxx > mnx &&
         mxy > mny;
    return FRect{mnx, mny, mxx, mxy, ok};
}

// NonZeroRect::to_int_rect
static inline IRect frect_to_int(FRect r) {
    int32_t w = (int32_t)std::max<uint32_t>(1u, as_u32(std::ceil(r.r - r.l)));
    int32_t h = (int32_t)std::max<uint32_t>(1u, as_u32(std::ceil(r.b - r.t)));
    return IRect{as_i32(std::floor(r.l)), as_i32(std::floor(r.t)), w, h};
}

// geom::fit_to_rect (crates/resvg/src/geom.rs:5-30)
static inline bool fit_to_rect(IRect r, IRect bounds, IRect &out) {
    int64_t l = std::max<int64_t>(r.x, bounds.x), t = std::max<int64_t>(r.y, bounds.y);
    int64_t rr = std::min<int64_t>((int64_t)r.x + r.w, (int64_t)bounds.x + bounds.w);
    int64_t bb = std::min<int64_t>((int64_t)r.y + r.h, (int64_t)bounds.y + bounds.h);
    if (rr <= l || bb <= t) return false;
    out = IRect{(int32_t)l, (int32_t)t, (int32_t)(rr - l), (int32_t)(bb - t)};
    return true;
}

// scale_coordinates (filter/mod.rs:1144-1147)
static inline void scale_coords(float x, float y, Ts ts, float &ox, float &oy) {
    float sx, sy;
    ts_get_scale(ts, sx, sy);
    ox = x * sx;
    oy = y * sy;
}

// resolve_std_dev (filter/mod.rs:1119-1142)
static inline bool resolve_std_dev(float sdx, float sdy, Ts ts, double &ox, double &oy, bool &box) {
    float x, y;
    scale_coords(sdx, sdy, ts, x, y);
    if (approx_eq_ulps(x, 0.0f, 4) && approx_eq_ulps(y, 0.0f, 4)) return false;
    if (x < 0.05f) x = 0.0f;
    if (y < 0.05f) y = 0.0f;
    box = x >= 2.0f || y >= 2.0f;
    ox = x;
    oy = y;
    return true;
}

// transform_light_source (filter/mod.rs:1063-1098)
static inline rvk_light_source transform_light(rvk_light_source l, IRect region, Ts ts) {
    const float SQRT_2 = 1.41421353816986083984375f;
    if (l.kind == RVK_LIGHT_POINT) {
        float x = l.x, y = l.y;
        ts_map(ts, x, y);
        l.x = x - (float)region.x;
        l.y = y - (float)region.y;
        l.z = l.z * std::sqrt(ts.sx * ts.sx + ts.sy * ts.sy) / SQRT_2;
    } else if (l.kind == RVK_LIGHT_SPOT) {
        float sz = std::sqrt(ts.sx * ts.sx + ts.sy * ts.sy) / SQRT_2;
        float x = l.x, y = l.y;
        ts_map(ts, x, y);
        l.x = x - (float)region.x;
        l.y = y - (float)region.y;
        l.z *= sz;
        x = l.points_at_x; y = l.points_at_y;
        ts_map(ts, x, y);
        l.points_at_x = x - (float)region.x;
        l.points_at_y = y - (float)region.y;
        l.points_at_z *= sz;
    }
    return l;
}

// drop-shadow flood of every pixel with the shadow colour at the pixel's alpha
// (filter/mod.rs:606-617); defined in pointwise.hip
rvk_status launch_shadow_flood(rvk_context *ctx, uint8_t *px, size_t n, const uint8_t color[4], float opacity);

}  // namespace rvkh

namespace rvk {
// Draws the children of scene group `root` with transform `ts` into the w x h device
// pixmap `px` (feImage content, filter/mod.rs:861-889); supplied by the render walk.
using FeImageHook = std::function<rvk_status(uint32_t root, rvkh::Ts ts, uint8_t *px, uint32_t w, uint32_t h)>;
rvk_status filter_apply(rvk_context *ctx, const rvk_filter *filter, const rvk_filter_data *data, rvkh::Ts ts,
                        uint8_t *px, uint32_t w, uint32_t h, rvk_status *out_error,
                        const FeImageHook *fe_image = nullptr);
}
