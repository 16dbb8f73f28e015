// Oracle geometry helpers: tiny-skia-path Transform / Rect arithmetic as used by
// resvg's render walk and filter region math.  Test infrastructure only.
#pragma once
#include "common.h"
#include "../include/resvg_amd.h"

namespace orc {

struct Ts { float sx, ky, kx, sy, tx, ty; };
static inline Ts ts_from(rvk_transform t) { return Ts{t.a, t.b, t.c, t.d, t.e, t.f}; }
static inline Ts ts_identity() { return Ts{1, 0, 0, 1, 0, 0}; }
static inline bool ts_is_identity(Ts t) {
    return t.sx == 1 && t.ky == 0 && t.kx == 0 && t.sy == 1 && t.tx == 0 && t.ty == 0;
}
static inline bool ts_has_skew(Ts t) { return t.kx != 0 || t.ky != 0; }
static inline bool ts_has_scale(Ts t) { return t.sx != 1 || t.sy != 1; }

// Transform::pre_concat / post_concat (a.pre_concat(b) = a * b).
static inline Ts ts_concat(Ts a, Ts b) {
    if (ts_is_identity(a)) return b;
    if (ts_is_identity(b)) return a;
    if (!ts_has_skew(a) && !ts_has_skew(b)) {
        return Ts{a.sx * b.sx, 0.0f, 0.0f, a.sy * b.sy, a.sx * b.tx + a.tx, a.sy * b.ty + a.ty};
    }
    auto mam = [](float p, float q, float r, float s) {
        return (float)((double)p * (double)q + (double)r * (double)s);
    };
    return Ts{mam(a.sx, b.sx, a.kx, b.ky), mam(a.ky, b.sx, a.sy, b.ky),
              mam(a.sx, b.kx, a.kx, b.sy), mam(a.ky, b.kx, a.sy, b.sy),
              mam(a.sx, b.tx, a.kx, b.ty) + a.tx, mam(a.ky, b.tx, a.sy, b.ty) + a.ty};
}

// Transform::map_points
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

// Transform::get_scale
static inline void ts_get_scale(Ts t, float &sx, float &sy) {
    sx = std::sqrt(t.sx * t.sx + t.kx * t.kx);
    sy = std::sqrt(t.ky * t.ky + t.sy * t.sy);
}

struct FRect { float l, t, r, b; bool ok; };

// NonZeroRect::from_xywh + NonZeroRect::transform (bounds of the 4 mapped corners).
static inline FRect rect_transform(rvk_rect rc, Ts ts) {
    float l = rc.x, t = rc.y, r = rc.x + rc.w, b = rc.y + rc.h;
    FRect out{l, t, r, b, std::isfinite(l) && std::isfinite(t) && std::isfinite(r) && std::isfinite(b) &&
                              r > l && b > t};
    if (!out.ok || ts_is_identity(ts)) return out;
    float xs[4] = {l, r, r, l}, ys[4] = {t, t, b, b};
    for (int i = 0; i < 4; i++) ts_map(ts, xs[i], ys[i]);
    float mnx = xs[0], mxx = xs[0], mny = ys[0], mxy = ys[0];
    for (int i = 1; i < 4; i++) {
        mnx = std::min(mnx, xs[i]); mxx = std::max(mxx, xs[i]);
        mny = std::min(mny, ys[i]); mxy = std::max(mxy, ys[i]);
    }
    out = FRect{mnx, mny, mxx, mxy, std::isfinite(mnx) && std::isfinite(mxx) && std::isfinite(mny) &&
                                        std::isfinite(mxy) && mxx > mnx && mxy > mny};
    return out;
}

// NonZeroRect::to_int_rect: floor(x), floor(y), max(1, ceil(width)), max(1, ceil(height)).
static inline IRect frect_to_int(FRect r) {
    int32_t w = (int32_t)std::max<uint32_t>(1u, as_u32(std::ceil(r.r - r.l)));
    int32_t h = (int32_t)std::max<uint32_t>(1u, as_u32(std::ceil(r.b - r.t)));
    return IRect{as_i32(std::floor(r.l)), as_i32(std::floor(r.t)), w, h};
}

// geom::fit_to_rect (crates/resvg/src/geom.rs:5-30); false when empty.
static inline bool fit_to_rect(IRect r, IRect bounds, IRect &out) {
    int64_t l = std::max<int64_t>(r.x, bounds.x), t = std::max<int64_t>(r.y, bounds.y);
    int64_t rr = std::min<int64_t>((int64_t)r.x + r.w, (int64_t)bounds.x + bounds.w);
    int64_t bb = std::min<int64_t>((int64_t)r.y + r.h, (int64_t)bounds.y + bounds.h);
    if (rr <= l || bb <= t) return false;
    out = IRect{(int32_t)l, (int32_t)t, (int32_t)(rr - l), (int32_t)(bb - t)};
    return true;
}

}  // namespace orc
