// tiny-skia's f32 ("highp") pixel stages on device: u8 load as v*(1/255), the 29
// blend modes, and store as round-to-nearest-even of clamp(v)*255.  Used by the
// draw_pixmap kernel (layer composites, filter blends) and by the fused path
// compositor.  Operation order matches the oracle's restatement.
#pragma once
#include "device_math.h"

namespace rvkd {

struct F4 { float r, g, b, a; };

__device__ __forceinline__ float from_u8f(uint32_t v) { return (float)v * (1.0f / 255.0f); }
__device__ __forceinline__ F4 load_px(uint32_t p) {
    return F4{from_u8f(ch(p, 0)), from_u8f(ch(p, 1)), from_u8f(ch(p, 2)), from_u8f(ch(p, 3))};
}
__device__ __forceinline__ uint32_t unnorm(float v) {
    v = fmaxf(v, 0.0f);  // NaN -> 0
    v = fminf(v, 1.0f);
    return (uint32_t)rintf(v * 255.0f);
}
__device__ __forceinline__ uint32_t store_px(F4 o) { return pack(unnorm(o.r), unnorm(o.g), unnorm(o.b), unnorm(o.a)); }

__device__ __forceinline__ float inv(float v) { return 1.0f - v; }
__device__ __forceinline__ float two(float v) { return v + v; }
__device__ __forceinline__ float lum(float r, float g, float b) { return r * 0.30f + (g * 0.59f + b * 0.11f); }
__device__ __forceinline__ float sat3(float r, float g, float b) {
    return fmaxf(r, fmaxf(g, b)) - fminf(r, fminf(g, b));
}
__device__ __forceinline__ void set_sat(float &r, float &g, float &b, float s) {
    float mn = fminf(r, fminf(g, b)), mx = fmaxf(r, fmaxf(g, b)), st = mx - mn;
    r = st == 0.0f ? 0.0f : (r - mn) * s / st;
    g = st == 0.0f ? 0.0f : (g - mn) * s / st;
    b = st == 0.0f ? 0.0f : (b - mn) * s / st;
}
__device__ __forceinline__ void set_lum(float &r, float &g, float &b, float l) {
    float d = l - lum(r, g, b);
    r += d; g += d; b += d;
}
// clip_color as tiny-skia evaluates it: only the upper clamp toward the luminance
// (the reference's mix-blend-mode/hue fixture rules out a lower `min < 0` clamp)
__device__ __forceinline__ float clip1(float c, float mx, float l, float a) {
    if (mx > a && mx - l != 0.0f) c = l + (c - l) * (a - l) / (mx - l);
    return fmaxf(c, 0.0f);
}
__device__ __forceinline__ void clip_color(float &r, float &g, float &b, float a) {
    float mx = fmaxf(r, fmaxf(g, b)), l = lum(r, g, b);
    r = clip1(r, mx, l, a); g = clip1(g, mx, l, a); b = clip1(b, mx, l, a);
}

__device__ __forceinline__ float sep_channel(uint32_t mode, float x, float y, float sa, float da) {
    switch (mode) {
        case RVK_BLEND_DARKEN: return x + y - fmaxf(x * da, y * sa);
        case RVK_BLEND_LIGHTEN: return x + y - fminf(x * da, y * sa);
        case RVK_BLEND_DIFFERENCE: return x + y - two(fminf(x * da, y * sa));
        case RVK_BLEND_EXCLUSION: return x + y - two(x * y);
        case RVK_BLEND_COLOR_BURN:
            if (y == da) return y + x * inv(da);
            if (x == 0.0f) return y * inv(sa);
            return sa * (da - fminf(da, (da - y) * sa * (1.0f / x))) + x * inv(da) + y * inv(sa);
        case RVK_BLEND_COLOR_DODGE:
            if (y == 0.0f) return x * inv(da);
            if (x == sa) return x + y * inv(sa);
            return sa * fminf(da, (y * sa) * (1.0f / (sa - x))) + x * inv(da) + y * inv(sa);
        case RVK_BLEND_HARD_LIGHT:
            return x * inv(da) + y * inv(sa) + (two(x) <= sa ? two(x * y) : sa * da - two((da - y) * (sa - x)));
        case RVK_BLEND_OVERLAY:
            return x * inv(da) + y * inv(sa) + (two(y) <= da ? two(x * y) : sa * da - two((da - y) * (sa - x)));
        case RVK_BLEND_SOFT_LIGHT: {
            float m = da > 0.0f ? y / da : 0.0f, s2 = two(x), m4 = two(two(m));
            float dark_src = y * (sa + (s2 - sa) * (1.0f - m));
            float dark_dst = (m4 * m4 + m4) * (m - 1.0f) + 7.0f * m;
            float lite_dst = sqrtf(m) - m;
            float lite_src = y * sa + da * (s2 - sa) * (two(two(y)) <= da ? dark_dst : lite_dst);
            return x * inv(da) + y * inv(sa) + (s2 <= sa ? dark_src : lite_src);
        }
        default: return x;
    }
}

__device__ __forceinline__ float pd_channel(uint32_t mode, float x, float y, float sa, float da) {
    switch (mode) {
        case RVK_BLEND_CLEAR: return 0.0f;
        case RVK_BLEND_SOURCE: return x;
        case RVK_BLEND_DESTINATION: return y;
        case RVK_BLEND_SOURCE_OVER: return y * inv(sa) + x;
        case RVK_BLEND_DESTINATION_OVER: return x * inv(da) + y;
        case RVK_BLEND_SOURCE_IN: return x * da;
        case RVK_BLEND_DESTINATION_IN: return y * sa;
        case RVK_BLEND_SOURCE_OUT: return x * inv(da);
        case RVK_BLEND_DESTINATION_OUT: return y * inv(sa);
        case RVK_BLEND_SOURCE_ATOP: return x * da + y * inv(sa);
        case RVK_BLEND_DESTINATION_ATOP: return y * sa + x * inv(da);
        case RVK_BLEND_XOR: return x * inv(da) + y * inv(sa);
        case RVK_BLEND_PLUS: return fminf(x + y, 1.0f);
        case RVK_BLEND_MODULATE: return x * y;
        case RVK_BLEND_SCREEN: return x + y - x * y;
        case RVK_BLEND_MULTIPLY: return x * inv(da) + y * inv(sa) + x * y;
        default: return x;
    }
}

__device__ __forceinline__ F4 blend(uint32_t mode, F4 s, F4 d) {
    const float sa = s.a, da = d.a;
    // SourceOver (nearly every layer): decided once rather than in each channel's switch
    if (mode == RVK_BLEND_SOURCE_OVER)
        return F4{d.r * inv(sa) + s.r, d.g * inv(sa) + s.g, d.b * inv(sa) + s.b, d.a * inv(sa) + s.a};
    if (mode <= RVK_BLEND_SCREEN || mode == RVK_BLEND_MULTIPLY) {
        return F4{pd_channel(mode, s.r, d.r, sa, da), pd_channel(mode, s.g, d.g, sa, da),
                  pd_channel(mode, s.b, d.b, sa, da), pd_channel(mode, s.a, d.a, sa, da)};
    }
    if (mode >= RVK_BLEND_HUE) {
        float R, G, B;
        if (mode == RVK_BLEND_HUE) {
            R = s.r * sa; G = s.g * sa; B = s.b * sa;
            set_sat(R, G, B, sat3(d.r, d.g, d.b) * sa);
            set_lum(R, G, B, lum(d.r, d.g, d.b) * sa);
        } else if (mode == RVK_BLEND_SATURATION) {
            R = d.r * sa; G = d.g * sa; B = d.b * sa;
            set_sat(R, G, B, sat3(s.r, s.g, s.b) * da);
            set_lum(R, G, B, lum(d.r, d.g, d.b) * sa);
        } else if (mode == RVK_BLEND_COLOR) {
            R = s.r * da; G = s.g * da; B = s.b * da;
            set_lum(R, G, B, lum(d.r, d.g, d.b) * sa);
        } else {
            R = d.r * sa; G = d.g * sa; B = d.b * sa;
            set_lum(R, G, B, lum(s.r, s.g, s.b) * da);
        }
        clip_color(R, G, B, sa * da);
        return F4{s.r * inv(da) + d.r * inv(sa) + R, s.g * inv(da) + d.g * inv(sa) + G,
                  s.b * inv(da) + d.b * inv(sa) + B, sa + da - sa * da};
    }
    return F4{sep_channel(mode, s.r, d.r, sa, da), sep_channel(mode, s.g, d.g, sa, da),
              sep_channel(mode, s.b, d.b, sa, da), d.a * inv(s.a) + s.a};
}

}  // namespace rvkd
