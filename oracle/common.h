// Oracle helpers shared by the restatement files: Rust numeric semantics and the
// strict-num ULP comparison.  Test infrastructure only.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <chrono>

#include <functional>

namespace orc {

// feImage content drawing (filter/mod.rs:861-889): set by the render walk while it
// applies a group's filters; draws the children of scene group `root` with the given
// transform (sx, ky, kx, sy, tx, ty) into a w x h pixmap; false on a bad group.
using FeImageHook = std::function<bool(uint32_t root, const float ts[6], uint8_t *px, uint32_t w, uint32_t h)>;
extern thread_local FeImageHook fe_image_hook;

// Pattern tiles (render_pattern_pixmap, path.rs:179-205): set by the render walk
// around each path it draws; draws the children of scene group `root` with the given
// transform into a w x h pixmap with the walk's Context; false on a bad group.
using PatternHook = std::function<bool(uint32_t root, const float ts[6], uint8_t *px, uint32_t w, uint32_t h)>;
extern thread_local PatternHook pattern_hook;

// Time split of a render by operation (filter primitive kinds, then fills, strokes,
// layer composites and whole filters), for the CPU baseline's per-primitive split.
// Off unless orc_prof_enable(1) was called; a Scope nested in another counts in both.
namespace prof {
enum Slot : int { kFill = 32, kStroke, kLayer, kFilter, kSlots };
extern bool enabled;
extern double ms[kSlots];
struct Scope {
    int slot;
    std::chrono::steady_clock::time_point t0;
    explicit Scope(int s) : slot(enabled ? s : -1) {
        if (slot >= 0) t0 = std::chrono::steady_clock::now();
    }
    ~Scope() {
        if (slot >= 0) ms[slot] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};
}  // namespace prof

// Rust `f32 as u8`: saturating, NaN -> 0, truncation toward zero.
static inline uint8_t as_u8(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}
static inline uint8_t as_u8d(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 255.0) return 255;
    return (uint8_t)v;
}
// Rust `f32 as i32` (saturating, NaN -> 0).
static inline int32_t as_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483647.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}
static inline int32_t as_i32d(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)v;
}
static inline int64_t as_i64d(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}
static inline uint32_t as_u32(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 4294967295.0f) return UINT32_MAX;
    return (uint32_t)v;
}
static inline size_t as_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551615.0f) return SIZE_MAX;
    return (size_t)v;
}

// f32_bound (crates/resvg/src/filter/mod.rs:242-254).
static inline float f32_bound(float mn, float v, float mx) {
    if (v > mx) return mx;
    if (v >= mn) return v;
    return mn;
}

static inline uint32_t f32_bits(float v) { uint32_t u; std::memcpy(&u, &v, 4); return u; }
static inline uint64_t f64_bits(double v) { uint64_t u; std::memcpy(&u, &v, 8); return u; }

// strict-num ApproxEqUlps for f32 (used through usvg::ApproxZeroUlps,
// crates/usvg/src/tree/geom.rs:9-24): equal, or same sign and within `ulps`
// representable steps.
static inline bool approx_eq_ulps(float a, float b, int32_t ulps) {
    if (a == b) return true;
    if (std::signbit(a) != std::signbit(b)) return false;
    int32_t d = (int32_t)(f32_bits(a) - f32_bits(b));
    return d >= -ulps && d <= ulps;
}
static inline bool approx_zero_ulps(float a, int32_t ulps) { return approx_eq_ulps(a, 0.0f, ulps); }
static inline bool approx_eq_ulps_d(double a, double b, int64_t ulps) {
    if (a == b) return true;
    if (std::signbit(a) != std::signbit(b)) return false;
    int64_t d = (int64_t)(f64_bits(a) - f64_bits(b));
    return d >= -ulps && d <= ulps;
}

// Rust f32::to_radians: self * (PI / 180.0) with the constant folded in f32.
static inline float to_radians(float deg) {
    const float rads_per_deg = 3.14159274101257324f / 180.0f;
    return deg * rads_per_deg;
}

struct IRect { int32_t x, y, w, h; };
static inline int32_t ir_right(IRect r) { return r.x + r.w; }
static inline int32_t ir_bottom(IRect r) { return r.y + r.h; }
static inline bool ir_eq(IRect a, IRect b) { return a.x == b.x && a.y == b.y && a.w == b.w && a.h == b.h; }

}  // namespace orc
