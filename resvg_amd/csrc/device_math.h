// Device-side numeric helpers with the reference's Rust semantics.  All HIP code
// is compiled with -ffp-contract=off so a*b+c is never fused, matching Rust.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rvkd {

// Rust `f32 as u8`: saturating, NaN -> 0, truncation toward zero.
__device__ __forceinline__ uint32_t as_u8(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 255.0f) return 255u;
    return (uint32_t)v;
}
__device__ __forceinline__ uint32_t as_u8d(double v) {
    if (!(v > 0.0)) return 0u;
    if (v >= 255.0) return 255u;
    return (uint32_t)v;
}
__device__ __forceinline__ int32_t as_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)v;
}
__device__ __forceinline__ int32_t as_i32d(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (-2147483647 - 1);
    return (int32_t)v;
}
__device__ __forceinline__ int64_t as_i64d(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return 9223372036854775807LL;
    if (v <= -9223372036854775808.0) return (-9223372036854775807LL - 1);
    return (int64_t)v;
}

// f32_bound (crates/resvg/src/filter/mod.rs:242-254)
__device__ __forceinline__ float f32_bound(float mn, float v, float mx) {
    if (v > mx) return mx;
    if (v >= mn) return v;
    return mn;
}

// strict-num ApproxEqUlps (crates/usvg/src/tree/geom.rs:9-24)
__device__ __forceinline__ bool approx_eq_ulps(float a, float b, int32_t ulps) {
    if (a == b) return true;
    if (signbit(a) != signbit(b)) return false;
    int32_t d = (int32_t)(__float_as_uint(a) - __float_as_uint(b));
    return d >= -ulps && d <= ulps;
}
__device__ __forceinline__ bool approx_zero_ulps(float a, int32_t ulps) { return approx_eq_ulps(a, 0.0f, ulps); }

__device__ __forceinline__ uint32_t pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return r | (g << 8) | (b << 16) | (a << 24);
}
__device__ __forceinline__ uint32_t ch(uint32_t p, int c) { return (p >> (8 * c)) & 0xffu; }

// multiply_alpha / demultiply_alpha on one packed pixel (filter/mod.rs:129-146)
__device__ __forceinline__ uint32_t premul_px(uint32_t p) {
    float a = (float)ch(p, 3) / 255.0f;
    uint32_t r = as_u8((float)ch(p, 0) * a + 0.5f);
    uint32_t g = as_u8((float)ch(p, 1) * a + 0.5f);
    uint32_t b = as_u8((float)ch(p, 2) * a + 0.5f);
    return pack(r, g, b, ch(p, 3));
}
__device__ __forceinline__ uint32_t demul_px(uint32_t p) {
    float a = (float)ch(p, 3) / 255.0f;
    uint32_t r = as_u8((float)ch(p, 0) / a + 0.5f);
    uint32_t g = as_u8((float)ch(p, 1) / a + 0.5f);
    uint32_t b = as_u8((float)ch(p, 2) / a + 0.5f);
    return pack(r, g, b, ch(p, 3));
}

}  // namespace rvkd
