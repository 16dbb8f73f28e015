// tiny_skia::PixmapMut::draw_pixmap with PixmapPaint{opacity, blend_mode, Nearest} and
// an identity transform (the only form resvg uses: render.rs:127-140,
// filter/mod.rs:624-1111): every destination pixel covered by the source rectangle
// goes through the f32 pipeline gather -> scale by opacity -> load dst -> blend -> store.
#include "rvk_internal.h"
#include "blend_ops.h"

using namespace rvkd;

namespace {

template <uint32_t MODE>
__global__ void k_draw(uint32_t *dst, uint32_t dw, const uint32_t *src, uint32_t sw, int32_t ox, int32_t oy,
                       uint32_t x0, uint32_t y0, uint32_t rw, uint32_t rh, float opacity, uint32_t mode) {
    size_t n = (size_t)rw * rh;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t x = x0 + (uint32_t)(i % rw), y = y0 + (uint32_t)(i / rw);
        uint32_t sp = src[(size_t)((int64_t)y - oy) * sw + (size_t)((int64_t)x - ox)];
        size_t di = (size_t)y * dw + x;
        F4 s = load_px(sp);
        if (opacity != 1.0f) { s.r *= opacity; s.g *= opacity; s.b *= opacity; s.a *= opacity; }
        if (MODE == RVK_BLEND_SOURCE_OVER) {
            // fast path: a transparent source leaves the destination unchanged
            if (s.a == 0.0f && s.r == 0.0f && s.g == 0.0f && s.b == 0.0f) continue;
            F4 d = load_px(dst[di]);
            dst[di] = store_px(blend(RVK_BLEND_SOURCE_OVER, s, d));
        } else {
            F4 d = load_px(dst[di]);
            dst[di] = store_px(blend(mode, s, d));
        }
    }
}

}  // namespace

namespace rvk {

rvk_status launch_draw_pixmap(rvk_context *ctx, uint8_t *dst, uint32_t dw, uint32_t dh, int32_t x, int32_t y,
                              const uint8_t *src, uint32_t sw, uint32_t sh, float opacity, uint32_t mode) {
    if (mode >= RVK_BLEND_COUNT) return RVK_ERR_INVALID_ARGUMENT;
    int64_t x0 = x > 0 ? x : 0, y0 = y > 0 ? y : 0;
    int64_t x1 = std::min<int64_t>((int64_t)x + sw, dw), y1 = std::min<int64_t>((int64_t)y + sh, dh);
    if (x1 <= x0 || y1 <= y0) return RVK_OK;
    if (mode == RVK_BLEND_DESTINATION) return RVK_OK;
    uint32_t rw = (uint32_t)(x1 - x0), rh = (uint32_t)(y1 - y0);
    size_t n = (size_t)rw * rh;
    size_t g = (n + 255) / 256;
    if (g > 65535u * 4u) g = 65535u * 4u;
    // source read + destination read and write: 12 bytes per pixel of the rect
    timed_launch(ctx, "draw_pixmap", 12.0 * (double)n, [&] {
        if (mode == RVK_BLEND_SOURCE_OVER)
            hipLaunchKernelGGL(k_draw<RVK_BLEND_SOURCE_OVER>, dim3((unsigned)g), dim3(256), 0, ctx->stream,
                               (uint32_t *)dst, dw, (const uint32_t *)src, sw, x, y, (uint32_t)x0, (uint32_t)y0, rw,
                               rh, opacity, mode);
        else
            hipLaunchKernelGGL(k_draw<RVK_BLEND_COUNT>, dim3((unsigned)g), dim3(256), 0, ctx->stream, (uint32_t *)dst,
                               dw, (const uint32_t *)src, sw, x, y, (uint32_t)x0, (uint32_t)y0, rw, rh, opacity, mode);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

}  // namespace rvk
