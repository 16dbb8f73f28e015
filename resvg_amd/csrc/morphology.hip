// Morphology (crates/resvg/src/filter/morphology.rs:15-73).  The reference takes a
// per-channel min (erode) or max (dilate) over the asymmetric window
// [x - t, x - t + n - 1] x [y - ty, y - ty + m - 1], skipping taps outside the image.
// Min/max are separable and exact, so the 2-D window is computed as a horizontal
// pass into LDS followed by a vertical pass, over a tile with its halo.  Channels
// are processed two at a time as packed u16 lanes.
#include "rvk_internal.h"
#include "device_math.h"

using namespace rvkd;

namespace {

constexpr int TW = 64, TH = 16, kThreads = 256;
constexpr int kMaxFusedN = 32;

// per-16-bit-lane min/max of (r|b<<16) and (g|a<<16) forms
__device__ __forceinline__ uint32_t lane_min(uint32_t a, uint32_t b) {
    uint32_t lo = min(a & 0xffffu, b & 0xffffu), hi = min(a >> 16, b >> 16);
    return lo | (hi << 16);
}
__device__ __forceinline__ uint32_t lane_max(uint32_t a, uint32_t b) {
    uint32_t lo = max(a & 0xffffu, b & 0xffffu), hi = max(a >> 16, b >> 16);
    return lo | (hi << 16);
}

template <bool ERODE>
__device__ __forceinline__ uint2 comb(uint2 a, uint2 b) {
    return ERODE ? make_uint2(lane_min(a.x, b.x), lane_min(a.y, b.y)) : make_uint2(lane_max(a.x, b.x), lane_max(a.y, b.y));
}

__device__ __forceinline__ uint2 split(uint32_t p) { return make_uint2(p & 0x00ff00ffu, (p >> 8) & 0x00ff00ffu); }
__device__ __forceinline__ uint32_t join(uint2 v) { return v.x | (v.y << 8); }

// identity element: out-of-image taps are skipped, which equals combining with it
template <bool ERODE>
__device__ __forceinline__ uint2 ident() { return ERODE ? make_uint2(0x00ff00ffu, 0x00ff00ffu) : make_uint2(0u, 0u); }

// Window of n columns starting at x - tx0; window of m rows starting at y - ty0.
template <bool ERODE>
__global__ __launch_bounds__(kThreads) void k_morph(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                    uint32_t w, uint32_t h, int n, int tx0, int m, int ty0) {
    extern __shared__ uint2 sm[];
    const int cw = TW + n - 1, chh = TH + m - 1;
    uint2 *in = sm;              // chh x cw
    uint2 *mid = sm + cw * chh;  // chh x TW
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    for (int i = threadIdx.x; i < cw * chh; i += kThreads) {
        int ly = i / cw, lx = i - ly * cw;
        int gx = x0 - tx0 + lx, gy = y0 - ty0 + ly;
        in[i] = (gx >= 0 && gx < (int)w && gy >= 0 && gy < (int)h) ? split(src[(size_t)gy * w + gx]) : ident<ERODE>();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < chh * TW; i += kThreads) {
        int ly = i / TW, lx = i - ly * TW;
        const uint2 *row = in + ly * cw + lx;
        uint2 v = row[0];
        for (int k = 1; k < n; k++) v = comb<ERODE>(v, row[k]);
        mid[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TH * TW; i += kThreads) {
        int ly = i / TW, lx = i - ly * TW;
        int gx = x0 + lx, gy = y0 + ly;
        if (gx >= (int)w || gy >= (int)h) continue;
        const uint2 *col = mid + ly * TW + lx;
        uint2 v = col[0];
        for (int k = 1; k < m; k++) v = comb<ERODE>(v, col[k * TW]);
        dst[(size_t)gy * w + gx] = join(v);
    }
}

// Large windows: separable passes straight from memory, one output per thread.
template <bool ERODE>
__global__ void k_morph_h(const uint32_t *src, uint32_t *dst, uint32_t w, uint32_t h, int n, int tx0) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    int x = (int)(i % w), y = (int)(i / w);
    uint2 v = ident<ERODE>();
    int a = max(0, x - tx0), b = min((int)w - 1, x - tx0 + n - 1);
    const uint32_t *row = src + (size_t)y * w;
    for (int k = a; k <= b; k++) v = comb<ERODE>(v, split(row[k]));
    dst[i] = join(v);
}

template <bool ERODE>
__global__ void k_morph_v(const uint32_t *src, uint32_t *dst, uint32_t w, uint32_t h, int m, int ty0) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    int x = (int)(i % w), y = (int)(i / w);
    uint2 v = ident<ERODE>();
    int a = max(0, y - ty0), b = min((int)h - 1, y - ty0 + m - 1);
    for (int k = a; k <= b; k++) v = comb<ERODE>(v, split(src[(size_t)k * w + x]));
    dst[i] = join(v);
}

}  // namespace

namespace rvk {

rvk_status launch_morphology(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, uint32_t op, float rx,
                             float ry) {
    if (op > RVK_MORPH_DILATE) return RVK_ERR_INVALID_ARGUMENT;
    // columns = min(ceil(rx) as u32 * 2, width); target = floor(columns / 2) (morphology.rs:17-20)
    auto as_u32 = [](float v) -> uint32_t {
        if (!(v > 0.0f)) return 0u;
        if (v >= 4294967295.0f) return 0xffffffffu;
        return (uint32_t)v;
    };
    uint32_t columns = std::min(as_u32(std::ceil(rx)) * 2u, w);
    uint32_t rows = std::min(as_u32(std::ceil(ry)) * 2u, h);
    uint32_t tx0 = as_u32(std::floor((float)columns / 2.0f));
    uint32_t ty0 = as_u32(std::floor((float)rows / 2.0f));
    const bool erode = op == RVK_MORPH_ERODE;
    if (columns == 0 || rows == 0) {
        // no taps: every pixel keeps the initial value (255 for erode, 0 for dilate)
        return launch_fill_sub(ctx, px, w, h, 0, 0, w, h, erode ? 0xffffffffu : 0u);
    }
    void *tmp = nullptr;
    size_t bytes = (size_t)w * h * 4;
    RVK_TRY(scratch_alloc(ctx, bytes, &tmp));
    rvk_status st = RVK_OK;
    if ((int)columns <= kMaxFusedN && (int)rows <= kMaxFusedN) {
        size_t lds = (size_t)((TW + columns - 1) * (TH + rows - 1) + (TH + rows - 1) * TW) * sizeof(uint2);
        dim3 grid(div_up(w, TW), div_up(h, TH));
        timed_launch(ctx, "morphology", 2.0 * (double)bytes, [&] {
            if (erode)
                hipLaunchKernelGGL(k_morph<true>, grid, dim3(kThreads), lds, ctx->stream, (const uint32_t *)px,
                                   (uint32_t *)tmp, w, h, (int)columns, (int)tx0, (int)rows, (int)ty0);
            else
                hipLaunchKernelGGL(k_morph<false>, grid, dim3(kThreads), lds, ctx->stream, (const uint32_t *)px,
                                   (uint32_t *)tmp, w, h, (int)columns, (int)tx0, (int)rows, (int)ty0);
        });
        if (hipGetLastError() != hipSuccess) st = RVK_ERR_HIP;
        if (st == RVK_OK) st = launch_copy(ctx, px, (const uint8_t *)tmp, bytes);
    } else {
        size_t n = (size_t)w * h;
        unsigned g = (unsigned)((n + 255) / 256);
        timed_launch(ctx, "morphology_h", 8.0 * (double)n, [&] {
            if (erode)
                hipLaunchKernelGGL(k_morph_h<true>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)px,
                                   (uint32_t *)tmp, w, h, (int)columns, (int)tx0);
            else
                hipLaunchKernelGGL(k_morph_h<false>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)px,
                                   (uint32_t *)tmp, w, h, (int)columns, (int)tx0);
        });
        timed_launch(ctx, "morphology_v", 8.0 * (double)n, [&] {
            if (erode)
                hipLaunchKernelGGL(k_morph_v<true>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)tmp,
                                   (uint32_t *)px, w, h, (int)rows, (int)ty0);
            else
                hipLaunchKernelGGL(k_morph_v<false>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)tmp,
                                   (uint32_t *)px, w, h, (int)rows, (int)ty0);
        });
        if (hipGetLastError() != hipSuccess) st = RVK_ERR_HIP;
    }
    scratch_release(ctx, tmp);
    return st;
}

}  // namespace rvk
