// PNG scanlines of a pixmap: the device part of tiny-skia's Pixmap::encode_png, which
// resvg's CLI calls to write its output (crates/resvg/src/main.rs:108 and :113).
//
// Every pixel is demultiplied the way PremultipliedColorU8::demultiply does it:
// alpha 255 unchanged, otherwise each colour channel c / (a / 255) + 0.5 in f64 and
// cast to u8 (saturating, NaN to 0) -- the reference's test harness restates the same
// rule at crates/resvg/tests/integration/main.rs:204-211.  The 256 x 256 results are
// a table built on the device in f64 before the rows run.  Each row is then filtered
// with one of the five PNG filter types (ISO/IEC 15948 section 9: None, Sub, Up,
// Average, Paeth; 4 bytes per pixel, bytes left of the row and the row above the
// first count as 0), chosen per row as the type whose filtered bytes, read as signed,
// have the smallest sum of magnitudes (the minimum-sum heuristic of adaptive PNG
// encoders; ties go to the lower type).  The output is the zlib input of the IDAT
// stream: per row its filter type byte, then its 4 * w filtered bytes.
//
// One 256-thread block per row: a first pass sums the five candidates over the row,
// a second writes the chosen one.  A pixel's left neighbour comes from the next lower
// lane (the wave's first lane reads it), the row above from memory (it is also the
// previous block's row, so it is an L2 hit); 8 bytes per pixel are algorithmic.
#include "rvk_internal.h"

namespace {

constexpr int kPngBlock = 256;

// tbl[a * 256 + c] = demultiplied c at alpha a
__global__ __launch_bounds__(256) void k_png_demul_table(uint8_t *tbl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 65536u) return;
    const uint32_t a = i >> 8, c = i & 0xffu;
    uint32_t v = c;
    if (a != 255u) {
        const double d = (double)c / ((double)a / 255.0) + 0.5;
        v = d != d ? 0u : d >= 255.0 ? 255u : d <= 0.0 ? 0u : (uint32_t)d;
    }
    tbl[i] = (uint8_t)v;
}

__device__ __forceinline__ uint32_t straight(const uint8_t *tbl, uint32_t p) {
    const uint32_t a = p >> 24;
    if (a == 255u) return p;
    const uint8_t *row = tbl + a * 256u;
    return (uint32_t)row[p & 0xffu] | ((uint32_t)row[(p >> 8) & 0xffu] << 8) |
           ((uint32_t)row[(p >> 16) & 0xffu] << 16) | (a << 24);
}

__device__ __forceinline__ uint32_t paeth(uint32_t a, uint32_t b, uint32_t c) {
    const int p = (int)a + (int)b - (int)c;
    const int pa = abs(p - (int)a), pb = abs(p - (int)b), pc = abs(p - (int)c);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

// the five filtered bytes of byte x with left a, above b, upper-left c
__device__ __forceinline__ void filter5(uint32_t x, uint32_t a, uint32_t b, uint32_t c, uint32_t f[5]) {
    f[0] = x;
    f[1] = (x - a) & 0xffu;
    f[2] = (x - b) & 0xffu;
    f[3] = (x - ((a + b) >> 1)) & 0xffu;
    f[4] = (x - paeth(a, b, c)) & 0xffu;
}

__global__ __launch_bounds__(kPngBlock) void k_png_rows(const uint32_t *px, uint32_t w, uint32_t h,
                                                       const uint8_t *tbl, uint8_t *out) {
    const uint32_t y = blockIdx.x;
    const int tid = (int)threadIdx.x, lane = tid % 64;
    const uint32_t *row = px + (size_t)y * w;
    const uint32_t *up = y > 0 ? row - w : nullptr;
    uint8_t *dst = out + (size_t)y * (4u * (size_t)w + 1u);
    __shared__ uint32_t part[kPngBlock / 64][5];
    __shared__ uint32_t chosen;

    // the straight pixel at column i of this row and of the row above, and of their
    // left neighbours (0 outside)
    auto load = [&](uint32_t i, uint32_t &x, uint32_t &b, uint32_t &l, uint32_t &c) {
        const bool in = i < w;
        x = in ? straight(tbl, row[i]) : 0u;
        b = in && up ? straight(tbl, up[i]) : 0u;
        l = (uint32_t)__shfl_up((int)x, 1, 64);
        c = (uint32_t)__shfl_up((int)b, 1, 64);
        if (lane == 0) {
            l = in && i > 0 ? straight(tbl, row[i - 1]) : 0u;
            c = in && i > 0 && up ? straight(tbl, up[i - 1]) : 0u;
        }
        if (i == 0) l = c = 0u;
    };

    uint32_t sum[5] = {0, 0, 0, 0, 0};
    for (uint32_t base = 0; base < w; base += kPngBlock) {
        const uint32_t i = base + (uint32_t)tid;
        uint32_t x, b, l, c;
        load(i, x, b, l, c);
        if (i < w) {
            for (int k = 0; k < 4; k++) {
                uint32_t f[5];
                filter5((x >> (8 * k)) & 0xffu, (l >> (8 * k)) & 0xffu, (b >> (8 * k)) & 0xffu,
                        (c >> (8 * k)) & 0xffu, f);
                for (int t = 0; t < 5; t++) sum[t] += (uint32_t)abs((int)(int8_t)(uint8_t)f[t]);
            }
        }
    }
    // the row's sums: waves reduce by shuffles, then across the block's four waves
    for (int t = 0; t < 5; t++) {
        uint32_t s = sum[t];
        for (int o = 32; o > 0; o >>= 1) s += (uint32_t)__shfl_xor((int)s, o, 64);
        if (lane == 0) part[tid / 64][t] = s;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t best = 0;
        uint64_t best_sum = ~0ull;
        for (uint32_t t = 0; t < 5; t++) {
            uint64_t s = 0;
            for (int wv = 0; wv < kPngBlock / 64; wv++) s += part[wv][t];
            if (s < best_sum) { best_sum = s; best = t; }
        }
        chosen = best;
        dst[0] = (uint8_t)best;
    }
    __syncthreads();
    const uint32_t t = chosen;
    for (uint32_t base = 0; base < w; base += kPngBlock) {
        const uint32_t i = base + (uint32_t)tid;
        uint32_t x, b, l, c;
        load(i, x, b, l, c);
        if (i < w) {
            uint8_t *o = dst + 1 + 4u * (size_t)i;
            for (int k = 0; k < 4; k++) {
                uint32_t f[5];
                filter5((x >> (8 * k)) & 0xffu, (l >> (8 * k)) & 0xffu, (b >> (8 * k)) & 0xffu,
                        (c >> (8 * k)) & 0xffu, f);
                o[k] = (uint8_t)f[t];
            }
        }
    }
}

}  // namespace

extern "C" {

rvk_status rvk_png_filter_rows(rvk_context *ctx, const uint8_t *px, uint32_t w, uint32_t h, uint8_t *out) {
    if (!ctx || !px || !out || !rvk::valid_dims(w, h)) return RVK_ERR_INVALID_ARGUMENT;
    RVK_HIP_CHECK(hipSetDevice(ctx->device));
    void *tbl = nullptr;
    RVK_TRY(rvk::scratch_alloc(ctx, 65536, &tbl));
    rvk::timed_launch(ctx, "png_table", 65536.0, [&] {
        hipLaunchKernelGGL(k_png_demul_table, dim3(256), dim3(256), 0, ctx->stream, (uint8_t *)tbl);
    });
    rvk_status st = RVK_OK;
    if (hipGetLastError() != hipSuccess) st = RVK_ERR_HIP;
    if (st == RVK_OK) {
        // algorithmic bytes: each pixel read once, 4 bytes per pixel and 1 per row written
        const double bytes = 4.0 * (double)w * h + (4.0 * (double)w + 1.0) * h;
        rvk::timed_launch(ctx, "png_rows", bytes, [&] {
            hipLaunchKernelGGL(k_png_rows, dim3(h), dim3(kPngBlock), 0, ctx->stream, (const uint32_t *)px, w, h,
                               (const uint8_t *)tbl, out);
        });
        if (hipGetLastError() != hipSuccess) st = RVK_ERR_HIP;
    }
    rvk::scratch_release(ctx, tbl);
    return st;
}

}  // extern "C"
