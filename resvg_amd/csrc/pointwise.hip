// Per-pixel kernels: colour-space conversion (filter/mod.rs:115-238), colour matrix
// (filter/color_matrix.rs), component transfer (filter/component_transfer.rs),
// arithmetic composite (filter/composite.rs), displacement map
// (filter/displacement_map.rs), fills and copies.  One thread handles one pixel
// per grid-stride step; pixels are moved as packed u32 (RGBA8) so a wave touches
// 256 contiguous bytes per access.
#include "rvk_internal.h"
#include "device_math.h"

using namespace rvkd;

namespace {

// SRGB_TO_LINEAR_RGB_TABLE (filter/mod.rs:162-179) and LINEAR_RGB_TO_SRGB_TABLE
// (filter/mod.rs:195-212): the reference's own data, kept verbatim.
__constant__ uint8_t c_srgb_to_linear[256] = {
    0,   0,   0,   0,   0,   0,  0,    1,   1,   1,   1,   1,   1,   1,   1,   1,
    1,   1,   2,   2,   2,   2,  2,    2,   2,   2,   3,   3,   3,   3,   3,   3,
    4,   4,   4,   4,   4,   5,  5,    5,   5,   6,   6,   6,   6,   7,   7,   7,
    8,   8,   8,   8,   9,   9,  9,   10,  10,  10,  11,  11,  12,  12,  12,  13,
    13,  13,  14,  14,  15,  15,  16,  16,  17,  17,  17,  18,  18,  19,  19,  20,
    20,  21,  22,  22,  23,  23,  24,  24,  25,  25,  26,  27,  27,  28,  29,  29,
    30,  30,  31,  32,  32,  33,  34,  35,  35,  36,  37,  37,  38,  39,  40,  41,
    41,  42,  43,  44,  45,  45,  46,  47,  48,  49,  50,  51,  51,  52,  53,  54,
    55,  56,  57,  58,  59,  60,  61,  62,  63,  64,  65,  66,  67,  68,  69,  70,
    71,  72,  73,  74,  76,  77,  78,  79,  80,  81,  82,  84,  85,  86,  87,  88,
    90,  91,  92,  93,  95,  96,  97,  99, 100, 101, 103, 104, 105, 107, 108, 109,
    111, 112, 114, 115, 116, 118, 119, 121, 122, 124, 125, 127, 128, 130, 131, 133,
    134, 136, 138, 139, 141, 142, 144, 146, 147, 149, 151, 152, 154, 156, 157, 159,
    161, 163, 164, 166, 168, 170, 171, 173, 175, 177, 179, 181, 183, 184, 186, 188,
    190, 192, 194, 196, 198, 200, 202, 204, 206, 208, 210, 212, 214, 216, 218, 220,
    222, 224, 226, 229, 231, 233, 235, 237, 239, 242, 244, 246, 248, 250, 253, 255,
};
__constant__ uint8_t c_linear_to_srgb[256] = {
    0,  13,  22,  28,  34,  38,  42,  46,  50,  53,  56,  59,  61,  64,  66,  69,
    71,  73,  75,  77,  79,  81,  83,  85,  86,  88,  90,  92,  93,  95,  96,  98,
    99, 101, 102, 104, 105, 106, 108, 109, 110, 112, 113, 114, 115, 117, 118, 119,
    120, 121, 122, 124, 125, 126, 127, 128, 129, 130, 131, 132, 133, 134, 135, 136,
    137, 138, 139, 140, 141, 142, 143, 144, 145, 146, 147, 148, 148, 149, 150, 151,
    152, 153, 154, 155, 155, 156, 157, 158, 159, 159, 160, 161, 162, 163, 163, 164,
    165, 166, 167, 167, 168, 169, 170, 170, 171, 172, 173, 173, 174, 175, 175, 176,
    177, 178, 178, 179, 180, 180, 181, 182, 182, 183, 184, 185, 185, 186, 187, 187,
    188, 189, 189, 190, 190, 191, 192, 192, 193, 194, 194, 195, 196, 196, 197, 197,
    198, 199, 199, 200, 200, 201, 202, 202, 203, 203, 204, 205, 205, 206, 206, 207,
    208, 208, 209, 209, 210, 210, 211, 212, 212, 213, 213, 214, 214, 215, 215, 216,
    216, 217, 218, 218, 219, 219, 220, 220, 221, 221, 222, 222, 223, 223, 224, 224,
    225, 226, 226, 227, 227, 228, 228, 229, 229, 230, 230, 231, 231, 232, 232, 233,
    233, 234, 234, 235, 235, 236, 236, 237, 237, 238, 238, 238, 239, 239, 240, 240,
    241, 241, 242, 242, 243, 243, 244, 244, 245, 245, 246, 246, 246, 247, 247, 248,
    248, 249, 249, 250, 250, 251, 251, 251, 252, 252, 253, 253, 254, 254, 255, 255,
};

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g > 65535u * 4u) g = 65535u * 4u;
    return (unsigned)(g ? g : 1);
}

// demultiply -> LUT -> premultiply in one read and one write (Image::into_color_space).
__global__ void k_into_cs(uint32_t *px, size_t n, int to_linear) {
    const uint8_t *lut = to_linear ? c_srgb_to_linear : c_linear_to_srgb;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t p = demul_px(px[i]);
        p = pack(lut[ch(p, 0)], lut[ch(p, 1)], lut[ch(p, 2)], ch(p, 3));
        px[i] = premul_px(p);
    }
}

__global__ void k_demul(uint32_t *px, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        px[i] = demul_px(px[i]);
}

__global__ void k_premul(uint32_t *px, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        px[i] = premul_px(px[i]);
}

__device__ __forceinline__ uint32_t from_normalized(float c) { return as_u8(f32_bound(0.0f, c, 1.0f) * 255.0f); }

struct ColorMatrixArgs { uint32_t kind; float m[20]; };

// color_matrix::apply (color_matrix.rs:11-110); with fuse_alpha the demultiply before
// and the multiply after (filter/mod.rs:912-914) run in the same pass.
__global__ void k_color_matrix(uint32_t *px, size_t n, ColorMatrixArgs a, int fuse_alpha) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t p = px[i];
        if (fuse_alpha) p = demul_px(p);
        float r = (float)ch(p, 0) / 255.0f, g = (float)ch(p, 1) / 255.0f, b = (float)ch(p, 2) / 255.0f,
              al = (float)ch(p, 3) / 255.0f;
        const float *m = a.m;
        uint32_t o;
        if (a.kind == RVK_COLOR_MATRIX_MATRIX) {
            float nr = r * m[0] + g * m[1] + b * m[2] + al * m[3] + m[4];
            float ng = r * m[5] + g * m[6] + b * m[7] + al * m[8] + m[9];
            float nb = r * m[10] + g * m[11] + b * m[12] + al * m[13] + m[14];
            float na = r * m[15] + g * m[16] + b * m[17] + al * m[18] + m[19];
            o = pack(from_normalized(nr), from_normalized(ng), from_normalized(nb), from_normalized(na));
        } else if (a.kind == RVK_COLOR_MATRIX_LUMINANCE_TO_ALPHA) {
            float na = r * 0.2125f + g * 0.7154f + b * 0.0721f;
            o = pack(0, 0, 0, from_normalized(na));
        } else {
            float nr = r * m[0] + g * m[1] + b * m[2];
            float ng = r * m[3] + g * m[4] + b * m[5];
            float nb = r * m[6] + g * m[7] + b * m[8];
            o = pack(from_normalized(nr), from_normalized(ng), from_normalized(nb), ch(p, 3));
        }
        if (fuse_alpha) o = premul_px(o);
        px[i] = o;
    }
}

// component transfer as four 256-entry tables built on the host from the
// reference's transfer() (component_transfer.rs:34-72): the function maps u8 to u8.
struct TransferLut { uint8_t t[4][256]; uint32_t active[4]; };

__global__ void k_component_transfer(uint32_t *px, size_t n, TransferLut lut, int fuse_alpha) {
    __shared__ uint8_t s[4][256];
    __shared__ uint32_t act[4];
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) s[i >> 8][i & 255] = lut.t[i >> 8][i & 255];
    if (threadIdx.x < 4) act[threadIdx.x] = lut.active[threadIdx.x];
    __syncthreads();
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t p = px[i];
        if (fuse_alpha) p = demul_px(p);
        uint32_t c0 = act[0] ? s[0][ch(p, 0)] : ch(p, 0);
        uint32_t c1 = act[1] ? s[1][ch(p, 1)] : ch(p, 1);
        uint32_t c2 = act[2] ? s[2][ch(p, 2)] : ch(p, 2);
        uint32_t c3 = act[3] ? s[3][ch(p, 3)] : ch(p, 3);
        p = pack(c0, c1, c2, c3);
        if (fuse_alpha) p = premul_px(p);
        px[i] = p;
    }
}

// composite::arithmetic (composite.rs:14-50)
__global__ void k_arithmetic(const uint32_t *s1, const uint32_t *s2, uint32_t *dst, size_t n, float k1, float k2,
                             float k3, float k4) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t a = s1[i], b = s2[i];
        auto calc = [&](uint32_t c1, uint32_t c2, float mx) {
            float i1 = (float)c1 / 255.0f, i2 = (float)c2 / 255.0f;
            float r = k1 * i1 * i2 + k2 * i1 + k3 * i2 + k4;
            return f32_bound(0.0f, r, mx);
        };
        float al = calc(ch(a, 3), ch(b, 3), 1.0f);
        if (approx_zero_ulps(al, 4)) continue;
        dst[i] = pack(as_u8(calc(ch(a, 0), ch(b, 0), al) * 255.0f), as_u8(calc(ch(a, 1), ch(b, 1), al) * 255.0f),
                      as_u8(calc(ch(a, 2), ch(b, 2), al) * 255.0f), as_u8(al * 255.0f));
    }
}

// displacement_map::apply (displacement_map.rs:15-62)
__global__ void k_displacement(const uint32_t *src, const uint32_t *map, uint32_t *dst, uint32_t w, uint32_t h,
                               float sx, float sy, float scale, uint32_t xch, uint32_t ych) {
    size_t n = (size_t)w * h;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(i % w), y = (uint32_t)(i / w);
        uint32_t m = map[i];
        float dx = (float)ch(m, (int)xch) / 255.0f - 0.5f;
        float dy = (float)ch(m, (int)ych) / 255.0f - 0.5f;
        int32_t ox = as_i32(roundf((float)x + dx * sx * scale));
        int32_t oy = as_i32(roundf((float)y + dy * sy * scale));
        if (ox >= 0 && ox < (int32_t)w && oy >= 0 && oy < (int32_t)h) dst[i] = src[(size_t)oy * w + ox];
    }
}

__global__ void k_fill_rect(uint32_t *px, uint32_t w, uint32_t x0, uint32_t y0, uint32_t rw, uint32_t rh,
                            uint32_t value) {
    size_t n = (size_t)rw * rh;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t x = x0 + (uint32_t)(i % rw), y = y0 + (uint32_t)(i / rw);
        px[(size_t)y * w + x] = value;
    }
}

// apply_tile's fill (filter/mod.rs:837-859): a Repeat pattern under an integer
// translation samples nearest, so pixel (x, y) takes the tile pixel at
// ((x + ox) mod tw, (y + oy) mod th); (ox, oy) is the tile phase, (tx0, ty0) the
// tile's corner in the source.  Columns are fixed per thread, rows strided.
__global__ void k_tile(uint32_t *dst, uint32_t w, uint32_t h, const uint32_t *src, uint32_t sw, uint32_t tx0,
                       uint32_t ty0, uint32_t tw, uint32_t th, uint32_t ox, uint32_t oy) {
    const uint32_t x = blockIdx.x * 64 + threadIdx.x;
    if (x >= w) return;
    const uint32_t col = tx0 + (x + ox) % tw;
    for (uint32_t y = blockIdx.y * 4 + threadIdx.y; y < h; y += gridDim.y * 4)
        dst[(size_t)y * w + x] = src[(size_t)(ty0 + (y + oy) % th) * sw + col];
}

__global__ void k_copy(uint4 *dst, const uint4 *src, size_t n16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Color::from_rgba8(..).apply_opacity(alpha/255).premultiply().to_color_u8() per pixel
// (filter/mod.rs:606-617)
__global__ void k_shadow_flood(uint32_t *px, size_t n, float cr, float cg, float cb, float ca) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float o = (float)(px[i] >> 24) / 255.0f;
        float a = f32_bound(0.0f, ca * f32_bound(0.0f, o, 1.0f), 1.0f);
        float r = cr, g = cg, b = cb;
        if (a != 1.0f) { r = r * a; g = g * a; b = b * a; }
        px[i] = pack(as_u8(r * 255.0f + 0.5f), as_u8(g * 255.0f + 0.5f), as_u8(b * 255.0f + 0.5f),
                     as_u8(a * 255.0f + 0.5f));
    }
}

// Pixmap::apply_mask with a mask made from a second pixmap (Mask::from_pixmap):
// every channel becomes div255(c * m) (lowp LoadMaskU8 + DestinationIn).  m is the
// source's alpha (kMaskAlpha), 255 minus it (kMaskInvAlpha: an inverted clip mask,
// clip.rs:25-27) or its luminance: the demultiplied colour with Rec. 709 weights,
// times alpha, scaled to 255 and rounded up (kMaskLuminance).
constexpr uint32_t kMaskAlpha = 0, kMaskInvAlpha = 1, kMaskLuminance = 2;

__device__ __forceinline__ uint32_t mask_value(uint32_t s, uint32_t mode) {
    const uint32_t a8 = s >> 24;
    if (mode == kMaskAlpha) return a8;
    if (mode == kMaskInvAlpha) return 255u - a8;
    float r = (float)(s & 0xff) / 255.0f, g = (float)((s >> 8) & 0xff) / 255.0f,
          b = (float)((s >> 16) & 0xff) / 255.0f;
    const float a = (float)a8 / 255.0f;
    if (a8 != 0) {
        r = r / a;
        g = g / a;
        b = b / a;
    }
    const float luma = r * 0.2126f + g * 0.7152f + b * 0.0722f;
    const float v = f32_bound(0.0f, (luma * a) * 255.0f, 255.0f);
    return (uint32_t)ceilf(v);
}

__global__ void k_apply_mask(uint32_t *px, const uint32_t *src, size_t n, uint32_t mode) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t m = mask_value(src[i], mode);
        const uint32_t d = px[i];
        if (m == 255u) continue;  // div255(c * 255) = c
        uint32_t o = 0;
        for (int k = 0; k < 4; k++) o |= ((((d >> (8 * k)) & 0xffu) * m + 255u) >> 8) << (8 * k);
        px[i] = o;
    }
}

// The tiny_skia::Mask operations on their own (a u8 plane, one byte per pixel).
// Each thread handles 4 consecutive pixels: one 16-byte pixel load and one 4-byte
// mask word when the group is whole and aligned, single pixels for the tail.
__global__ void k_mask_from_pixmap(const uint32_t *px, uint8_t *mask, size_t n, uint32_t mode) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; 4 * q < n; q += (size_t)gridDim.x * blockDim.x) {
        const size_t i = 4 * q;
        if (i + 4 <= n) {
            const uint4 p = *reinterpret_cast<const uint4 *>(px + i);
            const uint32_t w = mask_value(p.x, mode) | (mask_value(p.y, mode) << 8) | (mask_value(p.z, mode) << 16) |
                               (mask_value(p.w, mode) << 24);
            *reinterpret_cast<uint32_t *>(mask + i) = w;
        } else {
            for (size_t k = i; k < n; k++) mask[k] = (uint8_t)mask_value(px[k], mode);
        }
    }
}

__global__ void k_mask_invert(uint8_t *mask, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; 4 * q < n; q += (size_t)gridDim.x * blockDim.x) {
        const size_t i = 4 * q;
        if (i + 4 <= n) *reinterpret_cast<uint32_t *>(mask + i) ^= 0xffffffffu;
        else for (size_t k = i; k < n; k++) mask[k] = (uint8_t)(255u - mask[k]);
    }
}

__device__ __forceinline__ uint32_t scale_px(uint32_t d, uint32_t m) {
    if (m == 255u) return d;
    uint32_t o = 0;
    for (int k = 0; k < 4; k++) o |= ((((d >> (8 * k)) & 0xffu) * m + 255u) >> 8) << (8 * k);
    return o;
}

__global__ void k_apply_mask_u8(uint32_t *px, const uint8_t *mask, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; 4 * q < n; q += (size_t)gridDim.x * blockDim.x) {
        const size_t i = 4 * q;
        if (i + 4 <= n) {
            const uint32_t m = *reinterpret_cast<const uint32_t *>(mask + i);
            uint4 p = *reinterpret_cast<const uint4 *>(px + i);
            p.x = scale_px(p.x, m & 0xffu);
            p.y = scale_px(p.y, (m >> 8) & 0xffu);
            p.z = scale_px(p.z, (m >> 16) & 0xffu);
            p.w = scale_px(p.w, m >> 24);
            *reinterpret_cast<uint4 *>(px + i) = p;
        } else {
            for (size_t k = i; k < n; k++) px[k] = scale_px(px[k], mask[k]);
        }
    }
}

__global__ void k_fill_u8(uint8_t *p, size_t n, uint32_t v4) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; 4 * q < n; q += (size_t)gridDim.x * blockDim.x) {
        const size_t i = 4 * q;
        if (i + 4 <= n) *reinterpret_cast<uint32_t *>(p + i) = v4;
        else for (size_t k = i; k < n; k++) p[k] = (uint8_t)v4;
    }
}

__global__ void k_zero_rgb(uint32_t *px, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        px[i] &= 0xff000000u;
}

}  // namespace

namespace rvk {

static rvk_status post_launch(rvk_context *) {
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

// in-place per-pixel kernels read and write 4 bytes per pixel
static inline double rw_bytes(size_t n) { return 8.0 * (double)n; }

rvk_status launch_into_cs(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, int to_linear) {
    size_t n = (size_t)w * h;
    timed_launch(ctx, "into_color_space", rw_bytes(n), [&] {
        hipLaunchKernelGGL(k_into_cs, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, n, to_linear);
    });
    return post_launch(ctx);
}

rvk_status launch_demultiply(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h) {
    size_t n = (size_t)w * h;
    timed_launch(ctx, "demultiply", rw_bytes(n), [&] {
        hipLaunchKernelGGL(k_demul, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, n);
    });
    return post_launch(ctx);
}

rvk_status launch_multiply(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h) {
    size_t n = (size_t)w * h;
    timed_launch(ctx, "premultiply", rw_bytes(n), [&] {
        hipLaunchKernelGGL(k_premul, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, n);
    });
    return post_launch(ctx);
}

rvk_status launch_color_matrix(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p,
                               int fuse_alpha) {
    const auto &cm = p->u.color_matrix;
    ColorMatrixArgs a{};
    a.kind = cm.kind;
    std::memcpy(a.m, cm.values, sizeof a.m);
    // Saturate / hueRotate matrices are per-primitive constants: built on the host in
    // the reference's f32 order (color_matrix.rs:212-253).
    if (cm.kind == RVK_COLOR_MATRIX_SATURATE) {
        float v = cm.values[0] > 0.0f ? cm.values[0] : 0.0f;
        float s[9] = {0.213f + 0.787f * v, 0.715f - 0.715f * v, 0.072f - 0.072f * v,
                      0.213f - 0.213f * v, 0.715f + 0.285f * v, 0.072f - 0.072f * v,
                      0.213f - 0.213f * v, 0.715f - 0.715f * v, 0.072f + 0.928f * v};
        std::memcpy(a.m, s, sizeof s);
    } else if (cm.kind == RVK_COLOR_MATRIX_HUE_ROTATE) {
        const float rads_per_deg = 3.14159274101257324f / 180.0f;
        float ang = cm.values[0] * rads_per_deg;
        float a1 = std::cos(ang), a2 = std::sin(ang);
        float s[9] = {0.213f + 0.787f * a1 - 0.213f * a2, 0.715f - 0.715f * a1 - 0.715f * a2,
                      0.072f - 0.072f * a1 + 0.928f * a2, 0.213f - 0.213f * a1 + 0.143f * a2,
                      0.715f + 0.285f * a1 + 0.140f * a2, 0.072f - 0.072f * a1 - 0.283f * a2,
                      0.213f - 0.213f * a1 - 0.787f * a2, 0.715f - 0.715f * a1 + 0.715f * a2,
                      0.072f + 0.928f * a1 + 0.072f * a2};
        std::memcpy(a.m, s, sizeof s);
    } else if (cm.kind != RVK_COLOR_MATRIX_MATRIX && cm.kind != RVK_COLOR_MATRIX_LUMINANCE_TO_ALPHA) {
        return RVK_ERR_INVALID_ARGUMENT;
    }
    size_t n = (size_t)w * h;
    timed_launch(ctx, "color_matrix", rw_bytes(n), [&] {
        hipLaunchKernelGGL(k_color_matrix, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, n, a,
                           fuse_alpha);
    });
    return post_launch(ctx);
}

// transfer() of component_transfer.rs:34-72 evaluated for all 256 inputs.
static uint8_t host_as_u8(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}
static size_t host_as_usize(float v) {
    if (!(v > 0.0f)) return 0;
    return (size_t)v;
}

static void build_transfer_lut(const rvk_transfer_fn &f, const float *values, uint8_t *out, uint32_t *active) {
    bool dummy = f.kind == RVK_TRANSFER_IDENTITY ||
                 ((f.kind == RVK_TRANSFER_TABLE || f.kind == RVK_TRANSFER_DISCRETE) && f.table_len == 0);
    *active = dummy ? 0u : 1u;
    for (int i = 0; i < 256; i++) {
        if (dummy) { out[i] = (uint8_t)i; continue; }
        float c = (float)i / 255.0f;
        const float *v = values ? values + f.table_offset : nullptr;
        switch (f.kind) {
            case RVK_TRANSFER_TABLE: {
                size_t n = f.table_len - 1;
                size_t k = host_as_usize(std::floor(c * (float)n));
                if (k > n) k = n;
                if (k == n) c = v[k];
                else {
                    float vk = v[k], vk1 = v[k + 1], kf = (float)k, nf = (float)n;
                    c = vk + (c - kf / nf) * nf * (vk1 - vk);
                }
                break;
            }
            case RVK_TRANSFER_DISCRETE: {
                size_t n = f.table_len;
                size_t k = host_as_usize(std::floor(c * (float)n));
                c = v[k < n - 1 ? k : n - 1];
                break;
            }
            case RVK_TRANSFER_LINEAR: c = f.a * c + f.b; break;
            case RVK_TRANSFER_GAMMA: c = f.a * std::pow(c, f.b) + f.c; break;
            default: break;
        }
        float b = c > 1.0f ? 1.0f : (c >= 0.0f ? c : 0.0f);
        out[i] = host_as_u8(b * 255.0f);
    }
}

rvk_status launch_component_transfer(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h,
                                     const rvk_primitive *p, const float *values, int fuse_alpha) {
    TransferLut lut{};
    for (int c = 0; c < 4; c++) {
        const rvk_transfer_fn &f = p->u.component_transfer.func[c];
        if (f.kind > RVK_TRANSFER_GAMMA) return RVK_ERR_INVALID_ARGUMENT;
        if ((f.kind == RVK_TRANSFER_TABLE || f.kind == RVK_TRANSFER_DISCRETE) && f.table_len && !values)
            return RVK_ERR_INVALID_ARGUMENT;
        build_transfer_lut(f, values, lut.t[c], &lut.active[c]);
    }
    size_t n = (size_t)w * h;
    // the four tables travel by value in the kernel arguments (1040 bytes)
    timed_launch(ctx, "component_transfer", rw_bytes(n), [&] {
        hipLaunchKernelGGL(k_component_transfer, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, n,
                           lut, fuse_alpha);
    });
    return post_launch(ctx);
}

rvk_status launch_arithmetic(rvk_context *ctx, const uint8_t *s1, const uint8_t *s2, uint8_t *dst, size_t n,
                             float k1, float k2, float k3, float k4) {
    timed_launch(ctx, "composite_arithmetic", 12.0 * (double)n, [&] {  // two sources, one destination
        hipLaunchKernelGGL(k_arithmetic, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (const uint32_t *)s1,
                           (const uint32_t *)s2, (uint32_t *)dst, n, k1, k2, k3, k4);
    });
    return post_launch(ctx);
}

rvk_status launch_displacement(rvk_context *ctx, const uint8_t *src, const uint8_t *map, uint8_t *dst,
                               uint32_t w, uint32_t h, float sx, float sy, float scale, uint32_t xch,
                               uint32_t ych) {
    if (xch > 3 || ych > 3) return RVK_ERR_INVALID_ARGUMENT;
    size_t n = (size_t)w * h;
    timed_launch(ctx, "displacement_map", 12.0 * (double)n, [&] {  // source, map, destination
        hipLaunchKernelGGL(k_displacement, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (const uint32_t *)src,
                           (const uint32_t *)map, (uint32_t *)dst, w, h, sx, sy, scale, xch, ych);
    });
    return post_launch(ctx);
}

rvk_status launch_fill(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const uint8_t premul[4]) {
    uint32_t v = (uint32_t)premul[0] | ((uint32_t)premul[1] << 8) | ((uint32_t)premul[2] << 16) |
                 ((uint32_t)premul[3] << 24);
    return launch_fill_sub(ctx, px, w, h, 0, 0, w, h, v);
}

rvk_status launch_fill_sub(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, int64_t x0, int64_t y0,
                           int64_t x1, int64_t y1, uint32_t value) {
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > (int64_t)w) x1 = w;
    if (y1 > (int64_t)h) y1 = h;
    if (x1 <= x0 || y1 <= y0) return RVK_OK;
    size_t n = (size_t)(x1 - x0) * (size_t)(y1 - y0);
    timed_launch(ctx, "fill_rect", 4.0 * (double)n, [&] {  // writes only
        hipLaunchKernelGGL(k_fill_rect, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, w,
                           (uint32_t)x0, (uint32_t)y0, (uint32_t)(x1 - x0), (uint32_t)(y1 - y0), value);
    });
    return post_launch(ctx);
}

rvk_status launch_tile(rvk_context *ctx, uint8_t *dst, uint32_t w, uint32_t h, const uint8_t *src, uint32_t sw,
                       uint32_t sh, uint32_t tx0, uint32_t ty0, uint32_t tw, uint32_t th, uint32_t ox, uint32_t oy) {
    if (w == 0 || h == 0) return RVK_OK;
    if (!dst || !src || tw == 0 || th == 0 || (uint64_t)tx0 + tw > sw || (uint64_t)ty0 + th > sh || ox >= tw ||
        oy >= th || w > (1u << 30) || h > (1u << 30))
        return RVK_ERR_INVALID_ARGUMENT;
    const uint32_t gy = std::min<uint32_t>((h + 3) / 4, 65535u);
    timed_launch(ctx, "tile", 8.0 * (double)w * h, [&] {
        hipLaunchKernelGGL(k_tile, dim3((w + 63) / 64, gy), dim3(64, 4), 0, ctx->stream, (uint32_t *)dst, w, h,
                           (const uint32_t *)src, sw, tx0, ty0, tw, th, ox, oy);
    });
    return post_launch(ctx);
}

rvk_status launch_copy(rvk_context *ctx, uint8_t *dst, const uint8_t *src, size_t bytes) {
    if (bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0) {
        size_t n = bytes / 16;
        timed_launch(ctx, "copy", 2.0 * (double)bytes, [&] {
            hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint4 *)dst,
                               (const uint4 *)src, n);
        });
        return post_launch(ctx);
    }
    RVK_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return RVK_OK;
}

}  // namespace rvk

namespace rvkh {
rvk_status launch_shadow_flood(rvk_context *ctx, uint8_t *px, size_t n, const uint8_t color[4], float opacity) {
    float op = opacity < 0.0f ? 0.0f : (opacity > 1.0f ? 1.0f : opacity);
    uint8_t a8 = rvk::host_as_u8(op * 255.0f + 0.5f);
    rvk::timed_launch(ctx, "shadow_flood", 8.0 * (double)n, [&] {
        hipLaunchKernelGGL(k_shadow_flood, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, n,
                           (float)color[0] / 255.0f, (float)color[1] / 255.0f, (float)color[2] / 255.0f,
                           (float)a8 / 255.0f);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}
}  // namespace rvkh

namespace rvk {

rvk_status launch_apply_mask(rvk_context *ctx, uint8_t *px, const uint8_t *src, size_t n, uint32_t mode) {
    timed_launch(ctx, "apply_mask", 12.0 * (double)n, [&] {  // reads both pixmaps, writes one
        hipLaunchKernelGGL(k_apply_mask, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px,
                           (const uint32_t *)src, n, mode);
    });
    return post_launch(ctx);
}

// The mask kernels require 4-byte aligned masks and 16-byte aligned pixmaps (the
// callers' buffers: device allocations, torch tensors); the C entry points check it.
rvk_status launch_mask_from_pixmap(rvk_context *ctx, const uint8_t *px, uint8_t *mask, size_t n, uint32_t type) {
    const uint32_t mode = type == RVK_MASK_TYPE_LUMINANCE ? kMaskLuminance : kMaskAlpha;
    const size_t groups = (n + 3) / 4;
    timed_launch(ctx, "mask_from_pixmap", 5.0 * (double)n, [&] {  // 4 B read, 1 B written per pixel
        hipLaunchKernelGGL(k_mask_from_pixmap, dim3(grid_for(groups)), dim3(kBlock), 0, ctx->stream,
                           (const uint32_t *)px, mask, n, mode);
    });
    return post_launch(ctx);
}

rvk_status launch_mask_invert(rvk_context *ctx, uint8_t *mask, size_t n) {
    const size_t groups = (n + 3) / 4;
    timed_launch(ctx, "mask_invert", 2.0 * (double)n, [&] {
        hipLaunchKernelGGL(k_mask_invert, dim3(grid_for(groups)), dim3(kBlock), 0, ctx->stream, mask, n);
    });
    return post_launch(ctx);
}

rvk_status launch_apply_mask_u8(rvk_context *ctx, uint8_t *px, const uint8_t *mask, size_t n) {
    const size_t groups = (n + 3) / 4;
    timed_launch(ctx, "apply_mask", 9.0 * (double)n, [&] {  // pixel read + write, mask read
        hipLaunchKernelGGL(k_apply_mask_u8, dim3(grid_for(groups)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px,
                           mask, n);
    });
    return post_launch(ctx);
}

rvk_status launch_fill_u8(rvk_context *ctx, uint8_t *p, size_t n, uint8_t value) {
    const size_t groups = (n + 3) / 4;
    const uint32_t v4 = 0x01010101u * value;
    timed_launch(ctx, "fill_mask", (double)n, [&] {
        hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(groups)), dim3(kBlock), 0, ctx->stream, p, n, v4);
    });
    return post_launch(ctx);
}

rvk_status launch_zero_rgb(rvk_context *ctx, uint8_t *px, size_t n) {
    timed_launch(ctx, "zero_rgb", rw_bytes(n), [&] {
        hipLaunchKernelGGL(k_zero_rgb, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, (uint32_t *)px, n);
    });
    return post_launch(ctx);
}

}  // namespace rvk
