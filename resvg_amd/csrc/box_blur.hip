// Box blur (crates/resvg/src/filter/box_blur.rs).  Five rounds of
// [vertical pass, horizontal pass]; every pass rounds to u8, so the ten passes are
// run in the reference's order.  One kernel per round: a 2-D tile plus its halo is
// staged in LDS, the vertical running sums fill an LDS intermediate, and the
// horizontal running sums read it back, so each round reads and writes HBM once.
// The per-pass arithmetic is the reference's: exact integer window sums (zero
// outside the image), times f32 1/(2r+1), round-to-nearest-even, cast to u8.
#include "rvk_internal.h"
#include "device_math.h"
#include <mutex>

using namespace rvkd;

namespace {

constexpr int TW = 64;        // output tile width
constexpr int TH = 32;        // output tile height
constexpr int kThreads = 256;
constexpr int kMaxFusedR = 24;  // halo limit of the fused kernel (LDS budget)

// Back buffer of rvk_box_blur_noalloc: 64 MiB, a 4096 x 4096 RGBA8 image.
constexpr size_t kNoAllocScratchBytes = rvk::kNoAllocBoxBlurBytes;
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

__device__ __forceinline__ uint32_t round_pass(uint32_t s, float iarr) {
    // round() of box_blur.rs:327-331 is round-to-nearest-even on [0, 2^22].
    return (uint32_t)rintf((float)s * iarr);
}

// The same rounding, as the bits of a float: the product rounded to f32 (no fused
// multiply-add), plus 1.5 * 2^23, whose unit in the last place is 1, so the sum is
// rounded to the nearest integer (ties to even) and that integer, at most 255, is the
// low byte of the sum's bits.
__device__ __forceinline__ uint32_t round_pass_bits(uint32_t s, float iarr) {
#pragma clang fp contract(off)
    const float prod = (float)s * iarr;
    return __float_as_uint(prod + 12582912.0f);
}

// Sums four u8 channels of packed pixels into four 16-bit lanes of two u32s.
struct Acc {
    uint32_t rb, ga;  // (r | b << 16), (g | a << 16)
    // byte selects (v_perm_b32; selector 0x0c gives a zero byte)
    __device__ __forceinline__ static uint32_t lo(uint32_t p) { return __builtin_amdgcn_perm(0u, p, 0x0c020c00u); }
    __device__ __forceinline__ static uint32_t hi(uint32_t p) { return __builtin_amdgcn_perm(0u, p, 0x0c030c01u); }
    __device__ __forceinline__ void zero() { rb = 0; ga = 0; }
    __device__ __forceinline__ void add(uint32_t p) { rb += lo(p); ga += hi(p); }
    __device__ __forceinline__ void sub(uint32_t p) { rb -= lo(p); ga -= hi(p); }
    __device__ __forceinline__ uint32_t out(float iarr) const {
        const uint32_t r = round_pass_bits(rb & 0xffffu, iarr), g = round_pass_bits(ga & 0xffffu, iarr);
        const uint32_t b = round_pass_bits(rb >> 16, iarr), a = round_pass_bits(ga >> 16, iarr);
        return __builtin_amdgcn_perm(g, r, 0x0c0c0400u) | __builtin_amdgcn_perm(a, b, 0x04000c0cu);
    }
};

// One round: dst = H_rh(V_rv(src)).  Tile covers output [x0, x0+TW) x [y0, y0+TH).
// LDS: in[(TH + 2 rv) rows][(TW + 2 rh) cols], mid[TH rows][(TW + 2 rh) | 1 cols].
__global__ __launch_bounds__(kThreads) void k_box_round(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                        uint32_t w, uint32_t h, int rh, int rv) {
    extern __shared__ uint32_t lds[];
    const int cw = TW + 2 * rh;   // staged columns
    const int ch_ = TH + 2 * rv;  // staged rows
    const int ms = cw | 1;  // odd row stride of mid: the horizontal pass's lanes (rows) hit distinct banks
    uint32_t *in = lds;
    uint32_t *mid = lds + cw * ch_;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    // stage the input tile with zero padding outside the image (fv = lv = 0), in chunks
    // of 4 pixels aligned in memory: chunk c of a tile row holds global x in [a0 + 4c,
    // a0 + 4c + 3]; 16-byte loads where the image and its rows are 16-byte aligned and
    // the chunk lies inside the image, single pixels elsewhere
    {
        const int gx0 = x0 - rh, gy0 = y0 - rv;
        const int a0 = gx0 & ~3;
        const int nc = (gx0 + cw - a0 + 3) >> 2;
        const float inv_nc = 1.0f / (float)nc;  // (i + 0.5) / nc truncates exactly for these sizes
        const bool vec = (w & 3u) == 0 && (reinterpret_cast<uintptr_t>(src) & 15u) == 0;
        for (int i = threadIdx.x; i < ch_ * nc; i += kThreads) {
            const int ly = (int)(((float)i + 0.5f) * inv_nc), c = i - ly * nc;
            const int gy = gy0 + ly, gx = a0 + 4 * c;
            uint32_t v[4];
            if (gy >= 0 && gy < (int)h && vec && gx >= 0 && gx + 3 < (int)w) {
                const uint4 q = *reinterpret_cast<const uint4 *>(src + (size_t)gy * w + gx);
                v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++)
                    v[e] = (gy >= 0 && gy < (int)h && gx + e >= 0 && gx + e < (int)w) ? src[(size_t)gy * w + gx + e] : 0u;
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int lx = gx + e - gx0;
                if (lx >= 0 && lx < cw) in[ly * cw + lx] = v[e];
            }
        }
    }
    __syncthreads();
    // vertical pass: each thread walks one staged column over TH/segs rows
    if (rv == 0) {
        for (int i = threadIdx.x; i < cw * TH; i += kThreads) {
            const int ly = i / cw;
            mid[ly * ms + (i - ly * cw)] = in[i];
        }
    } else {
        const float iarr = 1.0f / (float)(2 * rv + 1);
        // two segments of TH / 2 rows per column: every column's 2 jobs fit one pass of
        // the block (cw <= 112), and the window sum is started (2 rv + 1 adds) only twice
        const int segs = 2, seg_len = TH / segs;
        for (int job = threadIdx.x; job < cw * segs; job += kThreads) {
            int col = job % cw, seg = job / cw;
            int t0 = seg * seg_len;
            Acc acc; acc.zero();
            for (int k = 0; k <= 2 * rv; k++) acc.add(in[(t0 + k) * cw + col]);
            mid[t0 * ms + col] = acc.out(iarr);
            for (int t = t0 + 1; t < t0 + seg_len; t++) {
                acc.add(in[(t + 2 * rv) * cw + col]);
                acc.sub(in[(t - 1) * cw + col]);
                mid[t * ms + col] = acc.out(iarr);
            }
        }
    }
    __syncthreads();
    // horizontal pass: each thread walks a row segment of 8 outputs
    constexpr int segw = 8, nseg = TW / segw;
    const bool vec_out = (w & 3u) == 0 && (reinterpret_cast<uintptr_t>(dst) & 15u) == 0;
    const float iarr = 1.0f / (float)(2 * rh + 1);
    for (int job = threadIdx.x; job < TH * nseg; job += kThreads) {
        int row = job % TH, seg = job / TH;
        int gy = y0 + row;
        if (gy >= (int)h) continue;
        const uint32_t *m = mid + row * ms;
        int c0 = seg * segw;
        uint32_t outv[segw];
        if (rh == 0) {
            for (int k = 0; k < segw; k++) outv[k] = m[c0 + k];
        } else {
            Acc acc; acc.zero();
            for (int k = 0; k <= 2 * rh; k++) acc.add(m[c0 + k]);
            outv[0] = acc.out(iarr);
            for (int k = 1; k < segw; k++) {
                acc.add(m[c0 + k + 2 * rh]);
                acc.sub(m[c0 + k - 1]);
                outv[k] = acc.out(iarr);
            }
        }
        int gx = x0 + c0;
        uint32_t *d = dst + (size_t)gy * w + gx;
        if (vec_out && gx + segw <= (int)w) {  // 32 bytes, 16-byte aligned
            reinterpret_cast<uint4 *>(d)[0] = make_uint4(outv[0], outv[1], outv[2], outv[3]);
            reinterpret_cast<uint4 *>(d)[1] = make_uint4(outv[4], outv[5], outv[6], outv[7]);
        } else {
            for (int k = 0; k < segw; k++)
                if (gx + k < (int)w) d[k] = outv[k];
        }
    }
}

// 32-bit per-channel sums for the large-radius fallback (255 * (2r + 1) may exceed 16 bits).
struct Acc32 {
    uint32_t c[4];
    __device__ __forceinline__ void zero() { c[0] = c[1] = c[2] = c[3] = 0; }
    __device__ __forceinline__ void add(uint32_t p) { for (int i = 0; i < 4; i++) c[i] += (p >> (8 * i)) & 0xffu; }
    __device__ __forceinline__ void sub(uint32_t p) { for (int i = 0; i < 4; i++) c[i] -= (p >> (8 * i)) & 0xffu; }
    __device__ __forceinline__ uint32_t out(float iarr) const {
        return pack(round_pass(c[0], iarr), round_pass(c[1], iarr), round_pass(c[2], iarr), round_pass(c[3], iarr));
    }
};

// Fallback for radii beyond the fused kernel's LDS budget: one thread per column
// (vertical) or per row (horizontal) with a running sum read straight from memory.
__global__ void k_box_vert(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint32_t w, uint32_t h,
                           int r, int chunk) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y0 = blockIdx.y * chunk;
    if (x >= w || y0 >= h) return;
    uint32_t y1 = min(h, y0 + (uint32_t)chunk);
    const float iarr = 1.0f / (float)(2 * r + 1);
    auto at = [&](int64_t y) -> uint32_t { return (y >= 0 && y < (int64_t)h) ? src[(size_t)y * w + x] : 0u; };
    Acc32 acc; acc.zero();
    for (int64_t k = (int64_t)y0 - r; k <= (int64_t)y0 + r; k++) acc.add(at(k));
    dst[(size_t)y0 * w + x] = acc.out(iarr);
    for (uint32_t y = y0 + 1; y < y1; y++) {
        acc.add(at((int64_t)y + r));
        acc.sub(at((int64_t)y - r - 1));
        dst[(size_t)y * w + x] = acc.out(iarr);
    }
}

__global__ void k_box_horz(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint32_t w, uint32_t h,
                           int r, int chunk) {
    uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t x0 = blockIdx.y * chunk;
    if (y >= h || x0 >= w) return;
    uint32_t x1 = min(w, x0 + (uint32_t)chunk);
    const float iarr = 1.0f / (float)(2 * r + 1);
    const uint32_t *row = src + (size_t)y * w;
    auto at = [&](int64_t x) -> uint32_t { return (x >= 0 && x < (int64_t)w) ? row[x] : 0u; };
    Acc32 acc; acc.zero();
    for (int64_t k = (int64_t)x0 - r; k <= (int64_t)x0 + r; k++) acc.add(at(k));
    dst[(size_t)y * w + x0] = acc.out(iarr);
    for (uint32_t x = x0 + 1; x < x1; x++) {
        acc.add(at((int64_t)x + r));
        acc.sub(at((int64_t)x - r - 1));
        dst[(size_t)y * w + x] = acc.out(iarr);
    }
}

}  // namespace

namespace rvk {

// create_box_gauss (box_blur.rs:37-71), evaluated in f32 exactly as the reference.
static void box_gauss_sizes(float sigma, int32_t out[5]) {
    if (sigma > 0.0f) {
        const float n = 5.0f;
        float w_ideal = std::sqrt(12.0f * sigma * sigma / n) + 1.0f;
        float fl = std::floor(w_ideal);
        int32_t wl = fl >= 2147483647.0f ? INT32_MAX : (int32_t)fl;
        if (wl % 2 == 0) wl -= 1;
        int32_t wu = wl + 2;
        float wlf = (float)wl;
        float m_ideal = (12.0f * sigma * sigma - n * wlf * wlf - 4.0f * n * wlf - 3.0f * n) / (-4.0f * wlf - 4.0f);
        float mr = std::round(m_ideal);
        size_t m = mr > 0.0f ? (size_t)mr : 0;
        for (size_t i = 0; i < 5; i++) out[i] = i < m ? wl : wu;
    } else {
        for (int i = 0; i < 5; i++) out[i] = 1;
    }
}

// The five rounds with `tmp` (w * h * 4 bytes of device memory) as the back buffer;
// allocates nothing.
static rvk_status box_blur_with(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy,
                                void *tmp) {
    int32_t bh[5], bv[5];
    box_gauss_sizes((float)sx, bh);
    box_gauss_sizes((float)sy, bv);
    const size_t bytes = (size_t)w * h * 4;
    uint32_t *a = (uint32_t *)px, *b = (uint32_t *)tmp;
    rvk_status st = RVK_OK;
    // algorithmic bytes: the image read once and written once per blur (8 B/px,
    // whatever the pass count), split evenly over the five rounds
    const double round_bytes = 8.0 * (double)w * h / 5.0;
    for (int i = 0; i < 5 && st == RVK_OK; i++) {
        int rh = (bh[i] - 1) / 2, rv = (bv[i] - 1) / 2;
        if (rh <= kMaxFusedR && rv <= kMaxFusedR) {
            size_t lds = (size_t)((TW + 2 * rh) * (TH + 2 * rv) + ((TW + 2 * rh) | 1) * TH) * 4;
            dim3 grid(div_up(w, TW), div_up(h, TH));
            timed_launch(ctx, "box_blur_round", round_bytes, [&] {
                hipLaunchKernelGGL(k_box_round, grid, dim3(kThreads), lds, ctx->stream, a, b, w, h, rh, rv);
            });
            if (hipGetLastError() != hipSuccess) { st = RVK_ERR_HIP; break; }
            std::swap(a, b);
        } else {
            const int chunk = 256;
            // vertical: a -> b, horizontal: b -> a
            if (rv == 0) {
                if (hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) { st = RVK_ERR_HIP; break; }
            } else {
                timed_launch(ctx, "box_blur_vert", round_bytes / 2.0, [&] {
                    hipLaunchKernelGGL(k_box_vert, dim3(div_up(w, 64), div_up(h, chunk)), dim3(64), 0, ctx->stream,
                                       a, b, w, h, rv, chunk);
                });
            }
            if (rh == 0) {
                if (hipMemcpyAsync(a, b, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) { st = RVK_ERR_HIP; break; }
            } else {
                timed_launch(ctx, "box_blur_horz", round_bytes / 2.0, [&] {
                    hipLaunchKernelGGL(k_box_horz, dim3(div_up(h, 64), div_up(w, chunk)), dim3(64), 0, ctx->stream,
                                       b, a, w, h, rh, chunk);
                });
            }
            if (hipGetLastError() != hipSuccess) { st = RVK_ERR_HIP; break; }
        }
    }
    if (st == RVK_OK && a != (uint32_t *)px) {
        if (hipMemcpyAsync(px, a, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) st = RVK_ERR_HIP;
    }
    return st;
}

rvk_status launch_box_blur(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    void *tmp = nullptr;
    RVK_TRY(scratch_alloc(ctx, (size_t)w * h * 4, &tmp));
    const rvk_status st = box_blur_with(ctx, px, w, h, sx, sy, tmp);
    scratch_release(ctx, tmp);
    return st;
}

// ---------------------------------------------------------------------------
// The no-allocation form: the back buffer is a file-scope device array, one per GPU
// for the life of the process.  Its address is looked up once per device, with that
// device current, when a context is created on it; host code uses only that address
// (the array's own name is a host-side stand-in, not a device pointer).
// ---------------------------------------------------------------------------
constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kMaxDevices] = {};

rvk_status noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return RVK_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return RVK_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        set_last_error("hipGetSymbolAddress(scratch)", __FILE__);
        return RVK_ERR_HIP;
    }
    g_noalloc_ptr[device] = p;
    return RVK_OK;
}

rvk_status launch_box_blur_noalloc(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    if ((size_t)w * h * 4 > kNoAllocScratchBytes) return RVK_ERR_TOO_LARGE;
    void *tmp = nullptr;
    if (ctx->device >= 0 && ctx->device < kMaxDevices) {
        std::lock_guard<std::mutex> g(g_noalloc_mu);
        tmp = g_noalloc_ptr[ctx->device];
    }
    if (!tmp) return RVK_ERR_HIP;  // the context's set-up could not look the array up
    return box_blur_with(ctx, px, w, h, sx, sy, tmp);
}

}  // namespace rvk
