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

// Channels are processed as two u32 words of 16-bit lanes, (r | b << 16) and
// (g | a << 16), so that one v_pk_min_u16 / v_pk_max_u16 combines two channels.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

template <bool ERODE>
__device__ __forceinline__ uint32_t pk(uint32_t a, uint32_t b) {
    const u16x2 x = __builtin_bit_cast(u16x2, a), y = __builtin_bit_cast(u16x2, b);
    return __builtin_bit_cast(uint32_t, ERODE ? __builtin_elementwise_min(x, y) : __builtin_elementwise_max(x, y));
}

template <bool ERODE>
__device__ __forceinline__ uint2 comb(uint2 a, uint2 b) {
    return make_uint2(pk<ERODE>(a.x, b.x), pk<ERODE>(a.y, b.y));
}

// byte selects (v_perm_b32; selector 0x0c gives a zero byte): (r, 0, b, 0) and (g, 0, a, 0)
__device__ __forceinline__ uint2 split(uint32_t p) {
    return make_uint2(__builtin_amdgcn_perm(0u, p, 0x0c020c00u), __builtin_amdgcn_perm(0u, p, 0x0c030c01u));
}
__device__ __forceinline__ uint32_t join(uint2 v) { return __builtin_amdgcn_perm(v.y, v.x, 0x06020400u); }

// identity element: out-of-image taps are skipped, which equals combining with it
template <bool ERODE>
__device__ __forceinline__ uint2 ident() { return ERODE ? make_uint2(0x00ff00ffu, 0x00ff00ffu) : make_uint2(0u, 0u); }

// Fused tile kernel: MTW x MTH outputs per block (64 x 32: at 4096^2 faster than
// 64 x 64 for every window, and than 64 x 16 from 8 x 8 up), the input tile and its
// halo in LDS (u32 RGBA8), a horizontal pass into a second LDS buffer, then a
// vertical pass to memory.  In each pass a thread produces MK consecutive outputs of one row (or
// column) from a sliding window of n values (van Herk / Gil-Werman within the
// thread's run: out[j] = comb(suffix of v[j .. MK-2], core = v[MK-1 .. n-1], prefix
// of v[n .. n+j-1]), about (n + 3 MK) / MK combines per output instead of n - 1.
constexpr int MTW = 64, MTH = 32, MK = 8, kMThreads = 256;
constexpr int kMaxFusedN = 32;

// Windows narrower than MK: the run's MK + N - 1 values are read once; for N >= 4
// the window is the combination of two overlapping windows of L (the largest power
// of two <= N), each built by doubling.
template <bool ERODE, int N, class Get>
__device__ __forceinline__ void slide_fixed(Get get, uint2 (&out)[MK]) {
    constexpr int V = MK + N - 1;
    uint2 v[V];
#pragma unroll
    for (int i = 0; i < V; i++) v[i] = get(i);
    if constexpr (N < 4) {
#pragma unroll
        for (int j = 0; j < MK; j++) {
            uint2 a = v[j];
#pragma unroll
            for (int t = 1; t < N; t++) a = comb<ERODE>(a, v[j + t]);
            out[j] = a;
        }
    } else {
        constexpr int L = 4;  // 4 <= N <= 7
        uint2 m2[V - 1], m4[V - 3];
#pragma unroll
        for (int i = 0; i < V - 1; i++) m2[i] = comb<ERODE>(v[i], v[i + 1]);
#pragma unroll
        for (int i = 0; i < V - 3; i++) m4[i] = comb<ERODE>(m2[i], m2[i + 2]);
#pragma unroll
        for (int j = 0; j < MK; j++) out[j] = N == L ? m4[j] : comb<ERODE>(m4[j], m4[j + N - L]);
    }
}

// out[j] = comb of v[j .. j + n - 1] for j < MK; get(i) returns v[i] (i < MK + n - 1)
template <bool ERODE, class Get>
__device__ __forceinline__ void slide(int n, Get get, uint2 (&out)[MK]) {
    if (n >= MK) {
        uint2 sfx[MK - 1];  // sfx[j] = comb of v[j .. MK - 2]
        sfx[MK - 2] = get(MK - 2);
#pragma unroll
        for (int j = MK - 3; j >= 0; j--) sfx[j] = comb<ERODE>(get(j), sfx[j + 1]);
        uint2 core = get(MK - 1);
        for (int i = MK; i < n; i++) core = comb<ERODE>(core, get(i));
        out[0] = comb<ERODE>(sfx[0], core);
        uint2 pre = get(n);  // comb of v[n .. n + j - 1] for output j
#pragma unroll
        for (int j = 1; j < MK; j++) {
            out[j] = j < MK - 1 ? comb<ERODE>(comb<ERODE>(sfx[j], core), pre) : comb<ERODE>(core, pre);
            if (j < MK - 1) pre = comb<ERODE>(pre, get(n + j));
        }
    } else {
        switch (n) {  // uniform: one static form per window width
            case 1: slide_fixed<ERODE, 1>(get, out); break;
            case 2: slide_fixed<ERODE, 2>(get, out); break;
            case 3: slide_fixed<ERODE, 3>(get, out); break;
            case 4: slide_fixed<ERODE, 4>(get, out); break;
            case 5: slide_fixed<ERODE, 5>(get, out); break;
            case 6: slide_fixed<ERODE, 6>(get, out); break;
            default: slide_fixed<ERODE, 7>(get, out); break;
        }
    }
}

// Window of n columns starting at x - tx0; window of m rows starting at y - ty0
// (n, m <= kMaxFusedN).  src and dst must not overlap.
template <bool ERODE>
__global__ __launch_bounds__(kMThreads) void k_morph(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                     uint32_t w, uint32_t h, int n, int tx0, int m, int ty0) {
    extern __shared__ uint32_t sm[];
    const int cw = MTW + n - 1, rows = MTH + m - 1;
    constexpr int ms = MTW + 1;  // odd strides: the lanes of a wave hit distinct banks
    const int cs = cw | 1;
    uint32_t *in = sm;              // rows x cs
    uint32_t *mid = sm + rows * cs;  // rows x ms
    const int x0 = blockIdx.x * MTW, y0 = blockIdx.y * MTH;
    const uint32_t id = ERODE ? 0xffffffffu : 0u;
    // the tile and its halo, in chunks of 4 pixels aligned in memory: chunk c of a row
    // holds global x in [a0 + 4c, a0 + 4c + 3]; 16-byte loads where the image and its
    // rows are 16-byte aligned and the chunk lies inside the image, single pixels
    // elsewhere
    const int gx0 = x0 - tx0, gy0 = y0 - ty0;
    const int a0 = gx0 & ~3;
    const int nc = (gx0 + cw - a0 + 3) >> 2;
    const float inv_nc = 1.0f / (float)nc;  // (i + 0.5) / nc truncates exactly for these sizes
    const bool vec = (w & 3u) == 0 && (reinterpret_cast<uintptr_t>(src) & 15u) == 0;
    for (int i = threadIdx.x; i < rows * nc; i += kMThreads) {
        const int ly = (int)(((float)i + 0.5f) * inv_nc), c = i - ly * nc;
        const int gy = gy0 + ly, gx = a0 + 4 * c;
        uint32_t v[4];
        if (gy >= 0 && gy < (int)h && vec && gx >= 0 && gx + 3 < (int)w) {
            const uint4 q = *reinterpret_cast<const uint4 *>(src + (size_t)gy * w + gx);
            v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++)
                v[e] = (gy >= 0 && gy < (int)h && gx + e >= 0 && gx + e < (int)w) ? src[(size_t)gy * w + gx + e] : id;
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int lx = gx + e - gx0;
            if (lx >= 0 && lx < cw) in[ly * cs + lx] = v[e];
        }
    }
    __syncthreads();
    // horizontal: task = (row r, run g); consecutive lanes take consecutive rows
    for (int t = threadIdx.x; t < rows * (MTW / MK); t += kMThreads) {
        const int r = t % rows, g = t / rows;
        const uint32_t *row = in + r * cs + g * MK;
        uint2 out[MK];
        slide<ERODE>(n, [&](int i) { return split(row[i]); }, out);
#pragma unroll
        for (int j = 0; j < MK; j++) mid[r * ms + g * MK + j] = join(out[j]);
    }
    __syncthreads();
    // vertical: task = (column c, run g); consecutive lanes take consecutive columns
    for (int t = threadIdx.x; t < MTW * (MTH / MK); t += kMThreads) {
        const int c = t % MTW, g = t / MTW;
        const int gx = x0 + c;
        const uint32_t *col = mid + g * MK * ms + c;
        uint2 out[MK];
        slide<ERODE>(m, [&](int i) { return split(col[i * ms]); }, out);
        if (gx < (int)w) {
#pragma unroll
            for (int j = 0; j < MK; j++) {
                const int gy = y0 + g * MK + j;
                if (gy < (int)h) dst[(size_t)gy * w + gx] = join(out[j]);
            }
        }
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

// src -> dst (distinct buffers, w x h RGBA8).
static rvk_status morphology_to(rvk_context *ctx, const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                uint32_t op, float rx, float ry) {
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
        return launch_fill_sub(ctx, dst, w, h, 0, 0, w, h, erode ? 0xffffffffu : 0u);
    }
    const size_t bytes = (size_t)w * h * 4;
    if ((int)columns <= kMaxFusedN && (int)rows <= kMaxFusedN) {
        const int cw = MTW + (int)columns - 1, rws = MTH + (int)rows - 1;
        const size_t lds = (size_t)rws * ((cw | 1) + MTW + 1) * sizeof(uint32_t);
        dim3 grid(div_up(w, MTW), div_up(h, MTH));
        timed_launch(ctx, "morphology", 2.0 * (double)bytes, [&] {
            if (erode)
                hipLaunchKernelGGL(k_morph<true>, grid, dim3(kMThreads), lds, ctx->stream, (const uint32_t *)src,
                                   (uint32_t *)dst, w, h, (int)columns, (int)tx0, (int)rows, (int)ty0);
            else
                hipLaunchKernelGGL(k_morph<false>, grid, dim3(kMThreads), lds, ctx->stream, (const uint32_t *)src,
                                   (uint32_t *)dst, w, h, (int)columns, (int)tx0, (int)rows, (int)ty0);
        });
        RVK_HIP_CHECK(hipGetLastError());
        return RVK_OK;
    }
    void *tmp = nullptr;
    RVK_TRY(scratch_alloc(ctx, bytes, &tmp));
    size_t np = (size_t)w * h;
    unsigned g = (unsigned)((np + 255) / 256);
    timed_launch(ctx, "morphology_h", 8.0 * (double)np, [&] {
        if (erode)
            hipLaunchKernelGGL(k_morph_h<true>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)src,
                               (uint32_t *)tmp, w, h, (int)columns, (int)tx0);
        else
            hipLaunchKernelGGL(k_morph_h<false>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)src,
                               (uint32_t *)tmp, w, h, (int)columns, (int)tx0);
    });
    timed_launch(ctx, "morphology_v", 8.0 * (double)np, [&] {
        if (erode)
            hipLaunchKernelGGL(k_morph_v<true>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)tmp,
                               (uint32_t *)dst, w, h, (int)rows, (int)ty0);
        else
            hipLaunchKernelGGL(k_morph_v<false>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)tmp,
                               (uint32_t *)dst, w, h, (int)rows, (int)ty0);
    });
    rvk_status st = hipGetLastError() == hipSuccess ? RVK_OK : RVK_ERR_HIP;
    scratch_release(ctx, tmp);
    return st;
}

rvk_status launch_morphology_to(rvk_context *ctx, const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                uint32_t op, float rx, float ry) {
    return morphology_to(ctx, src, dst, w, h, op, rx, ry);
}

// In place: through a scratch copy of the input.
rvk_status launch_morphology(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, uint32_t op, float rx,
                             float ry) {
    void *tmp = nullptr;
    const size_t bytes = (size_t)w * h * 4;
    RVK_TRY(scratch_alloc(ctx, bytes, &tmp));
    rvk_status st = launch_copy(ctx, (uint8_t *)tmp, px, bytes);
    if (st == RVK_OK) st = morphology_to(ctx, (const uint8_t *)tmp, px, w, h, op, rx, ry);
    scratch_release(ctx, tmp);
    return st;
}

}  // namespace rvk
