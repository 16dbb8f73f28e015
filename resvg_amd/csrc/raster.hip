// Device rasterizer: anti-aliased coverage and the fused shade/composite of draw
// lists.
//
// Coverage (k_coverage): one wave per (coverage job, pixel row).  For each of the
// row's four sub-scanlines the wave gathers the active line pieces of the row's
// band, computes each crossing's 16.16 x, sorts the crossings by (x, edge id) with
// a bitonic sort in LDS, scans the windings to find the spans tiny-skia's walk
// would blit, and accumulates the SuperBlitter contributions (16 per covered
// sub-column of a partial pixel, 64 or 63 for a fully covered pixel) into an LDS
// row; full-pixel runs go through a difference array.  The row is then prefix-summed,
// clamped to 255 and written as u8.
//
// Compositing (k_draw): one thread per target pixel of a 64x4 tile, looping over
// the draw items in order with the pixel in a register: fill/stroke coverage ->
// gradient or solid shade (tiny-skia lowp stages) -> SourceOver into the target
// (direct items) or into an elided layer pixel that is then composited in f32 with
// the group's opacity and blend mode (draw_pixmap semantics).
#include "rvk_internal.h"
#include "blend_ops.h"
#include "raster.h"

using namespace rvkd;
using rvkr::CovJob;
using rvkr::DrawItem;
using rvkr::Piece;
using rvkr::ShaderParams;

namespace {

constexpr int kWave = 64;
constexpr int kMaxX = 1024;       // crossings per sub-scanline a wave can sort
constexpr int kRowWindow = 4608;  // pixels accumulated per pass over a row

struct CovArgs {
    const CovJob *jobs;
    uint32_t njobs;
    const Piece *pieces;
    const uint32_t *band_start;
    const uint32_t *band_items;
    uint8_t *arena;
    int32_t *overflow;
};

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
    for (int d = 1; d < kWave; d <<= 1) {
        int t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t find_job(const CovJob *jobs, uint32_t n, uint32_t row) {
    uint32_t lo = 0, hi = n;  // last job with row_start <= row
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].row_start <= row) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kWave) void k_coverage(CovArgs a) {
    __shared__ uint64_t keys[kMaxX];
    __shared__ int32_t ev[kMaxX];
    __shared__ int32_t row[kRowWindow + 1];  // acc (low 16 bits) + diff (high 16 bits, signed)
    const int lane = threadIdx.x;
    const uint32_t gr = blockIdx.x;
    const uint32_t ji = find_job(a.jobs, a.njobs, gr);
    const CovJob job = a.jobs[ji];
    const int32_t r = (int32_t)(gr - job.row_start);
    if (r < 0 || r >= job.h) return;
    const int shift = job.aa ? 2 : 0;
    const int32_t scale = 1 << shift;
    const int32_t py = job.y0 + r;
    const uint32_t stride = ((uint32_t)job.w + 3u) & ~3u;
    uint8_t *out = a.arena + job.cov_offset + (size_t)r * stride;
    const bool evenodd = job.rule == RVK_FILL_EVENODD;

    for (int32_t wx0 = 0; wx0 < job.w; wx0 += kRowWindow) {
        const int32_t ww = min(kRowWindow, job.w - wx0);
        for (int i = lane; i <= ww; i += kWave) row[i] = 0;
        __syncthreads();
        // absolute supersampled column of the window's left edge
        const int32_t colbase = (job.x0 + wx0) * scale;
        const int32_t colmax = ww * scale;
        for (int s = 0; s < scale; s++) {
            const int32_t sy = py * scale + s;
            // ---- gather active pieces of the row's band ----
            uint32_t n = 0;
            const int32_t band = (sy - job.band_y0) >> rvkr::kBandShift;
            if (band >= 0 && (uint32_t)band < job.band_count) {
                const uint32_t b0 = a.band_start[job.band_offset + band], b1 = a.band_start[job.band_offset + band + 1];
                for (uint32_t base = b0; base < b1; base += kWave) {
                    const uint32_t i = base + lane;
                    bool act = false;
                    uint64_t key = 0;
                    if (i < b1) {
                        const Piece p = a.pieces[job.piece_offset + a.band_items[i]];
                        if (p.first_y <= sy && sy <= p.last_y) {
                            int32_t x = (int32_t)((uint32_t)p.x + (uint32_t)(sy - p.first_y) * (uint32_t)p.dx);
                            key = ((uint64_t)((uint32_t)x ^ 0x80000000u) << 32) | ((uint64_t)(uint32_t)p.id << 1) |
                                  (p.winding > 0 ? 1u : 0u);
                            act = true;
                        }
                    }
                    const uint64_t m = __ballot(act);
                    const uint32_t pos = n + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                    if (act && pos < (uint32_t)kMaxX) keys[pos] = key;
                    n += (uint32_t)__popcll(m);
                }
            }
            if (n > (uint32_t)kMaxX) {
                if (lane == 0) atomicAdd(a.overflow, 1);
                n = kMaxX;
            }
            if (n == 0) continue;
            // ---- bitonic sort of keys[0, npow) ----
            uint32_t npow = 1;
            while (npow < n) npow <<= 1;
            for (uint32_t i = n + lane; i < npow; i += kWave) keys[i] = ~0ull;
            __syncthreads();
            for (uint32_t k = 2; k <= npow; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t i = lane; i < npow; i += kWave) {
                        const uint32_t ixj = i ^ j;
                        if (ixj > i) {
                            const uint64_t x0 = keys[i], x1 = keys[ixj];
                            const bool up = (i & k) == 0;
                            if ((x0 > x1) == up) { keys[i] = x1; keys[ixj] = x0; }
                        }
                    }
                    __syncthreads();
                }
            }
            // ---- winding scan -> span boundary events (start, end, start, end, ...) ----
            int carry = 0;
            uint32_t nev = 0;
            for (uint32_t base = 0; base < n; base += kWave) {
                const uint32_t i = base + lane;
                int w = 0;
                uint64_t key = 0;
                if (i < n) { key = keys[i]; w = (key & 1) ? 1 : -1; }
                const int incl = wave_incl_scan(w, lane);
                const int before = carry + incl - w, after = carry + incl;
                const bool in_b = evenodd ? (before & 1) != 0 : before != 0;
                const bool in_a = evenodd ? (after & 1) != 0 : after != 0;
                const bool evt = i < n && (in_b != in_a);
                const uint64_t m = __ballot(evt);
                if (evt) {
                    const int32_t x = (int32_t)((uint32_t)(key >> 32) ^ 0x80000000u);
                    ev[nev + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (x + 0x8000) >> 16;
                }
                nev += (uint32_t)__popcll(m);
                carry += __shfl(incl, kWave - 1, kWave);
            }
            __syncthreads();
            // ---- SuperBlitter contributions of each span ----
            const int32_t maxv = job.aa ? (64 - (((s & 3) + 1) >> 2)) : 255;
            for (uint32_t k = lane; 2 * k + 1 < nev; k += kWave) {
                int32_t L = ev[2 * k] - colbase, R = ev[2 * k + 1] - colbase;
                if (R <= L) continue;
                L = max(L, 0);
                R = min(R, colmax);
                if (R <= L) continue;
                if (!job.aa) {
                    atomicAdd(&row[L], 255 << 16);
                    atomicAdd(&row[R], -(255 << 16));
                    continue;
                }
                const int32_t fb = L & 3, fe = R & 3;
                int32_t cnt = (R >> 2) - (L >> 2) - 1;
                if (cnt < 0) {
                    atomicAdd(&row[L >> 2], (fe - fb) * 16);
                } else {
                    int32_t st = L >> 2;
                    if (fb == 0) cnt += 1;
                    else { atomicAdd(&row[st], (4 - fb) * 16); st += 1; }
                    if (cnt > 0) {
                        atomicAdd(&row[st], maxv << 16);
                        atomicAdd(&row[st + cnt], -(maxv << 16));
                    }
                    if (fe) atomicAdd(&row[R >> 2], fe * 16);
                }
            }
            __syncthreads();
        }
        // ---- prefix-sum the run differences, clamp, write u8 ----
        int carry = 0;
        for (int32_t base = 0; base < ww; base += 4 * kWave) {
            const int32_t p0 = base + 4 * lane;
            int v[4], d[4];
            int loc = 0;
            for (int q = 0; q < 4; q++) {
                const int32_t p = p0 + q;
                const int32_t packed = p < ww ? row[p] : 0;
                v[q] = packed & 0xffff;
                d[q] = packed >> 16;
                loc += d[q];
            }
            const int incl = wave_incl_scan(loc, lane);
            int run = carry + incl - loc;
            uint32_t word = 0;
            for (int q = 0; q < 4; q++) {
                run += d[q];
                const int c = min(255, v[q] + run);
                word |= (uint32_t)c << (8 * q);
            }
            const int32_t gx = wx0 + p0;
            if (gx + 3 < job.w && ((wx0 + p0) & 3) == 0) {
                *(uint32_t *)(out + gx) = word;
            } else {
                for (int q = 0; q < 4; q++)
                    if (gx + q < job.w) out[gx + q] = (uint8_t)(word >> (8 * q));
            }
            carry += __shfl(incl, kWave - 1, kWave);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// shading (tiny-skia lowp) and compositing
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t div255(uint32_t v) { return (v + 255u) >> 8; }
__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// premultiplied u8 source colour of a shader at pixel (px, py)
__device__ __forceinline__ void shade(const ShaderParams &s, int32_t px, int32_t py, uint32_t o[4]) {
    if (s.kind == 1) {
        o[0] = s.rgba & 0xff; o[1] = (s.rgba >> 8) & 0xff; o[2] = (s.rgba >> 16) & 0xff; o[3] = s.rgba >> 24;
        return;
    }
    const float x = (float)px + 0.5f, y = (float)py + 0.5f;
    const float tx = x * s.sx + (y * s.kx + s.tx);
    const float ty = x * s.ky + (y * s.sy + s.ty);
    float t = tx;
    if (s.radial) {
        t = sqrtf(tx * tx + ty * ty);
        if (s.r_twist) t = (t - s.r_scale) / s.r_bias;
    }
    if (s.spread == RVK_SPREAD_REFLECT) {
        const float u = t - 1.0f;
        t = clamp01(fabsf(u - 2.0f * floorf(u * 0.5f) - 1.0f));
    } else if (s.spread == RVK_SPREAD_REPEAT) {
        t = clamp01(t - floorf(t));
    } else if (s.pad_clamp) {
        t = clamp01(t);
    }
    float c[4];
    if (s.two_stop) {
        for (int k = 0; k < 4; k++) c[k] = t * s.f[0][k] + s.b[0][k];
    } else {
        uint32_t idx = 0;
        for (uint32_t i = 1; i < s.nstops; i++) idx += t >= s.t[i] ? 1u : 0u;
        for (int k = 0; k < 4; k++) c[k] = t * s.f[idx][k] + s.b[idx][k];
    }
    for (int k = 0; k < 4; k++) {
        const float v = (c[k] != c[k]) ? 0.0f : clamp01(c[k]);
        o[k] = (uint32_t)(v * 255.0f + 0.5f);
    }
    if (!s.opaque) for (int k = 0; k < 3; k++) o[k] = div255(o[k] * o[3]);
}

// lowp: scale by coverage (unless full), then SourceOver onto `d`; an opaque shader
// composites with Source instead (full coverage stores, partial coverage lerps).
__device__ __forceinline__ uint32_t over_lowp(const uint32_t src_in[4], uint32_t cov, uint32_t d, bool as_source) {
    if (as_source) {
        if (cov == 255) return src_in[0] | (src_in[1] << 8) | (src_in[2] << 16) | (src_in[3] << 24);
        const uint32_t ic = 255u - cov;
        uint32_t o = 0;
        for (int k = 0; k < 4; k++) o |= div255(((d >> (8 * k)) & 0xffu) * ic + src_in[k] * cov) << (8 * k);
        return o;
    }
    uint32_t src[4] = {src_in[0], src_in[1], src_in[2], src_in[3]};
    if (cov != 255) for (int k = 0; k < 4; k++) src[k] = div255(src[k] * cov);
    const uint32_t ia = 255u - src[3];
    uint32_t o = 0;
    for (int k = 0; k < 4; k++) o |= (src[k] + div255(((d >> (8 * k)) & 0xffu) * ia)) << (8 * k);
    return o;
}

__device__ __forceinline__ uint32_t job_cov(const CovJob *jobs, const uint8_t *arena, int32_t j, int32_t x,
                                            int32_t y) {
    if (j < 0) return 0;
    const CovJob &jb = jobs[j];
    const int32_t cx = x - jb.x0, cy = y - jb.y0;
    if (cx < 0 || cy < 0 || cx >= jb.w || cy >= jb.h) return 0;
    const uint32_t stride = ((uint32_t)jb.w + 3u) & ~3u;
    return arena[jb.cov_offset + (size_t)cy * stride + cx];
}

constexpr int TX = 64, TY = 4;

__global__ __launch_bounds__(TX * TY) void k_draw(uint32_t *px, uint32_t w, uint32_t h, int32_t ox, int32_t oy,
                                                 const DrawItem *items, uint32_t nitems, const CovJob *jobs,
                                                 const uint8_t *arena, const ShaderParams *shaders) {
    const int32_t x = ox + (int32_t)(blockIdx.x * TX + threadIdx.x % TX);
    const int32_t y = oy + (int32_t)(blockIdx.y * TY + threadIdx.x / TX);
    const int32_t tx0 = ox + (int32_t)(blockIdx.x * TX), ty0 = oy + (int32_t)(blockIdx.y * TY);
    const bool inside = x < (int32_t)w && y < (int32_t)h;
    uint32_t d = inside ? px[(size_t)y * w + x] : 0u;
    const uint32_t d0 = d;
    for (uint32_t it = 0; it < nitems; it++) {
        const DrawItem item = items[it];
        if (item.x1 <= tx0 || item.x0 >= tx0 + TX || item.y1 <= ty0 || item.y0 >= ty0 + TY) continue;
        if (!inside || x < item.x0 || x >= item.x1 || y < item.y0 || y >= item.y1) continue;
        const int32_t lx = x - item.lx, ly = y - item.ly;
        const uint32_t cf = job_cov(jobs, arena, item.fill_job, lx, ly);
        const uint32_t cs = job_cov(jobs, arena, item.stroke_job, lx, ly);
        if ((cf | cs) == 0) continue;
        uint32_t target = item.layer ? 0u : d;
        for (int pass = 0; pass < 2; pass++) {
            const bool stroke = (pass == 0) == (item.stroke_first != 0);
            const uint32_t c = stroke ? cs : cf;
            if (!c) continue;
            uint32_t src[4];
            const ShaderParams &sh = shaders[stroke ? item.stroke_shader : item.fill_shader];
            shade(sh, lx, ly, src);
            target = over_lowp(src, c, target, sh.as_source != 0);
        }
        if (!item.layer) {
            d = target;
        } else {
            F4 s = load_px(target);
            if (item.opacity != 1.0f) {
                s.r *= item.opacity; s.g *= item.opacity; s.b *= item.opacity; s.a *= item.opacity;
            }
            d = store_px(blend(item.blend_mode, s, load_px(d)));
        }
    }
    if (inside && d != d0) px[(size_t)y * w + x] = d;
}

}  // namespace

namespace rvk {

rvk_status launch_coverage(rvk_context *ctx, const CovJob *jobs, uint32_t njobs, uint32_t total_rows,
                           const Piece *pieces, const uint32_t *band_start, const uint32_t *band_items, uint8_t *arena,
                           int32_t *overflow, double algo_bytes) {
    if (total_rows == 0) return RVK_OK;
    CovArgs a{jobs, njobs, pieces, band_start, band_items, arena, overflow};
    timed_launch(ctx, "coverage", algo_bytes,
                 [&] { hipLaunchKernelGGL(k_coverage, dim3(total_rows), dim3(kWave), 0, ctx->stream, a); });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

rvk_status launch_draw_items(rvk_context *ctx, uint8_t *target, uint32_t w, uint32_t h, int32_t x0, int32_t y0,
                             int32_t x1, int32_t y1, const DrawItem *items, uint32_t nitems, const CovJob *jobs,
                             const uint8_t *arena, const ShaderParams *shaders, double algo_bytes) {
    if (nitems == 0 || x1 <= x0 || y1 <= y0) return RVK_OK;
    dim3 grid(div_up((uint32_t)(x1 - x0), TX), div_up((uint32_t)(y1 - y0), TY));
    timed_launch(ctx, "draw_items", algo_bytes, [&] {
        hipLaunchKernelGGL(k_draw, grid, dim3(TX * TY), 0, ctx->stream, (uint32_t *)target, w, h, x0, y0, items, nitems,
                           jobs, arena, shaders);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

}  // namespace rvk
