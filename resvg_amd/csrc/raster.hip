// Device rasterizer: curve expansion, anti-aliased coverage and the fused
// shade/composite of draw lists.
//
// Expansion (k_expand): one thread per edge runs tiny-skia's quadratic / cubic
// update loop (Skia's forward differencing in 16.16 fixed point) from the state the
// host left after the first piece, writing every non-empty line piece in order.
//
// Coverage (k_coverage): one 64-lane wave per (coverage job, pixel row).  The row's
// edge window comes from the job's table; each lane takes an edge, binary-searches
// its pieces for the row and produces the crossings of the row's sub-scanlines.
// Per sub-scanline the crossings (keys: x, edge id, winding) are sorted with a
// register-resident bitonic network (up to 8 keys per lane), the windings are
// scanned into span events (tiny-skia's walk_edges spans), and the SuperBlitter
// contributions (16 per covered sub-column of a partial pixel, 64 or 63 for a fully
// covered pixel) are accumulated per 1024-pixel window in LDS, prefix-summed,
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
using rvkr::EdgeCov;
using rvkr::EdgeRec;
using rvkr::HairJob;
using rvkr::HairLine;
using rvkr::PieceDev;
using rvkr::ShaderParams;
using rvkr::ShaderHead;
using rvkr::JobRef;
using rvkr::kMaxStops;

namespace {

constexpr int kWave = 64;
constexpr int kMaxKeys = (int)rvkr::kCoverageKeys[2];  // crossings per sub-scanline of the widest variant
constexpr int kWin = 512;         // pixels accumulated per window (256, 1024, 2048: slower)
constexpr int kMaxSub = 4;        // sub-scanlines per pixel row

// ---------------------------------------------------------------------------
// expansion
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t fdot6_round(int32_t x) { return (x + 32) >> 6; }
__device__ __forceinline__ int32_t fdot6_to_fixed(int32_t x) { return (int32_t)((uint32_t)x << 10); }
__device__ __forceinline__ int32_t fixed_mul(int32_t a, int32_t b) {
    return (int32_t)(((int64_t)a * (int64_t)b) >> 16);
}
__device__ __forceinline__ int32_t fdot6_div(int32_t a, int32_t b) {
    if ((int32_t)(int16_t)a == a) return (int32_t)((uint32_t)a << 16) / b;
    int64_t v = ((int64_t)a << 16) / (int64_t)b;
    return (int32_t)max((int64_t)INT32_MIN, min((int64_t)INT32_MAX, v));
}

// SkEdge::updateLine on 16.16 inputs
__device__ __forceinline__ bool line_piece(int32_t x0, int32_t y0, int32_t x1, int32_t y1, PieceDev &p) {
    y0 >>= 10;
    y1 >>= 10;
    const int32_t top = fdot6_round(y0), bot = fdot6_round(y1);
    if (top == bot) return false;
    x0 >>= 10;
    x1 >>= 10;
    const int32_t slope = fdot6_div(x1 - x0, y1 - y0);
    const int32_t dy = ((top << 6) + 32) - y0;
    p.x = fdot6_to_fixed(x0 + fixed_mul(slope, dy));
    p.dx = slope;
    p.first_y = top;
    p.last_y = bot - 1;
    return true;
}

struct ExpandArgs {
    const EdgeRec *edges;
    uint32_t nedges;
    const EdgeCov *cov;
    PieceDev *pieces;
    uint32_t *npieces;
    int32_t *overflow;
};

__global__ __launch_bounds__(256) void k_expand(ExpandArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nedges) return;
    EdgeRec e = a.edges[i];
    PieceDev *out = a.pieces + e.piece_offset;
    const int32_t ymax = a.cov[i].ymax;
    uint32_t n = 0;
    out[n++] = PieceDev{e.x, e.dx, e.first_y, e.last_y};
    int32_t oldx = e.cx, oldy = e.cy;
    if (e.type == 1) {  // SkQuadraticEdge::updateQuadratic, repeated until exhausted
        int32_t count = e.count, dx = e.d1x, dy = e.d1y;
        while (count > 0) {
            int32_t newx, newy;
            if (--count > 0) {
                newx = oldx + (dx >> e.shift);
                dx += e.d2x;
                newy = oldy + (dy >> e.shift);
                dy += e.d2y;
            } else {
                newx = e.lx;
                newy = e.ly;
            }
            PieceDev p;
            if (line_piece(oldx, oldy, newx, newy, p)) out[n++] = p;
            oldx = newx;
            oldy = newy;
        }
    } else if (e.type == 2) {  // SkCubicEdge::updateCubic
        int32_t count = e.count;
        int32_t d1x = e.d1x, d1y = e.d1y, d2x = e.d2x, d2y = e.d2y;
        while (count < 0) {
            int32_t newx, newy;
            if (++count < 0) {
                newx = oldx + (d1x >> e.dshift);
                d1x += d2x >> e.shift;
                d2x += e.d3x;
                newy = oldy + (d1y >> e.dshift);
                d1y += d2y >> e.shift;
                d2y += e.d3y;
            } else {
                newx = e.lx;
                newy = e.ly;
            }
            if (newy < oldy) newy = oldy;
            PieceDev p;
            if (line_piece(oldx, oldy, newx, newy, p)) out[n++] = p;
            oldx = newx;
            oldy = newy;
        }
    }
    a.npieces[i] = n;
    // the host's row bound must cover every piece (else rows would miss the edge)
    if (out[n - 1].last_y > ymax) atomicAdd(a.overflow, 1);
}

// ---------------------------------------------------------------------------
// coverage
// ---------------------------------------------------------------------------
struct CovArgs {
    const uint2 *groups;  // (job, first row in the job) of each row group this launch handles
    const CovJob *jobs;
    uint32_t njobs;
    const EdgeCov *edges;
    const uint2 *windows;
    const PieceDev *pieces;
    const uint32_t *npieces;
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

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, kWave);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// Lane exchange across xor distance LD (1..32) for the bitonic network: an
// immediate-pattern ds_swizzle within 32-lane halves, ds_bpermute across them.
template <int LD>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (LD < 32) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (LD << 10) | 0x1F);
    } else {
        return (uint32_t)__shfl_xor((int)v, LD, kWave);
    }
}

// One stage (k, jj) of an ascending bitonic network over S independent arrays of
// 64*K keys, element e = lane*K + j of array s in v[s][j].
template <int K, int S, int k, int jj>
__device__ __forceinline__ void bitonic_stage(uint64_t (&v)[S][K], int lane) {
    if constexpr (jj >= K) {
        constexpr int ld = jj / K;
        // e & k and e & jj depend on the lane only; keep the smaller key when the pair
        // is ascending and this is its lower element, or descending and upper
        const bool up = k >= kWave * K ? true : (lane & (k / K)) == 0;
        const bool keep_min = up == ((lane & ld) == 0);
#pragma unroll
        for (int s = 0; s < S; s++)
#pragma unroll
            for (int j = 0; j < K; j++) {
                const uint64_t x = v[s][j];
                const uint64_t o = ((uint64_t)xor_lane<ld>((uint32_t)(x >> 32)) << 32) | xor_lane<ld>((uint32_t)x);
                v[s][j] = ((o < x) == keep_min) ? o : x;
            }
    } else {
#pragma unroll
        for (int j = 0; j < K; j++) {
            if ((j & jj) == 0) {
                const int p = j | jj;
                const bool up = k >= K ? (k >= kWave * K ? true : (lane & (k / K)) == 0) : (j & k) == 0;
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const uint64_t x = v[s][j], y = v[s][p];
                    const bool sw = (x > y) == up;
                    v[s][j] = sw ? y : x;
                    v[s][p] = sw ? x : y;
                }
            }
        }
    }
}

template <int K, int S, int k, int jj>
__device__ __forceinline__ void bitonic_from(uint64_t (&v)[S][K], int lane) {
    bitonic_stage<K, S, k, jj>(v, lane);
    if constexpr (jj > 1) bitonic_from<K, S, k, jj / 2>(v, lane);
    else if constexpr (k < kWave * K) bitonic_from<K, S, k * 2, k>(v, lane);
}

// Sorts the S sub-scanlines' keys (n[s] each, in keys[s]) together and writes each
// one's span events (pixel-rounded supersampled columns; starts and ends alternate)
// over its own key area.  Returns the event counts in nev.
template <int K, int S, int MK>
__device__ __forceinline__ void spans_of(uint64_t (*keys)[MK], const int *n, bool evenodd, int *nev, int lane) {
    uint64_t v[S][K];
#pragma unroll
    for (int s = 0; s < S; s++)
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int e = lane * K + j;
            v[s][j] = e < n[s] ? keys[s][e] : ~0ull;
        }
    __syncthreads();  // every key is in registers before the key areas are reused for events
    bitonic_from<K, S, 2, 1>(v, lane);
#pragma unroll
    for (int s = 0; s < S; s++) {
        int w[K], local = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int e = lane * K + j;
            w[j] = e < n[s] ? ((v[s][j] & 1) ? 1 : -1) : 0;
            local += w[j];
        }
        const int incl = wave_incl_scan(local, lane);
        int before = incl - local;
        bool evt[K];
        int nevt = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int after = before + w[j];
            const bool in_b = evenodd ? (before & 1) != 0 : before != 0;
            const bool in_a = evenodd ? (after & 1) != 0 : after != 0;
            evt[j] = w[j] != 0 && in_b != in_a;
            nevt += evt[j] ? 1 : 0;
            before = after;
        }
        const int eincl = wave_incl_scan(nevt, lane);
        int pos = eincl - nevt;
        int32_t *ev = reinterpret_cast<int32_t *>(keys[s]);
#pragma unroll
        for (int j = 0; j < K; j++)
            if (evt[j]) {
                const int32_t x = (int32_t)((uint32_t)(v[s][j] >> 32) ^ 0x80000000u);
                ev[pos++] = (x + 0x8000) >> 16;
            }
        nev[s] = __shfl(eincl, kWave - 1, kWave);
    }
}

// MK: crossings per sub-scanline the launch can hold (128, 256 or 512; the keys take
// 4, 8 or 16 KB of LDS, so the smaller variants run two to three times the waves per
// CU).  The host sends a row group to the smallest variant that holds the number of
// edges whose rows meet the group, an upper bound on any of its sub-scanlines'
// crossings.  Block b handles row groups[b / kGroupRows].y + b % kGroupRows of job
// groups[b / kGroupRows].x.
template <int MK>
__global__ __launch_bounds__(kWave) void k_coverage(CovArgs a) {
    __shared__ uint64_t keys[kMaxSub][MK];  // crossings; reused for span events
    __shared__ __align__(16) int32_t acc[kWin + 4];  // acc (low 16 bits) + run difference (high 16 bits)
    const int lane = threadIdx.x;
    // the group's first row picks the job (a job's last group can be partial)
    const uint2 gs = a.groups[blockIdx.x / rvkr::kGroupRows];
    const CovJob job = a.jobs[gs.x];
    const int32_t r = (int32_t)gs.y + (int32_t)(blockIdx.x % rvkr::kGroupRows);
    if (r < 0 || r >= job.h) return;
    const int shift = job.aa ? 2 : 0;
    const int nsub = 1 << shift;
    const int32_t sy0 = (job.y0 + r) << shift;
    const uint32_t stride = ((uint32_t)job.w + 3u) & ~3u;
    uint8_t *out = a.arena + job.cov_offset + (size_t)r * stride;
    const bool evenodd = job.rule == RVK_FILL_EVENODD;

    // ---- gather the crossings of the row's sub-scanlines ----
    const uint2 win = a.windows[job.window_offset + (uint32_t)r / rvkr::kGroupRows];
    int cnt[kMaxSub] = {0, 0, 0, 0};
    for (uint32_t base = win.x; base < win.y; base += kWave) {
        const uint32_t i = base + lane;
        uint64_t key[kMaxSub];
        bool act[kMaxSub] = {false, false, false, false};
        if (i < win.y) {
            const EdgeCov e = a.edges[job.edge_offset + i];
            if (e.first_y <= sy0 + nsub - 1 && e.ymax >= sy0) {
                const PieceDev *pc = a.pieces + e.piece_offset;
                const uint32_t n = a.npieces[job.edge_offset + i];
                uint32_t lo = 0, hi = n;  // first piece with last_y >= sy0
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (pc[mid].last_y < sy0) lo = mid + 1;
                    else hi = mid;
                }
                uint32_t k = lo;
                for (int s = 0; s < nsub; s++) {
                    const int32_t sy = sy0 + s;
                    while (k < n && pc[k].last_y < sy) k++;
                    if (k < n && pc[k].first_y <= sy) {
                        const PieceDev p = pc[k];
                        const int32_t x = (int32_t)((uint32_t)p.x + (uint32_t)(sy - p.first_y) * (uint32_t)p.dx);
                        key[s] = ((uint64_t)((uint32_t)x ^ 0x80000000u) << 32) | ((uint64_t)(uint32_t)e.id << 1) |
                                 (e.winding > 0 ? 1u : 0u);
                        act[s] = true;
                    }
                }
            }
        }
        for (int s = 0; s < nsub; s++) {
            const uint64_t m = __ballot(act[s]);
            const int pos = cnt[s] + __popcll(m & ((1ull << lane) - 1ull));
            if (act[s] && pos < MK) keys[s][pos] = key[s];
            cnt[s] += __popcll(m);
        }
    }
    __syncthreads();
    // ---- sort the sub-scanlines together and turn them into span events ----
    int nev[kMaxSub] = {0, 0, 0, 0};
    int nmax = 0;
    for (int s = 0; s < nsub; s++) {
        if (cnt[s] > MK) {
            if (lane == 0) atomicAdd(a.overflow, 1);
            cnt[s] = MK;
        }
        nmax = max(nmax, cnt[s]);
    }
    // keys per lane: the smallest network that holds the fullest sub-scanline (cnt is
    // at most MK, so the branches past MK are never taken)
    constexpr int K4 = MK >= 256 ? 4 : 2, K8 = MK >= 512 ? 8 : K4;
    if (nsub == 1) {
        if (nmax == 0) {}
        else if (nmax <= 64) spans_of<1, 1, MK>(keys, cnt, evenodd, nev, lane);
        else if (nmax <= 128) spans_of<2, 1, MK>(keys, cnt, evenodd, nev, lane);
        else if (nmax <= 256) spans_of<K4, 1, MK>(keys, cnt, evenodd, nev, lane);
        else spans_of<K8, 1, MK>(keys, cnt, evenodd, nev, lane);
    } else {
        if (nmax == 0) {}
        else if (nmax <= 64) spans_of<1, kMaxSub, MK>(keys, cnt, evenodd, nev, lane);
        else if (nmax <= 128) spans_of<2, kMaxSub, MK>(keys, cnt, evenodd, nev, lane);
        else if (nmax <= 256) spans_of<K4, kMaxSub, MK>(keys, cnt, evenodd, nev, lane);
        else spans_of<K8, kMaxSub, MK>(keys, cnt, evenodd, nev, lane);
    }
    __syncthreads();
    // ---- accumulate per window, prefix-sum, clamp, write u8 ----
    for (int32_t wx0 = 0; wx0 < job.w; wx0 += kWin) {
        const int32_t ww = min(kWin, job.w - wx0);
        for (int i = lane; i <= ww; i += kWave) acc[i] = 0;
        __syncthreads();
        const int32_t colbase = (job.x0 + wx0) << shift;
        const int32_t colmax = ww << shift;
        for (int s = 0; s < nsub; s++) {
            const int32_t *ev = reinterpret_cast<const int32_t *>(keys[s]);
            const int32_t maxv = job.aa ? (64 - (((s & 3) + 1) >> 2)) : 255;
            for (int k = lane; 2 * k + 1 < nev[s]; k += kWave) {
                int32_t L = ev[2 * k] - colbase, R = ev[2 * k + 1] - colbase;
                if (R <= L) continue;
                L = max(L, 0);
                R = min(R, colmax);
                if (R <= L) continue;
                if (!job.aa) {
                    atomicAdd(&acc[L], 255 << 16);
                    atomicAdd(&acc[R], -(255 << 16));
                    continue;
                }
                const int32_t fb = L & 3, fe = R & 3;
                int32_t c = (R >> 2) - (L >> 2) - 1;
                if (c < 0) {
                    atomicAdd(&acc[L >> 2], (fe - fb) * 16);
                } else {
                    int32_t st = L >> 2;
                    if (fb == 0) c += 1;
                    else { atomicAdd(&acc[st], (4 - fb) * 16); st += 1; }
                    if (c > 0) {
                        atomicAdd(&acc[st], maxv << 16);
                        atomicAdd(&acc[st + c], -(maxv << 16));
                    }
                    if (fe) atomicAdd(&acc[R >> 2], fe * 16);
                }
            }
        }
        __syncthreads();
        int carry = 0;
        for (int32_t base = 0; base < ww; base += 4 * kWave) {
            const int32_t p0 = base + 4 * lane;
            // one 16-byte LDS read per lane (acc is 16-byte aligned; p0 + 3 <= kWin + 3
            // stays inside the padded array, entries past ww are zeroed below)
            const int4 q4 = *reinterpret_cast<const int4 *>(&acc[p0]);
            const int32_t packed4[4] = {q4.x, q4.y, q4.z, q4.w};
            int v[4], d[4], loc = 0;
            for (int q = 0; q < 4; q++) {
                const int32_t packed = p0 + q < ww ? packed4[q] : 0;
                v[q] = packed & 0xffff;
                d[q] = packed >> 16;
                loc += d[q];
            }
            const int incl = wave_incl_scan(loc, lane);
            int run = carry + incl - loc;
            uint32_t word = 0;
            for (int q = 0; q < 4; q++) {
                run += d[q];
                word |= (uint32_t)min(255, v[q] + run) << (8 * q);
            }
            const int32_t gx = wx0 + p0;
            if (gx + 3 < job.w) {
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
__device__ __forceinline__ void shade(const ShaderParams &sp, const ShaderHead &s, int32_t px, int32_t py,
                                      uint32_t o[4]) {
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
        for (int k = 0; k < 4; k++) c[k] = t * sp.f[0][k] + sp.b[0][k];
    } else {
        // the thresholds are read four at a time (t has kMaxStops + 2 entries)
        uint32_t idx = 0;
        for (uint32_t i = 1; i < s.nstops; i += 4) {
            float q[4];
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = sp.t[min(i + (uint32_t)k, (uint32_t)kMaxStops + 1u)];
#pragma unroll
            for (int k = 0; k < 4; k++) idx += i + (uint32_t)k < s.nstops && t >= q[k] ? 1u : 0u;
        }
        for (int k = 0; k < 4; k++) c[k] = t * sp.f[idx][k] + sp.b[idx][k];
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

// one covered pixel of a draw item: Clear, or the shaded colour composited with
// over_lowp (the colour is shaded only when needed)
__device__ __forceinline__ uint32_t paint_px(const ShaderParams &sp, const ShaderHead &sh, int32_t lx, int32_t ly,
                                             uint32_t c, uint32_t d, uint32_t src[4], bool &shaded) {
    if (sh.clear) {
        const uint32_t ic = 255u - c;
        uint32_t o = 0;
        for (int k = 0; k < 4; k++) o |= div255(((d >> (8 * k)) & 0xffu) * ic) << (8 * k);
        return o;
    }
    if (!shaded) {
        shade(sp, sh, lx, ly, src);
        shaded = true;
    }
    return over_lowp(src, c, d, sh.as_source != 0);
}

// coverage of a job at layer pixel (x, y); 0 outside its rect (and for no job, whose
// rect is empty)
__device__ __forceinline__ uint32_t ref_cov(const JobRef &r, const uint8_t *arena, int32_t x, int32_t y) {
    const uint32_t cx = (uint32_t)(x - r.x0), cy = (uint32_t)(y - r.y0);
    return cx < (uint32_t)r.w && cy < (uint32_t)r.h ? arena[r.cov_offset + (size_t)cy * r.stride + cx] : 0u;
}

// Coverage of one hairline blit at pixel (x, y) (see HairLine); 0 when the line does
// not blit the pixel.
__device__ __forceinline__ uint32_t hair_alpha(const HairLine &L, int32_t x, int32_t y) {
    const bool ymaj = (L.flags & rvkr::kHairYMajor) != 0;
    const int32_t major = ymaj ? y : x, minor = ymaj ? x : y;
    if (major < L.istart || major >= L.istop) return 0;
    if ((L.flags & rvkr::kHairClipped) && (x < L.cl || x >= L.cr || y < L.ct || y >= L.cb)) return 0;
    const int32_t k = major - L.istart;
    int32_t mod = 64;
    if (k == 0) mod = L.scale_start;
    else if (L.scale_stop > 0 && major == L.istop - 1) mod = L.scale_stop;
    const int32_t f = (int32_t)((uint32_t)L.fstart + (uint32_t)k * (uint32_t)L.slope + 0x8000u);
    const int32_t lo = f >> 16;
    const int32_t a = (f >> 8) & 0xff;
    if (minor == lo) return (uint32_t)((a * mod) >> 6);
    if (minor == lo - 1) return (uint32_t)(((255 - a) * mod) >> 6);
    return 0;
}

struct DrawArgs {
    const DrawItem *items;
    uint32_t nitems;
    const CovJob *jobs;
    const uint8_t *arena;
    const ShaderParams *shaders;
    const HairJob *hair_jobs;
    const uint32_t *hair_tiles, *hair_bins;
    const HairLine *hair_lines;
};

// One thread per target pixel; each 64-lane wave covers an 8 x 8 pixel square and a
// block of 4 waves a 16 x 16 tile.  Items are culled per block and per wave; a wave
// runs an item's shading only when one of its pixels is covered, and compact squares
// cross path boundaries far less often than 64-pixel row segments do (16 x 4 and
// 4 x 16 waves, with fewer or more cache rows per access, measured no faster).
constexpr int kWaveW = 8, kWaveH = 8;  // pixels of one 64-lane wave
constexpr int TX = 16, TY = 16;         // pixels of one block (four waves)
constexpr int kWavesX = TX / kWaveW;
static_assert(kWaveW * kWaveH == 64 && (TX / kWaveW) * (TY / kWaveH) == 4, "four 64-pixel waves per tile");

// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.  Block b is
// mapped to a tile so that each XCD gets runs of kChunk horizontally adjacent
// row-major tiles (which read the same coverage and pixel cache lines), the runs
// interleaved over the draw rect so the XCDs share the work evenly; tiles past the
// last whole round of runs keep their own index.  Runs of 4 tiles (64 pixels)
// measured fastest on the full-filter document (1, 2, 8, 16 and 32 all slower).
constexpr uint32_t kXcds = 8, kChunk = 4;

__global__ __launch_bounds__(256) void k_draw(uint32_t *px, uint32_t w, uint32_t h, int32_t ox, int32_t oy,
                                              uint32_t tiles_x, DrawArgs a) {
    const uint32_t b = blockIdx.x;
    const uint32_t round = kXcds * kChunk, full = gridDim.x / round * round;
    const uint32_t xcd = b % kXcds, k = b / kXcds;
    const uint32_t tile = b < full ? (k / kChunk) * round + xcd * kChunk + k % kChunk : b;
    const uint32_t bx = tile % tiles_x, by = tile / tiles_x;
    const int lane = (int)(threadIdx.x % 64), wv = (int)(threadIdx.x / 64);
    const int32_t tx0 = ox + (int32_t)(bx * TX), ty0 = oy + (int32_t)(by * TY);
    const int32_t wx0 = tx0 + (wv % kWavesX) * kWaveW, wy0 = ty0 + (wv / kWavesX) * kWaveH;
    const int32_t x = wx0 + lane % kWaveW, y = wy0 + lane / kWaveW;
    const bool inside = x < (int32_t)w && y < (int32_t)h;
    uint32_t d = inside ? px[(size_t)y * w + x] : 0u;
    const uint32_t d0 = d;
    for (uint32_t it = 0; it < a.nitems; it++) {
        const DrawItem item = a.items[it];  // with its jobs' rects and coverage offsets
        const bool hit = item.x1 > tx0 && item.x0 < tx0 + TX && item.y1 > ty0 && item.y0 < ty0 + TY &&
                         item.x1 > wx0 && item.x0 < wx0 + kWaveW && item.y1 > wy0 && item.y0 < wy0 + kWaveH;
        if (hit && inside && x >= item.x0 && x < item.x1 && y >= item.y0 && y < item.y1) {
            const int32_t lx = x - item.lx, ly = y - item.ly;
            const uint32_t cf = ref_cov(item.fj, a.arena, lx, ly);
            uint32_t cs = 0, hb = 0, he = 0;
            if (item.stroke_hair) {
                // the hairline lines binned to this pixel's tile, in drawing order
                const uint32_t cx = (uint32_t)(lx - item.sj.x0), cy = (uint32_t)(ly - item.sj.y0);
                if (cx < (uint32_t)item.sj.w && cy < (uint32_t)item.sj.h) {
                    const uint32_t t = item.sj.tile_offset + cy / rvkr::kHairTile * item.sj.stride +
                                       cx / rvkr::kHairTile;
                    hb = a.hair_tiles[t];
                    he = a.hair_tiles[t + 1];
                }
            } else {
                cs = ref_cov(item.sj, a.arena, lx, ly);
            }
            if ((cf | cs) != 0 || hb != he) {
                uint32_t target = item.layer ? 0u : d;
                for (int pass = 0; pass < 2; pass++) {
                    const bool stroke = (pass == 0) == (item.stroke_first != 0);
                    const ShaderParams &sp = a.shaders[stroke ? item.stroke_shader : item.fill_shader];
                    const ShaderHead &sh = sp;
                    uint32_t src[4];
                    bool shaded = false;
                    if (stroke && item.stroke_hair) {
                        // every blit of the hairline at this pixel, composited in order
                        for (uint32_t k = hb; k < he; k++) {
                            const uint32_t c = hair_alpha(a.hair_lines[a.hair_bins[k]], lx, ly);
                            if (c) target = paint_px(sp, sh, lx, ly, c, target, src, shaded);
                        }
                        continue;
                    }
                    const uint32_t c = stroke ? cs : cf;
                    if (c) target = paint_px(sp, sh, lx, ly, c, target, src, shaded);
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
        }
    }
    if (inside && d != d0) px[(size_t)y * w + x] = d;
}

}  // namespace

namespace rvk {

rvk_status launch_expand(rvk_context *ctx, const EdgeRec *edges, uint32_t nedges, const EdgeCov *cov,
                         PieceDev *pieces, uint32_t *npieces, int32_t *overflow, double algo_bytes) {
    if (nedges == 0) return RVK_OK;
    ExpandArgs a{edges, nedges, cov, pieces, npieces, overflow};
    timed_launch(ctx, "expand", algo_bytes,
                 [&] { hipLaunchKernelGGL(k_expand, dim3(div_up(nedges, 256)), dim3(256), 0, ctx->stream, a); });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

rvk_status launch_coverage(rvk_context *ctx, const CovJob *jobs, uint32_t njobs, const uint2 *groups128,
                           uint32_t n128, const uint2 *groups256, uint32_t n256, const uint2 *groups512,
                           uint32_t n512, const EdgeCov *edges, const uint2 *windows, const PieceDev *pieces,
                           const uint32_t *npieces, uint8_t *arena, int32_t *overflow, double algo_bytes) {
    // the algorithmic bytes are split between the launches by their row-group counts
    const double total = (double)n128 + (double)n256 + (double)n512;
    auto launch = [&](auto kernel, const uint2 *groups, uint32_t n) -> rvk_status {
        if (n == 0) return RVK_OK;
        CovArgs a{groups, jobs, njobs, edges, windows, pieces, npieces, arena, overflow};
        timed_launch(ctx, "coverage", algo_bytes * (double)n / total, [&] {
            hipLaunchKernelGGL(kernel, dim3(n * rvkr::kGroupRows), dim3(kWave), 0, ctx->stream, a);
        });
        RVK_HIP_CHECK(hipGetLastError());
        return RVK_OK;
    };
    RVK_TRY(launch(k_coverage<(int)rvkr::kCoverageKeys[0]>, groups128, n128));
    RVK_TRY(launch(k_coverage<(int)rvkr::kCoverageKeys[1]>, groups256, n256));
    RVK_TRY(launch(k_coverage<(int)rvkr::kCoverageKeys[2]>, groups512, n512));
    return RVK_OK;
}

rvk_status launch_draw_items(rvk_context *ctx, uint8_t *target, uint32_t w, uint32_t h, int32_t x0, int32_t y0,
                             int32_t x1, int32_t y1, const DrawItem *items, uint32_t nitems, const CovJob *jobs,
                             const uint8_t *arena, const ShaderParams *shaders, const HairJob *hair_jobs,
                             const uint32_t *hair_tiles, const uint32_t *hair_bins, const HairLine *hair_lines,
                             double algo_bytes) {
    if (nitems == 0 || x1 <= x0 || y1 <= y0) return RVK_OK;
    const uint32_t tiles_x = div_up((uint32_t)(x1 - x0), TX), tiles_y = div_up((uint32_t)(y1 - y0), TY);
    const dim3 grid(tiles_x * tiles_y);
    const DrawArgs a{items, nitems, jobs, arena, shaders, hair_jobs, hair_tiles, hair_bins, hair_lines};
    timed_launch(ctx, "draw_items", algo_bytes, [&] {
        // four 8 x 8 waves per 16 x 16 tile
        hipLaunchKernelGGL(k_draw, grid, dim3(256), 0, ctx->stream, (uint32_t *)target, w, h, x0, y0, tiles_x, a);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

}  // namespace rvk
