// Device rasterizer: curve expansion, anti-aliased coverage and the fused
// shade/composite of draw lists.
//
// Expansion (k_expand): one thread per edge runs tiny-skia's quadratic / cubic
// update loop (Skia's forward differencing in 16.16 fixed point) from the state the
// host left after the first piece, writing every non-empty line piece in order.
//
// Coverage (k_coverage): one 64-lane wave per (coverage job, band of 8 pixel rows).
// Per row, the band's edge window comes from the job's table; each lane takes an
// edge, binary-searches its pieces for the row and produces the crossings of the
// row's sub-scanlines.  Per sub-scanline the crossings (keys: x, edge id, winding)
// are sorted with a register-resident bitonic network (up to 8 keys per lane), the
// windings are scanned into span events (tiny-skia's walk_edges spans), and the
// SuperBlitter contributions (16 per covered sub-column of a partial pixel, 64 or 63
// for a fully covered pixel) are accumulated per 512-pixel window in LDS,
// prefix-summed and clamped to 255.  Coverage leaves the kernel as 8 x 8 tiles
// aligned to the target: a u16 table entry per tile (empty, full, or the slot of a
// partial tile's 64 bytes), so only tiles that an edge crosses cost memory.
//
// Compositing (k_draw): one thread per target pixel, one wave per 8 x 8 tile,
// looping over the draw items in order with the pixel in a register: tile-table
// entry -> coverage -> gradient or solid shade (tiny-skia lowp stages) -> SourceOver
// into the target (direct items) or into an elided layer pixel that is then
// composited in f32 with the group's opacity and blend mode (draw_pixmap semantics).
#include "rvk_internal.h"
#include "blend_ops.h"
#include "raster.h"

using namespace rvkd;
using rvkr::CovJob;
using rvkr::DrawCull;
using rvkr::DrawItem;
using rvkr::EdgeCov;
using rvkr::EdgeRec;
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
    const uint2 *bands;  // (job, band) of each block of this launch
    const CovJob *jobs;
    const EdgeCov *edges;
    const uint2 *windows;
    const PieceDev *pieces;
    const uint32_t *npieces;
    uint16_t *table;
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
// A column chunk of the row (crossings rounded to [c0, c1) only) starts from the
// backdrop winding base[s] of the crossings left of it: inside a span there, the
// events get a leading start at c0, and a span still open after the last crossing
// gets an end at c1.  The whole row is the chunk with base 0 (closed paths end
// outside).
template <int K, int S, int MK>
__device__ __forceinline__ void spans_of(uint64_t (*keys)[MK], const int *n, bool evenodd, int *nev, int lane,
                                         const int *base, int32_t c0, int32_t c1) {
    static_assert(K * 64 <= MK, "each lane's K keys stay inside its sub-scanline's key row");
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
        int before = base[s] + incl - local;
        const bool in0 = evenodd ? (base[s] & 1) != 0 : base[s] != 0;
        const int total = base[s] + __shfl(incl, kWave - 1, kWave);
        const bool in1 = evenodd ? (total & 1) != 0 : total != 0;
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
        int pos = eincl - nevt + (in0 ? 1 : 0);
        int32_t *ev = reinterpret_cast<int32_t *>(keys[s]);
#pragma unroll
        for (int j = 0; j < K; j++)
            if (evt[j]) {
                const int32_t x = (int32_t)((uint32_t)(v[s][j] >> 32) ^ 0x80000000u);
                ev[pos++] = (x + 0x8000) >> 16;
            }
        const int ne = (in0 ? 1 : 0) + __shfl(eincl, kWave - 1, kWave);
        if (lane == 0 && in0) ev[0] = c0;
        if (lane == 0 && in1) ev[ne] = c1;
        nev[s] = ne + (in1 ? 1 : 0);
    }
}

// tile state of the coverage kernel (u16): 0..255 a tile whose rows so far are all
// that value, kStatePartial + k a partial tile holding slot k
constexpr uint32_t kStatePartial = 256u;
static_assert(rvkr::kMaxTilesX + kStatePartial <= 65536u, "a partial tile's state fits 16 bits");
static_assert(kWin == 8 * kWave, "a lane holds one 8-pixel tile row of a window");

// MK: crossings per sub-scanline the launch can hold (128, 256 or 512; the keys take
// 4, 8 or 16 KB of LDS, so the smaller variants run two to three times the waves per
// CU).  The host sends a band to the smallest variant that holds the number of edges
// whose rows meet the band, an upper bound on any of its sub-scanlines' crossings.
//
// Block b handles band bands[b].y (8 pixel rows) of job bands[b].x.  At the start of
// the band the edges of its window that meet its rows are gathered into registers
// (lane l holds active edges l, l + 64, ...) with the index of the piece each one is
// on; row by row every active edge steps its piece forward and emits its crossings,
// the sub-scanlines are sorted together and turned into spans, and the spans are
// accumulated per 512-pixel window in LDS (each window visits only the spans that
// reach it), prefix-summed and clamped to 255.  Each lane then holds one 8-pixel row
// of a tile, which is tracked as uniform (every pixel so far one value) or partial: a
// tile turns partial at the first row that differs, takes the band's next slot, and
// from then on its rows are written there (the earlier, uniform rows are filled in at
// that point); at the end of the band the tile table gets empty, full or the slot.
template <int MK>
__global__ __launch_bounds__(kWave) void k_coverage(CovArgs a) {
    constexpr int J = MK / kWave;  // active edges per lane
    __shared__ uint64_t keys[kMaxSub][MK];  // crossings; reused for span events
    __shared__ __align__(16) int32_t acc[kWin + 8];  // acc (low 16 bits) + run difference (high 16 bits)
    __shared__ uint16_t tstate[rvkr::kChunkTiles];
    const int lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint2 bj = a.bands[blockIdx.x];
    const CovJob job = a.jobs[bj.x];
    const uint32_t band = bj.y;
    const int shift = job.aa ? 2 : 0;
    const int nsub_aa = 1 << shift;
    const bool evenodd = job.rule == RVK_FILL_EVENODD;
    const int32_t row0 = job.gy0 + (int32_t)band * rvkr::kBandRows;
    const int32_t ry0 = max(row0, job.y0), ry1 = min(row0 + rvkr::kBandRows, job.y0 + job.h);  // covered rows
    uint16_t *tab = a.table + job.table_offset + (size_t)band * job.tiles_x;
    uint8_t *slots = a.arena + job.cov_offset + (size_t)band * job.tiles_x * 64u;
    const int32_t rx0 = job.x0, rx1 = job.x0 + job.w;  // covered columns
    const EdgeCov *edges = a.edges + job.edge_offset;
    const uint32_t *npieces = a.npieces + job.edge_offset;

    for (int i = lane; i < kWin + 8; i += kWave) acc[i] = 0;  // each window clears it behind its read
    // ---- the band's active edges: the window's edges whose rows meet the band ----
    const int32_t sb0 = ry0 * nsub_aa, sb1 = ry1 * nsub_aa - 1;  // covered sub-scanlines
    uint32_t eidx[J], ek[J];  // an active edge (index in the job) and the piece it is on
    int nact = 0;
    const uint2 win = a.windows[job.window_offset + band];
    {
        uint32_t *list = reinterpret_cast<uint32_t *>(&keys[0][0]);
        for (uint32_t base = win.x; base < win.y; base += kWave) {
            const uint32_t i = base + (uint32_t)lane;
            bool meets = false;
            if (i < win.y) {
                const EdgeCov e = edges[i];
                meets = e.first_y <= sb1 && e.ymax >= sb0;
            }
            const uint64_t m = __ballot(meets);
            const int pos = nact + __popcll(m & below);
            if (meets && pos < MK) list[pos] = i;
            nact += __popcll(m);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int e = lane + kWave * j;
            eidx[j] = e < nact && nact <= MK ? list[e] : 0u;
            ek[j] = 0;
        }
        __syncthreads();  // the key area is free again
        // the first piece that reaches the band's first sub-scanline (an edge that
        // starts inside the band starts on its first piece)
#pragma unroll
        for (int j = 0; j < J; j++) {
            if (lane + kWave * j < nact && nact <= MK) {
                const EdgeCov e = edges[eidx[j]];
                if (e.first_y < sb0) {
                    const PieceDev *pc = a.pieces + e.piece_offset;
                    uint32_t lo = 0, hi = npieces[eidx[j]];
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (pc[mid].last_y < sb0) lo = mid + 1;
                        else hi = mid;
                    }
                    ek[j] = lo;
                }
            }
        }
    }

    uint32_t nslots = 0;  // partial tiles of the band so far
    // The band is processed in column chunks of up to kChunkTiles tiles.  Normally a
    // row's every crossing is sorted; when a sub-scanline holds more crossings than the
    // variant's key area, the chunk is redone at half the width, and from then on each
    // chunk keeps only the crossings that round into its columns and starts from the
    // winding of those left of it (down to single tiles).
    uint32_t chunk_tiles = rvkr::kChunkTiles;
    bool cut = false;  // chunks keep only their own crossings
    for (uint32_t ct0 = 0; ct0 < job.tiles_x;) {
        const uint32_t ctn = min(chunk_tiles, job.tiles_x - ct0);
        const int32_t cx0 = job.gx0 + (int32_t)ct0 * rvkr::kTile;  // first column of the chunk
        const int32_t cw = (int32_t)ctn * rvkr::kTile;
        const int32_t c0s = cut ? cx0 * nsub_aa : INT32_MIN, c1s = cut ? (cx0 + cw) * nsub_aa : INT32_MAX;
        const uint32_t nslots0 = nslots;
        bool redo = false;
        // a chunk after the first walks the band's rows again from its start
        uint32_t k0[J];
#pragma unroll
        for (int j = 0; j < J; j++) k0[j] = ek[j];
        for (int r = 0; r < rvkr::kBandRows && !redo; r++) {
            const int32_t y = row0 + r;
            const bool row_in = y >= ry0 && y < ry1;
            int nev[kMaxSub] = {0, 0, 0, 0};
            if (row_in) {
                // ---- the crossings of the row's sub-scanlines ----
                const int32_t sy0 = y * nsub_aa;
                int cnt[kMaxSub] = {0, 0, 0, 0};
                int bd[kMaxSub] = {0, 0, 0, 0};  // winding of the crossings left of the chunk
                // the crossings of edge i on the row's sub-scanlines, stepping its piece
                // index k forward; those in the chunk are compacted into the key areas in
                // lane order, those left of it are summed into bd
                auto gather = [&](bool valid, uint32_t i, uint32_t &k) {
                    uint64_t key[kMaxSub];
                    bool act[kMaxSub] = {false, false, false, false};
                    int left[kMaxSub] = {0, 0, 0, 0};
                    if (valid) {
                        const EdgeCov e = edges[i];
                        const PieceDev *pc = a.pieces + e.piece_offset;
                        const uint32_t n = npieces[i];
                        if (e.first_y <= sy0 + nsub_aa - 1 && e.ymax >= sy0 && k < n) {
                            PieceDev p = pc[k];
                            for (int s = 0; s < nsub_aa; s++) {
                                const int32_t sy = sy0 + s;
                                while (p.last_y < sy && k + 1 < n) p = pc[++k];
                                if (p.first_y <= sy && p.last_y >= sy) {
                                    const int32_t x = (int32_t)((uint32_t)p.x + (uint32_t)(sy - p.first_y) * (uint32_t)p.dx);
                                    const int32_t xr = (x + 0x8000) >> 16;  // the sub-column of its event
                                    if (xr < c0s) {
                                        left[s] = e.winding > 0 ? 1 : -1;
                                    } else if (xr < c1s) {
                                        key[s] = ((uint64_t)((uint32_t)x ^ 0x80000000u) << 32) |
                                                 ((uint64_t)(uint32_t)e.id << 1) | (e.winding > 0 ? 1u : 0u);
                                        act[s] = true;
                                    }
                                }
                            }
                        }
                    }
                    for (int s = 0; s < nsub_aa; s++) {
                        const uint64_t m = __ballot(act[s]);
                        const int pos = cnt[s] + __popcll(m & below);
                        if (act[s] && pos < MK) keys[s][pos] = key[s];
                        cnt[s] += __popcll(m);
                        if (cut) bd[s] += __popcll(__ballot(left[s] > 0)) - __popcll(__ballot(left[s] < 0));
                    }
                };
                if (nact <= MK) {
                    // the band's active edges, each on the piece it reached
#pragma unroll
                    for (int j = 0; j < J; j++) {
                        if (kWave * j >= nact) break;
                        gather(lane + kWave * j < nact, eidx[j], k0[j]);
                    }
                } else {
                    // more edges meet the band than the registers hold (many short edges,
                    // such as the arcs of round joins): every edge of the window, its
                    // piece found by binary search
                    for (uint32_t base = win.x; base < win.y; base += kWave) {
                        const uint32_t i = base + (uint32_t)lane;
                        uint32_t k = 0;
                        if (i < win.y) {
                            const EdgeCov e = edges[i];
                            if (e.first_y < sy0) {
                                const PieceDev *pc = a.pieces + e.piece_offset;
                                uint32_t hi = npieces[i];
                                while (k < hi) {
                                    const uint32_t mid = (k + hi) >> 1;
                                    if (pc[mid].last_y < sy0) k = mid + 1;
                                    else hi = mid;
                                }
                            }
                        }
                        gather(i < win.y, i, k);
                    }
                }
                __syncthreads();
                // ---- sort the sub-scanlines together and turn them into span events ----
                int nmax = 0;
                for (int s = 0; s < nsub_aa; s++) nmax = max(nmax, cnt[s]);
                if (nmax > MK) {
                    if (ctn > 1) {  // too many crossings: this chunk again, narrower
                        redo = true;
                        break;
                    }
                    // more than MK crossings round into one tile's columns: kept
                    // truncated, and the render reports it
                    if (lane == 0) atomicAdd(a.overflow, 1);
                    for (int s = 0; s < nsub_aa; s++) cnt[s] = min(cnt[s], MK);
                    nmax = MK;
                }
                // keys per lane: the smallest network that holds the fullest sub-scanline
                // (cnt is at most MK, so no branch needs more than MK / 64 keys per lane)
                constexpr int K4 = MK >= 256 ? 4 : 2, K8 = MK >= 512 ? 8 : K4;
                static_assert(K8 * kWave <= MK && K4 * kWave <= MK, "a network never reads past the key area");
                // (a cut chunk with no crossings inside can still lie inside a span)
                if (nmax == 0 && !cut) {
                } else if (nsub_aa == 1) {
                    if (nmax <= 64) spans_of<1, 1, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                    else if (nmax <= 128) spans_of<2, 1, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                    else if (nmax <= 256) spans_of<K4, 1, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                    else spans_of<K8, 1, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                } else {
                    if (nmax <= 64) spans_of<1, kMaxSub, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                    else if (nmax <= 128) spans_of<2, kMaxSub, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                    else if (nmax <= 256) spans_of<K4, kMaxSub, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                    else spans_of<K8, kMaxSub, MK>(keys, cnt, evenodd, nev, lane, bd, c0s, c1s);
                }
                __syncthreads();
            }
            const int nsub = row_in ? nsub_aa : 0;
            int kst[kMaxSub] = {0, 0, 0, 0};  // per sub-scanline: the first span not yet behind the windows
            for (int32_t wx0 = 0; wx0 < cw; wx0 += kWin) {
                const int32_t ww = min(kWin, cw - wx0);  // a multiple of kTile
                // ---- the spans that reach this window, cut to it and to the job's
                // columns (coverage is 0 outside them) ----
                const int32_t wabs = cx0 + wx0;
                const int32_t colbase = wabs * nsub_aa;
                const int32_t wend = colbase + ww * nsub_aa;  // first sub-column past the window
                const int32_t lo = max(0, rx0 - wabs) * nsub_aa, hi = max(0, min(ww, rx1 - wabs)) * nsub_aa;
                // lanes 16 s .. 16 s + 15 look at the next 16 spans of sub-scanline s: the
                // spans that start before the window's end are a prefix of them (spans
                // are ordered), as are those that end at or before it (behind every
                // later window); all the window's spans are then accumulated in one
                // pass, one per lane
                int nin[kMaxSub] = {0, 0, 0, 0}, off[kMaxSub + 1] = {0, 0, 0, 0, 0};
                bool more = false;
                if (nsub > 0 && hi > lo) {
                    const int sl = lane >> 4;
                    const int kk = kst[sl] + (lane & 15);
                    const int nspl = sl < nsub ? nev[sl] >> 1 : 0;
                    int32_t La = 0, Ra = 0;
                    if (kk < nspl) {
                        const int32_t *ev = reinterpret_cast<const int32_t *>(keys[sl]);
                        La = ev[2 * kk];
                        Ra = ev[2 * kk + 1];
                    }
                    const uint64_t min_ = __ballot(kk < nspl && La < wend);
                    const uint64_t mdone = __ballot(kk < nspl && Ra <= wend);
                    for (int q = 0; q < kMaxSub; q++) {
                        nin[q] = __popcll((min_ >> (16 * q)) & 0xffffull);
                        more = more || nin[q] == 16;
                        off[q + 1] = off[q] + nin[q];
                    }
                    for (int q = 0; q < kMaxSub; q++) kst[q] += __popcll((mdone >> (16 * q)) & 0xffffull);
                    // the accumulation of one span (window-relative sub-columns)
                    auto add_span = [&](int q, int32_t La2, int32_t Ra2) {
                        const int32_t maxv = job.aa ? (64 - (((q & 3) + 1) >> 2)) : 255;
                        int32_t L = max(La2 - colbase, lo), R = min(Ra2 - colbase, hi);
                        if (R <= L) return;
                        if (!job.aa) {
                            atomicAdd(&acc[L], 255 << 16);
                            atomicAdd(&acc[R], -(255 << 16));
                            return;
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
                    };
                    if (lane < off[kMaxSub]) {
                        const int q = (lane >= off[1]) + (lane >= off[2]) + (lane >= off[3]);
                        const int k = kst[q] - __popcll((mdone >> (16 * q)) & 0xffffull) + (lane - off[q]);
                        const int32_t *ev = reinterpret_cast<const int32_t *>(keys[q]);
                        add_span(q, ev[2 * k], ev[2 * k + 1]);
                    }
                    if (more) {
                        // a sub-scanline with 16 or more spans in this window: the rest of
                        // them, 64 at a time
                        for (int q = 0; q < nsub; q++) {
                            if (nin[q] < 16) continue;
                            const int32_t *ev = reinterpret_cast<const int32_t *>(keys[q]);
                            const int nsp = nev[q] >> 1;
                            const int k0q = kst[q] - __popcll((mdone >> (16 * q)) & 0xffffull) + 16;
                            int extra_done = 0;
                            for (int kb = k0q; kb < nsp; kb += kWave) {
                                const int k = kb + lane;
                                int32_t La2 = 0, Ra2 = 0;
                                if (k < nsp) {
                                    La2 = ev[2 * k];
                                    Ra2 = ev[2 * k + 1];
                                }
                                const bool in = k < nsp && La2 < wend;
                                extra_done += __popcll(__ballot(k < nsp && Ra2 <= wend));
                                if (in) add_span(q, La2, Ra2);
                                if (__ballot(in) != ~0ull) break;
                            }
                            kst[q] += extra_done;
                        }
                    }
                }
                __syncthreads();
                // ---- prefix sum: lane l holds pixels [8 l, 8 l + 8) of the window, one
                // tile row; the window's accumulator is cleared behind the read ----
                const int32_t p0 = 8 * lane;
                const bool valid = p0 < ww;
                const int4 qa = *reinterpret_cast<const int4 *>(&acc[p0]);
                const int4 qb = *reinterpret_cast<const int4 *>(&acc[p0 + 4]);
                *reinterpret_cast<int4 *>(&acc[p0]) = make_int4(0, 0, 0, 0);
                *reinterpret_cast<int4 *>(&acc[p0 + 4]) = make_int4(0, 0, 0, 0);
                if (lane == 0) acc[kWin] = 0;
                const int32_t packed8[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
                int v[8], d[8], loc = 0;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int32_t packed = valid ? packed8[q] : 0;
                    v[q] = packed & 0xffff;
                    d[q] = packed >> 16;
                    loc += d[q];
                }
                const int incl = wave_incl_scan(loc, lane);
                int run = incl - loc;
                uint32_t w0 = 0, w1 = 0;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    run += d[q];
                    const uint32_t c = (uint32_t)min(255, v[q] + run);
                    if (q < 4) w0 |= c << (8 * q);
                    else w1 |= c << (8 * (q - 4));
                }
                // ---- the tile of this row of 8 pixels ----
                const uint32_t t = (uint32_t)wx0 / rvkr::kTile + (uint32_t)lane;  // tile in the chunk
                const uint32_t v0 = w0 & 0xffu;
                const bool tuni = w0 == v0 * 0x01010101u && w1 == w0;
                const uint32_t st = valid && r > 0 ? (uint32_t)tstate[t] : 0u;
                const bool was_partial = r > 0 && st >= kStatePartial;
                const bool fresh = valid && !was_partial && (r == 0 ? !tuni : !(tuni && v0 == st));
                const uint64_t m = __ballot(fresh);
                const uint32_t k = was_partial ? st - kStatePartial : nslots + (uint32_t)__popcll(m & below);
                nslots += (uint32_t)__popcll(m);
                if (valid && (was_partial || fresh)) {
                    uint2 *tp = reinterpret_cast<uint2 *>(slots + (size_t)k * 64u);
                    tp[r] = make_uint2(w0, w1);
                    if (fresh) {  // the tile's earlier rows were all st
                        const uint32_t u = st * 0x01010101u;
                        for (int rr = 0; rr < r; rr++) tp[rr] = make_uint2(u, u);
                        tstate[t] = (uint16_t)(kStatePartial + k);
                    }
                } else if (valid && r == 0) {
                    tstate[t] = (uint16_t)v0;
                }
                __syncthreads();
            }
        }
        if (redo) {
            chunk_tiles = max(ctn / 2, 1u);
            cut = true;
            nslots = nslots0;
            __syncthreads();
            continue;
        }
        // ---- the chunk's tile table entries ----
        for (uint32_t t0 = 0; t0 < ctn; t0 += kWave) {
            const uint32_t t = t0 + (uint32_t)lane;
            const bool valid = t < ctn;
            const uint32_t st = valid ? (uint32_t)tstate[t] : 0u;
            // a uniform tile of another value than 0 or 255 (rows of different sub-row
            // spans, or a partial pixel value everywhere) gets a slot of its own
            const bool fill = valid && st < kStatePartial && st != 0u && st != 255u;
            const uint64_t m = __ballot(fill);
            const uint32_t k = nslots + (uint32_t)__popcll(m & below);
            nslots += (uint32_t)__popcll(m);
            uint32_t e = rvkr::kTileEmpty;
            if (st >= kStatePartial) {
                e = rvkr::kTileSlot0 + (st - kStatePartial);
            } else if (fill) {
                uint4 *tp = reinterpret_cast<uint4 *>(slots + (size_t)k * 64u);
                const uint32_t v4 = st * 0x01010101u;
                for (int q = 0; q < 4; q++) tp[q] = make_uint4(v4, v4, v4, v4);
                e = rvkr::kTileSlot0 + k;
            } else if (st == 255u) {
                e = rvkr::kTileFull;
            }
            if (valid) tab[ct0 + t] = (uint16_t)e;
        }
        __syncthreads();
        ct0 += ctn;
    }
}

// ---------------------------------------------------------------------------
// shading (tiny-skia lowp) and compositing
// ---------------------------------------------------------------------------
// lowp arithmetic on packed pixels, two channels per 24-bit multiply (one full-rate
// VALU operation): channels 0 and 2 sit in bits 0-7 and 16-23 of one word, 1 and 3 in
// bits 8-15 and 24-31 of the other, and each 16-bit field's value (at most
// 255 * 255 + 255) never carries into the next.  div255(v) = (v + 255) >> 8.
constexpr uint32_t kEven = 0x00ff00ffu;

// div255(c_k * m) for every channel k of c (m <= 255)
__device__ __forceinline__ uint32_t scale_px(uint32_t c, uint32_t m) {
    const uint32_t lo = ((__umul24(c & kEven, m) + kEven) >> 8) & kEven;
    const uint32_t hi = (__umul24((c >> 8) & kEven, m) + kEven) & ~kEven;
    return lo | hi;
}

// div255(d_k * (255 - m) + s_k * m) for every channel k (lowp lerp, m <= 255)
__device__ __forceinline__ uint32_t lerp_px(uint32_t d, uint32_t s, uint32_t m) {
    const uint32_t im = 255u - m;
    const uint32_t lo = ((__umul24(d & kEven, im) + __umul24(s & kEven, m) + kEven) >> 8) & kEven;
    const uint32_t hi = (__umul24((d >> 8) & kEven, im) + __umul24((s >> 8) & kEven, m) + kEven) & ~kEven;
    return lo | hi;
}
__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// the spread mode applied to a gradient coordinate (m: the shader's mode word)
__device__ __forceinline__ float spread_t(uint32_t m, float t) {
    const uint32_t spread = (m >> rvkr::kModeSpreadShift) & 3u;
    if (spread == RVK_SPREAD_REFLECT) {
        const float u = t - 1.0f;
        t = clamp01(fabsf(u - 2.0f * floorf(u * 0.5f) - 1.0f));
    } else if (spread == RVK_SPREAD_REPEAT) {
        t = clamp01(t - floorf(t));
    } else if (m & rvkr::kModePadClamp) {
        t = clamp01(t);
    }
    return t;
}

// gradient coordinate t of a linear or concentric radial shader at the pixel centre
// (x, y)
__device__ __forceinline__ float grad_t(const ShaderHead &s, uint32_t m, float x, float y) {
    const float tx = x * s.sx + (y * s.kx + s.tx);
    float t = tx;
    if (m & rvkr::kModeRadial) {
        const float ty = x * s.ky + (y * s.sy + s.ty);
        t = sqrtf(tx * tx + ty * ty);
        if (m & rvkr::kModeRTwist) t = (t - s.r_scale) / s.r_bias;
    }
    return spread_t(m, t);
}

// gradient coordinate t of a two-point conical shader; `degenerate` is set where the
// gradient is undefined (tiny-skia mask_2pt_conical_degenerates: t <= 0 or NaN, with t
// forced to 0)
__device__ __forceinline__ float grad_t_focal(const ShaderHead &s, int32_t px, int32_t py, bool &degenerate) {
    const float x = (float)px + 0.5f, y = (float)py + 0.5f;
    const float tx = x * s.sx + (y * s.kx + s.tx);
    const float ty = x * s.ky + (y * s.sy + s.ty);
    float t;
    if (s.focal == rvkr::kFocalOnCircle) t = tx + ty * ty / tx;
    else if (s.focal == rvkr::kFocalWellBehaved) t = sqrtf(tx * tx + ty * ty) - tx * s.p0;
    else t = sqrtf(tx * tx - ty * ty) - tx * s.p0;
    degenerate = false;
    if (s.focal != rvkr::kFocalWellBehaved) {
        degenerate = t <= 0.0f || t != t;
        if (degenerate) t = 0.0f;
    }
    return spread_t(s.mode, t);
}

// the gradient's colour at t (tiny-skia's gradient / evenly_spaced_2_stop_gradient
// stages, f32, before any clamp)
__device__ __forceinline__ void grad_color(const ShaderParams &sp, uint32_t m, float t, float c[4]) {
    if (m & rvkr::kModeTwoStop) {
        for (int k = 0; k < 4; k++) c[k] = t * sp.f[0][k] + sp.b[0][k];
    } else {
        // the thresholds are read four at a time (t has kMaxStops + 2 entries)
        const uint32_t nstops = m >> rvkr::kModeStopsShift;
        uint32_t idx = 0;
        for (uint32_t i = 1; i < nstops; i += 4) {
            float q[4];
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = sp.t[min(i + (uint32_t)k, (uint32_t)kMaxStops + 1u)];
#pragma unroll
            for (int k = 0; k < 4; k++) idx += i + (uint32_t)k < nstops && t >= q[k] ? 1u : 0u;
        }
        for (int k = 0; k < 4; k++) c[k] = t * sp.f[idx][k] + sp.b[idx][k];
    }
}

// premultiplied u8 source colour of a shader at the pixel centre (x, y), packed
// (tiny-skia lowp)
__device__ __forceinline__ uint32_t shade(const ShaderParams &sp, uint32_t m, float x, float y) {
    if ((m & rvkr::kModeKind) == 1u) return sp.rgba;
    float c[4];
    grad_color(sp, m, grad_t(sp, m, x, y), c);
    uint32_t o = 0;
    for (int k = 0; k < 4; k++) {
        const float v = clamp01(c[k]);  // fmaxf(NaN, 0) = 0
        o |= (uint32_t)(v * 255.0f + 0.5f) << (8 * k);
    }
    // premultiply: every colour channel times alpha / 255 (alpha's own field is empty)
    if (!(m & rvkr::kModeOpaque)) o = scale_px(o & 0x00ffffffu, o >> 24) | (o & 0xff000000u);
    return o;
}

// One texel of a pattern pixmap at pixmap coordinates (u, v): Repeat tiling (none for
// Pad), then gather_ix (clamped to the last texel) and load_8888 (v * (1/255)).
__device__ __forceinline__ F4 tile_texel(const ShaderHead &s, const uint32_t *tile, float u, float v) {
    const float fw = (float)s.tile_w, fh = (float)s.tile_h;
    if (s.spread == RVK_SPREAD_REPEAT) {
        u = u - floorf(u * s.tile_iw) * fw;
        v = v - floorf(v * s.tile_ih) * fh;
    }
    // the largest float below the size (ulp_sub)
    u = fminf(fmaxf(u, 0.0f), __uint_as_float(__float_as_uint(fw) - 1u));
    v = fminf(fmaxf(v, 0.0f), __uint_as_float(__float_as_uint(fh) - 1u));
    const uint32_t p = tile[(uint32_t)(int32_t)v * s.tile_w + (uint32_t)(int32_t)u];
    const float k = 1.0f / 255.0f;
    return F4{(float)(p & 0xffu) * k, (float)((p >> 8) & 0xffu) * k, (float)((p >> 16) & 0xffu) * k,
              (float)(p >> 24) * k};
}

// The premultiplied colour of a pattern shader at pixel (px, py) (tiny-skia's f32
// seed_shader and transform, then repeat (Repeat spread) + gather for nearest
// sampling, bilinear, or bicubic + clamp_0 + clamp_a, then scale_1_float by the
// opacity).
__device__ F4 pattern_color(const ShaderHead &s, int32_t px, int32_t py) {
    const float x = (float)px + 0.5f, y = (float)py + 0.5f;
    const float tx = x * s.sx + (y * s.kx + s.tx);
    const float ty = x * s.ky + (y * s.sy + s.ty);
    const uint32_t *tile = (const uint32_t *)s.tile;
    F4 o;
    if (s.quality == 0) {
        o = tile_texel(s, tile, tx, ty);
    } else if (s.quality == 1) {
        // bilinear: weights 1 - f, f of the fractions of x + 0.5 and y + 0.5 over the
        // 2 x 2 texels from (x - 0.5, y - 0.5); no clamp stages follow
        const float fx = (tx + 0.5f) - floorf(tx + 0.5f), fy = (ty + 0.5f) - floorf(ty + 0.5f);
        o = F4{0.0f, 0.0f, 0.0f, 0.0f};
        float v = ty - 0.5f;
#pragma unroll 1
        for (int j = 0; j < 2; j++) {
            float u = tx - 0.5f;
            const float wyj = j == 0 ? 1.0f - fy : fy;
#pragma unroll 1
            for (int i = 0; i < 2; i++) {
                const F4 c = tile_texel(s, tile, u, v);
                const float wgt = (i == 0 ? 1.0f - fx : fx) * wyj;
                o.r = wgt * c.r + o.r; o.g = wgt * c.g + o.g; o.b = wgt * c.b + o.b; o.a = wgt * c.a + o.a;
                u = u + 1.0f;
            }
            v = v + 1.0f;
        }
    } else {
        // weights far(1 - f), near(1 - f), near(f), far(f) of the fractions of x + 0.5
        // and y + 0.5: near(t) = ((-21/18 t + 27/18) t + 9/18) t + 1/18, far(t) =
        // t^2 (7/18 t - 6/18); texels from (x - 1.5, y - 1.5) in steps of 1
        auto near_w = [](float t) {
            return t * (t * ((-21.0f / 18.0f) * t + 27.0f / 18.0f) + 9.0f / 18.0f) + 1.0f / 18.0f;
        };
        auto far_w = [](float t) { return (t * t) * ((7.0f / 18.0f) * t + -6.0f / 18.0f); };
        const float fx = (tx + 0.5f) - floorf(tx + 0.5f), fy = (ty + 0.5f) - floorf(ty + 0.5f);
        auto weight = [&](int k, float f) {
            return k == 0 ? far_w(1.0f - f) : k == 1 ? near_w(1.0f - f) : k == 2 ? near_w(f) : far_w(f);
        };
        o = F4{0.0f, 0.0f, 0.0f, 0.0f};
        float v = ty - 1.5f;
        // rolled loops with the weights formed in place keep the compositor's register
        // budget (the kernel is sized for the deepest call it makes)
#pragma unroll 1
        for (int j = 0; j < 4; j++) {
            float u = tx - 1.5f;
            const float wyj = weight(j, fy);
#pragma unroll 1
            for (int i = 0; i < 4; i++) {
                const F4 c = tile_texel(s, tile, u, v);
                const float wgt = weight(i, fx) * wyj;
                o.r = wgt * c.r + o.r; o.g = wgt * c.g + o.g; o.b = wgt * c.b + o.b; o.a = wgt * c.a + o.a;
                u = u + 1.0f;
            }
            v = v + 1.0f;
        }
        o.r = fmaxf(o.r, 0.0f); o.g = fmaxf(o.g, 0.0f); o.b = fmaxf(o.b, 0.0f); o.a = fmaxf(o.a, 0.0f);
        o.a = fminf(o.a, 1.0f);
        o.r = fminf(o.r, o.a); o.g = fminf(o.g, o.a); o.b = fminf(o.b, o.a);
    }
    if (s.pattern_opacity != 1.0f) {
        o.r *= s.pattern_opacity; o.g *= s.pattern_opacity; o.b *= s.pattern_opacity; o.a *= s.pattern_opacity;
    }
    return o;
}

// A covered pixel under a shader of tiny-skia's f32 pipeline (a two-point conical
// gradient: gradient, premultiply, apply_vector_mask; or a pattern), then
// scale_1_float by c / 255, SourceOver onto d, store (kept out of line: the
// compositor's common paths do not carry it)
template <bool kPatterns>
__device__ __noinline__ uint32_t paint_px_hp(const ShaderParams &sp, const ShaderHead &s, int32_t px, int32_t py,
                                             uint32_t c, uint32_t d) {
    F4 src;
    if (kPatterns && s.kind == 3) {
        src = pattern_color(s, px, py);
    } else {
        bool degenerate;
        const float t = grad_t_focal(s, px, py, degenerate);
        float col[4];
        grad_color(sp, s.mode, t, col);
        if (!s.opaque) { col[0] *= col[3]; col[1] *= col[3]; col[2] *= col[3]; }
        src = degenerate ? F4{0.0f, 0.0f, 0.0f, 0.0f} : F4{col[0], col[1], col[2], col[3]};
    }
    if (c != 255u) {
        const float cf = (float)c * (1.0f / 255.0f);
        src.r *= cf; src.g *= cf; src.b *= cf; src.a *= cf;
    }
    return store_px(blend(RVK_BLEND_SOURCE_OVER, src, load_px(d)));
}

// lowp: scale by coverage (unless full), then SourceOver onto `d`; an opaque shader
// composites with Source instead (full coverage stores, partial coverage lerps).  A
// premultiplied colour never exceeds its alpha, so SourceOver's per-channel sums stay
// within 255.
__device__ __forceinline__ uint32_t over_lowp(uint32_t src, uint32_t cov, uint32_t d, bool as_source) {
    if (as_source) return cov == 255 ? src : lerp_px(d, src, cov);
    if (cov != 255) src = scale_px(src, cov);
    return src + scale_px(d, 255u - (src >> 24));
}

// one covered pixel (lx, ly), centre (cx, cy), of a draw item: Clear, or the shaded
// colour composited with over_lowp, or for a shader of the f32 pipeline (two-point
// conical gradients, patterns) in f32 (scale_1_float by c / 255, SourceOver, store);
// the colour is shaded only when needed
template <bool kPatterns>
__device__ __forceinline__ uint32_t paint_px(const ShaderParams &sp, const ShaderHead &sh, int32_t lx, int32_t ly,
                                             float cx, float cy, uint32_t c, uint32_t d, uint32_t &src,
                                             bool &shaded) {
    const uint32_t m = sh.mode;
    if (m & rvkr::kModeClear) return scale_px(d, 255u - c);
    if (m & rvkr::kModeF32) return paint_px_hp<kPatterns>(sp, sh, lx, ly, c, d);
    if (!shaded) {
        src = shade(sp, m, cx, cy);
        shaded = true;
    }
    return over_lowp(src, c, d, (m & rvkr::kModeAsSource) != 0);
}

// tile-table entry of target tile (tx, ty) in a job's grid (origin gx0, gy0, gw x gh
// tiles); empty outside the grid (and for no job, whose grid is 0 x 0)
__device__ __forceinline__ uint32_t tile_entry(const uint16_t *table, int32_t gx0, int32_t gy0, uint32_t gw,
                                               uint32_t gh, uint32_t first, int32_t tx, int32_t ty) {
    const uint32_t cx = (uint32_t)(tx - gx0), cy = (uint32_t)(ty - gy0);
    return cx < gw && cy < gh ? (uint32_t)table[first + cy * gw + cx] : rvkr::kTileEmpty;
}

// coverage of a pixel of target tile row ty from its tile-table entry e (`pix` is the
// pixel's index within its 8 x 8 tile)
__device__ __forceinline__ uint32_t tile_cov(const JobRef &r, const uint8_t *arena, uint32_t e, int32_t ty,
                                             uint32_t pix) {
    if (e == rvkr::kTileEmpty) return 0u;
    if (e == rvkr::kTileFull) return 255u;
    const size_t slot = (size_t)(uint32_t)(ty - r.y0) * r.w + (e - rvkr::kTileSlot0);
    return arena[r.cov_offset + slot * 64u + pix];
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
    const DrawCull *cull;
    uint32_t nitems;
    const uint16_t *table;
    const uint8_t *arena;
    const ShaderParams *shaders;
    const uint32_t *hair_tiles, *hair_bins;
    const HairLine *hair_lines;
};

// One thread per target pixel; each 64-lane wave covers one 8 x 8 coverage tile and a
// block of 4 waves a 16 x 16 square.  A wave first culls a run of 64 items, one lane
// per item (its tile rect and its jobs' tile-table entries for the wave's tile, loaded
// side by side), then composites the items that touch its tile in order: an empty
// tile costs nothing, a full one no coverage load, a partial one one 64-byte load.
constexpr int kWaveW = rvkr::kTile, kWaveH = rvkr::kTile;  // pixels of one 64-lane wave
constexpr int TX = 16, TY = 16;                            // pixels of one block (four waves)
constexpr int kWavesX = TX / kWaveW;
static_assert(kWaveW * kWaveH == 64 && (TX / kWaveW) * (TY / kWaveH) == 4, "four 64-pixel waves per tile");

// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.  Block b is
// mapped to a square so that each XCD gets runs of kChunk horizontally adjacent
// row-major squares (which read neighbouring coverage and pixel cache lines), the
// runs interleaved over the draw rect so the XCDs share the work evenly; squares past
// the last whole round of runs keep their own index.
constexpr uint32_t kXcds = 8, kChunk = 4;

// kPatterns: the draw list holds pattern shaders.  The pattern sampler makes the
// out-of-line f32 path (and with it the whole kernel, which is sized for its deepest
// call) need more registers, so lists without patterns use the variant without it.
template <bool kPatterns>
__global__ __launch_bounds__(256) void k_draw(uint32_t *px, uint32_t w, uint32_t h, int32_t otx, int32_t oty,
                                              uint32_t blocks_x, DrawArgs a) {
    const uint32_t b = blockIdx.x;
    const uint32_t round = kXcds * kChunk, full = gridDim.x / round * round;
    const uint32_t xcd = b % kXcds, k = b / kXcds;
    const uint32_t sq = b < full ? (k / kChunk) * round + xcd * kChunk + k % kChunk : b;
    const uint32_t bx = sq % blocks_x, by = sq / blocks_x;
    const int lane = (int)(threadIdx.x % 64);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));  // uniform: tile math stays scalar
    // this wave's target tile and pixel
    const int32_t ttx = otx + (int32_t)bx * kWavesX + wv % kWavesX, tty = oty + (int32_t)by * (TY / kWaveH) + wv / kWavesX;
    const int32_t x = ttx * kWaveW + lane % kWaveW, y = tty * kWaveH + lane / kWaveW;
    const uint32_t pix = (uint32_t)lane;  // = (y % 8) * 8 + x % 8
    const float fx = (float)x + 0.5f, fy = (float)y + 0.5f;
    const bool inside = x < (int32_t)w && y < (int32_t)h;
    uint32_t d = inside ? px[(size_t)y * w + x] : 0u;
    const uint32_t d0 = d;
    for (uint32_t base = 0; base < a.nitems; base += kWave) {
        // ---- cull 64 items, one per lane ----
        const uint32_t i = base + (uint32_t)lane;
        uint32_t ef = 0, es = 0;
        bool keep = false;
        if (i < a.nitems) {
            const DrawCull c = a.cull[i];
            if (ttx >= c.tx0 && ttx < c.tx1 && tty >= c.ty0 && tty < c.ty1) {
                ef = tile_entry(a.table, c.fx0, c.fy0, c.fw, c.fh, c.ftable, ttx, tty);
                es = tile_entry(a.table, c.sx0, c.sy0, c.sw, c.sh, c.stable, ttx, tty);
                keep = ef != rvkr::kTileEmpty || es != rvkr::kTileEmpty || c.hair != 0;
            }
        }
        // ---- composite the kept items in order ----
        for (uint64_t m = __ballot(keep); m != 0; m &= m - 1) {
            const int kl = __builtin_ctzll(m);
            const uint32_t efk = (uint32_t)__builtin_amdgcn_readlane((int)ef, kl);
            const uint32_t esk = (uint32_t)__builtin_amdgcn_readlane((int)es, kl);
            const DrawItem item = a.items[base + (uint32_t)kl];  // with its jobs' grids and tables
            if (!inside) continue;
            const int32_t lx = x - item.lx, ly = y - item.ly;
            // the pixel centre in layer coordinates, exact in f32 (|x|, |lx| < 2^23)
            const float cx = fx - (float)item.lx, cy = fy - (float)item.ly;
            const uint32_t cf = tile_cov(item.fj, a.arena, efk, tty, pix);
            uint32_t cs = 0, hb = 0, he = 0;
            if (item.stroke_hair) {
                // the hairline lines binned to this pixel's tile, in drawing order
                const uint32_t cx = (uint32_t)(lx - item.sj.x0), cy = (uint32_t)(ly - item.sj.y0);
                if (cx < item.sj.w && cy < item.sj.h) {
                    const uint32_t t = item.sj.table + cy / rvkr::kHairTile * item.sj.stride + cx / rvkr::kHairTile;
                    hb = a.hair_tiles[t];
                    he = a.hair_tiles[t + 1];
                }
            } else {
                cs = tile_cov(item.sj, a.arena, esk, tty, pix);
            }
            if ((cf | cs) != 0 || hb != he) {
                uint32_t target = item.layer ? 0u : d;
                for (int pass = 0; pass < 2; pass++) {
                    const bool stroke = (pass == 0) == (item.stroke_first != 0);
                    const ShaderParams &sp = a.shaders[stroke ? item.stroke_shader : item.fill_shader];
                    const ShaderHead &sh = sp;
                    uint32_t src = 0;
                    bool shaded = false;
                    if (stroke && item.stroke_hair) {
                        // every blit of the hairline at this pixel, composited in order
                        for (uint32_t kk = hb; kk < he; kk++) {
                            const uint32_t c = hair_alpha(a.hair_lines[a.hair_bins[kk]], lx, ly);
                            if (c) target = paint_px<kPatterns>(sp, sh, lx, ly, cx, cy, c, target, src, shaded);
                        }
                        continue;
                    }
                    const uint32_t c = stroke ? cs : cf;
                    if (c) target = paint_px<kPatterns>(sp, sh, lx, ly, cx, cy, c, target, src, shaded);
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

rvk_status launch_coverage(rvk_context *ctx, const CovJob *jobs, uint32_t njobs, const uint2 *bands128,
                           uint32_t n128, const uint2 *bands256, uint32_t n256, const uint2 *bands512,
                           uint32_t n512, const EdgeCov *edges, const uint2 *windows, const PieceDev *pieces,
                           const uint32_t *npieces, uint16_t *table, uint8_t *arena, int32_t *overflow,
                           double algo_bytes) {
    (void)njobs;
    // the algorithmic bytes are split between the launches by their band counts
    const double total = (double)n128 + (double)n256 + (double)n512;
    auto launch = [&](auto kernel, const uint2 *bands, uint32_t n) -> rvk_status {
        if (n == 0) return RVK_OK;
        CovArgs a{bands, jobs, edges, windows, pieces, npieces, table, arena, overflow};
        timed_launch(ctx, "coverage", algo_bytes * (double)n / total,
                     [&] { hipLaunchKernelGGL(kernel, dim3(n), dim3(kWave), 0, ctx->stream, a); });
        RVK_HIP_CHECK(hipGetLastError());
        return RVK_OK;
    };
    RVK_TRY(launch(k_coverage<(int)rvkr::kCoverageKeys[0]>, bands128, n128));
    RVK_TRY(launch(k_coverage<(int)rvkr::kCoverageKeys[1]>, bands256, n256));
    RVK_TRY(launch(k_coverage<(int)rvkr::kCoverageKeys[2]>, bands512, n512));
    return RVK_OK;
}

rvk_status launch_draw_items(rvk_context *ctx, uint8_t *target, uint32_t w, uint32_t h, int32_t x0, int32_t y0,
                             int32_t x1, int32_t y1, const DrawItem *items, const DrawCull *cull, uint32_t nitems,
                             const uint16_t *table, const uint8_t *arena, const ShaderParams *shaders,
                             const uint32_t *hair_tiles, const uint32_t *hair_bins, const HairLine *hair_lines,
                             bool patterns, double algo_bytes) {
    if (nitems == 0 || x1 <= x0 || y1 <= y0 || x0 < 0 || y0 < 0) return RVK_OK;
    // the draw rect in whole coverage tiles (the tile grids are aligned to the target)
    const int32_t otx = x0 / rvkr::kTile, oty = y0 / rvkr::kTile;
    const uint32_t ntx = (uint32_t)(div_up((uint32_t)x1, (uint32_t)rvkr::kTile) - (uint32_t)otx);
    const uint32_t nty = (uint32_t)(div_up((uint32_t)y1, (uint32_t)rvkr::kTile) - (uint32_t)oty);
    const uint32_t blocks_x = div_up(ntx, (uint32_t)kWavesX), blocks_y = div_up(nty, (uint32_t)(TY / kWaveH));
    const DrawArgs a{items, cull, nitems, table, arena, shaders, hair_tiles, hair_bins, hair_lines};
    timed_launch(ctx, "draw_items", algo_bytes, [&] {
        if (patterns)
            hipLaunchKernelGGL(k_draw<true>, dim3(blocks_x * blocks_y), dim3(256), 0, ctx->stream, (uint32_t *)target,
                               w, h, otx, oty, blocks_x, a);
        else
            hipLaunchKernelGGL(k_draw<false>, dim3(blocks_x * blocks_y), dim3(256), 0, ctx->stream, (uint32_t *)target,
                               w, h, otx, oty, blocks_x, a);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

}  // namespace rvk
