// IIR Gaussian blur (crates/resvg/src/filter/iir_blur.rs).  The reference runs, per
// channel, 4 steps of a causal then an anti-causal first-order recurrence along every
// row and then along every column, in f64, and truncates (v*255) to u8.  Every value
// here is the one the reference's sequential loop produces, bit for bit.
//
// Chains of up to kChainMax values run on chip (k_iir_chains): one wave per (chain,
// channel), the chain in LDS and each lane's chunk of it in registers.  A sweep is a
// sequential recurrence y[x] = b[x] + dnu * y[x-1] (product, then sum; never fused),
// so a lane cannot know its chunk's incoming value before the lanes to its left have
// finished.  Each lane therefore first speculates: it runs the recurrence over the K
// values left of its chunk starting from b itself (|dnu| <= 0.27 for the sigmas that
// reach this blur, so the start's influence decays below an f64 ulp within K steps
// wherever the values are not themselves vanishing), and computes its chunk from
// that.  Then the wave iterates to a fixpoint: every lane compares, bit for bit, the
// incoming value it used with the last value of its left neighbour, and the lanes
// where they differ recompute their chunk from the neighbour's value.  The first lane
// starts from the chain's true first value, so after r rounds lanes 0..r are exact
// and the loop ends within 64 rounds; at the fixpoint every lane's incoming value is
// its neighbour's outgoing one, which makes the chunks the sequential recurrence.
// Speculation fails mostly in decaying tails next to transparent runs, where the
// values shrink as fast as the error does; those are the rounds the loop pays for.
//
// Longer chains use the streaming form: one lane per (row or column, channel) running
// its chain sequentially over four f64 planes in HBM, laid out so that lanes of a wave
// touch consecutive addresses at every step (column-major for the row sweeps,
// row-major for the column sweeps).
#include "rvk_internal.h"
#include "device_math.h"

using namespace rvkd;

namespace {

constexpr int T = 32;  // transpose tile

// u8 RGBA (row-major) -> four f64 planes, column-major: plane[c][x*h + y] = px/255.
__global__ void k_to_f64_colmajor(const uint32_t *px, double *planes, uint32_t w, uint32_t h) {
    __shared__ uint32_t tile[T][T + 1];
    uint32_t bx = blockIdx.x * T, by = blockIdx.y * T;
    for (int j = threadIdx.y; j < T; j += blockDim.y) {
        uint32_t x = bx + threadIdx.x, y = by + j;
        tile[j][threadIdx.x] = (x < w && y < h) ? px[(size_t)y * w + x] : 0u;
    }
    __syncthreads();
    size_t plane = (size_t)w * h;
    for (int j = threadIdx.y; j < T; j += blockDim.y) {
        uint32_t x = bx + j, y = by + threadIdx.x;  // write along y (contiguous in column-major)
        if (x < w && y < h) {
            uint32_t p = tile[threadIdx.x][j];
            for (int c = 0; c < 4; c++) planes[c * plane + (size_t)x * h + y] = (double)ch(p, c) / 255.0;
        }
    }
}

// The eight sweeps of one chain b[0], b[s], ..., b[(n-1) s] in the reference's order
// of f64 operations: per step, the product dnu * previous, then the sum (no fused
// multiply-add).  Values ahead of the sweep are not written by it, so each lane loads
// the next kBlock values while it computes the current kBlock: the recurrence then
// waits for memory once per block instead of once per step.
constexpr int kBlock = 32;
__device__ __forceinline__ void iir_chain(double *b, size_t s, uint32_t n, double dnu) {
    if (n < 2) return;
    double cur[kBlock], nxt[kBlock];
    for (int step = 0; step < 4; step++) {
        // causal: b[x] += dnu * b[x - 1] for x = 1 .. n-1
        double prev = b[0];
#pragma unroll
        for (int k = 0; k < kBlock; k++) cur[k] = 1 + k < (int)n ? b[(size_t)(1 + k) * s] : 0.0;
        for (uint32_t x0 = 1; x0 < n; x0 += kBlock) {
            const uint32_t x1 = x0 + kBlock;
#pragma unroll
            for (int k = 0; k < kBlock; k++) nxt[k] = x1 + k < n ? b[(size_t)(x1 + k) * s] : 0.0;
#pragma unroll
            for (int k = 0; k < kBlock; k++) {
                if (x0 + k < n) {
                    const double v = cur[k] + dnu * prev;
                    b[(size_t)(x0 + k) * s] = v;
                    prev = v;
                }
            }
#pragma unroll
            for (int k = 0; k < kBlock; k++) cur[k] = nxt[k];
        }
        // anti-causal: b[x - 1] += dnu * b[x] for x = n-1 .. 1
        prev = b[(size_t)(n - 1) * s];
#pragma unroll
        for (int k = 0; k < kBlock; k++) cur[k] = (int)n - 2 - k >= 0 ? b[(size_t)(n - 2 - k) * s] : 0.0;
        for (int64_t x0 = (int64_t)n - 2; x0 >= 0; x0 -= kBlock) {
            const int64_t x1 = x0 - kBlock;
#pragma unroll
            for (int k = 0; k < kBlock; k++) nxt[k] = x1 - k >= 0 ? b[(size_t)(x1 - k) * s] : 0.0;
#pragma unroll
            for (int k = 0; k < kBlock; k++) {
                if (x0 - k >= 0) {
                    const double v = cur[k] + dnu * prev;
                    b[(size_t)(x0 - k) * s] = v;
                    prev = v;
                }
            }
#pragma unroll
            for (int k = 0; k < kBlock; k++) cur[k] = nxt[k];
        }
    }
}

// Row sweeps on column-major planes: lane = (row y, channel c).
__global__ __launch_bounds__(64) void k_iir_rows(double *planes, uint32_t w, uint32_t h, double dnu) {
    uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = blockIdx.y;
    if (y >= h) return;
    iir_chain(planes + (size_t)c * w * h + y, h, w, dnu);
}

// column-major planes -> row-major planes
__global__ void k_transpose_planes(const double *src, double *dst, uint32_t w, uint32_t h) {
    __shared__ double tile[T][T + 1];
    uint32_t c = blockIdx.z;
    size_t plane = (size_t)w * h;
    const double *s = src + c * plane;
    double *d = dst + c * plane;
    uint32_t bx = blockIdx.x * T, by = blockIdx.y * T;  // x tile, y tile
    for (int j = threadIdx.y; j < T; j += blockDim.y) {
        uint32_t x = bx + j, y = by + threadIdx.x;  // read contiguous in y
        if (x < w && y < h) tile[j][threadIdx.x] = s[(size_t)x * h + y];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < T; j += blockDim.y) {
        uint32_t y = by + j, x = bx + threadIdx.x;  // write contiguous in x
        if (x < w && y < h) d[(size_t)y * w + x] = tile[threadIdx.x][j];
    }
}

// Column sweeps on row-major planes: lane = (column x, channel c).
__global__ __launch_bounds__(64) void k_iir_cols(double *planes, uint32_t w, uint32_t h, double dnu) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = blockIdx.y;
    if (x >= w) return;
    iir_chain(planes + (size_t)c * w * h + x, w, h, dnu);
}

// scale + `(v * 255.0) as u8`, from row-major planes
__global__ void k_from_f64_rowmajor(const double *planes, uint32_t *px, size_t n, double post) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t o = 0;
        for (int c = 0; c < 4; c++) {
            double v = planes[c * n + i] * post;
            o |= as_u8d(v * 255.0) << (8 * c);
        }
        px[i] = o;
    }
}

// scale + cast from column-major planes (no column pass ran)
__global__ void k_from_f64_colmajor(const double *planes, uint32_t *px, uint32_t w, uint32_t h, double post) {
    __shared__ uint32_t tile[T][T + 1];
    uint32_t bx = blockIdx.x * T, by = blockIdx.y * T;
    size_t plane = (size_t)w * h;
    for (int j = threadIdx.y; j < T; j += blockDim.y) {
        uint32_t x = bx + j, y = by + threadIdx.x;
        uint32_t o = 0;
        if (x < w && y < h)
            for (int c = 0; c < 4; c++) {
                double v = planes[c * plane + (size_t)x * h + y] * post;
                o |= as_u8d(v * 255.0) << (8 * c);
            }
        tile[threadIdx.x][j] = o;
    }
    __syncthreads();
    for (int j = threadIdx.y; j < T; j += blockDim.y) {
        uint32_t y = by + j, x = bx + threadIdx.x;
        if (x < w && y < h) px[(size_t)y * w + x] = tile[j][threadIdx.x];
    }
}

// ---- on-chip chains ----

constexpr uint32_t kChainMax = 4096;  // 64 lanes x 64 values

__device__ __forceinline__ uint64_t bits_of(double v) { return (uint64_t)__double_as_longlong(v); }

// v / 255.0 for every byte, folded (correctly rounded) at compile time
struct Inv255 {
    double v[256];
    constexpr Inv255() : v() {
        for (int i = 0; i < 256; i++) v[i] = (double)i / 255.0;
    }
};
__device__ constexpr Inv255 kInv255{};

// LDS layout of a chain: chunks of CT values at a stride of P = CT | 1 doubles (the
// 64 lanes' t-th values fall in distinct bank pairs), lane j's chunk at chunk A + j,
// with A chunks of zeros before the chain and after its 64 chunks, so that warm-ups
// read past either end without a bounds check.  A chunk's offset from another is a
// compile-time constant wherever the sweeps use one.
template <int CT, int K>
struct ChainLayout {
    static constexpr int P = CT | 1;
    static constexpr int A = (K + CT - 1) / CT;
    static constexpr int kDoubles = (64 + 2 * A) * P;
    // position, relative to a lane's chunk start, of the element rel values after it
    // (rel may be negative)
    static constexpr int rel_pos(int rel) {
        const int d = rel >= 0 ? rel / CT : -((-rel + CT - 1) / CT);
        return d * P + (rel - d * CT);
    }
    __device__ static uint32_t pos(uint32_t e) { return (A + e / CT) * P + e % CT; }
};

// One sweep rightwards (y[0] = b[0]; y[x] = b[x] + dnu * y[x-1]) or leftwards
// (y[n-1] = b[n-1]; y[x-1] = b[x-1] + dnu * y[x]) of a chain whose lane j holds
// positions x0 = j * CT .. x0 + CT - 1 in r (0 past the chain's end).  Values are
// never negative, so a recurrence started from y = 0 over zeros stays +0 and its
// first real step gives b exactly: the chain's ends need no special case, and a
// lane whose warm-up reaches past an end computes its true incoming value.
template <int CT, int K, bool kRight>
__device__ __forceinline__ void chain_sweep(double *buf, uint32_t pb, double (&r)[CT], int n, int x0, double dnu) {
    using Lay = ChainLayout<CT, K>;
    __syncthreads();  // every read of the previous sweep's values has been made
    double *mine = buf + pb;
#pragma unroll
    for (int t = 0; t < CT; t++) mine[t] = r[t];
    __syncthreads();
    // warm-up over the K values before the chunk in the sweep's direction (eight LDS
    // loads in flight ahead of eight dependent steps)
    double y = 0.0;
#pragma unroll
    for (int i = 0; i < K; i += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            v[u] = mine[Lay::rel_pos(kRight ? i + u - K : CT + K - 1 - i - u)];
#pragma unroll
        for (int u = 0; u < 8; u++) y = v[u] + dnu * y;
    }
    // exact: the warm-up reached the chain's start, or the lane lies past the end of
    // a rightward chain (its values are discarded)
    const bool exact = kRight ? (x0 <= K || x0 >= n) : x0 + CT + K >= n;
    double yin = y;
    if (kRight) {
#pragma unroll
        for (int t = 0; t < CT; t++) {
            y = r[t] + dnu * y;
            r[t] = y;
        }
    } else {
#pragma unroll
        for (int t = CT - 1; t >= 0; t--) {
            y = r[t] + dnu * y;
            r[t] = y;
        }
    }
    // fixpoint: recompute the chunks whose incoming value differs from the
    // neighbour's outgoing one (exact lanes never do)
    for (;;) {
        const double nin = kRight ? __shfl_up(y, 1, 64) : __shfl_down(y, 1, 64);
        const bool redo = !exact && bits_of(nin) != bits_of(yin);
        if (__ballot(redo) == 0ull) break;
        if (redo) {
            yin = nin;
            y = yin;
            if (kRight) {
#pragma unroll
                for (int t = 0; t < CT; t++) {
                    y = mine[t] + dnu * y;
                    r[t] = y;
                }
            } else {
#pragma unroll
                for (int t = CT - 1; t >= 0; t--) {
                    y = mine[t] + dnu * y;
                    r[t] = y;
                }
            }
        }
    }
    // a rightward sweep carries values past the chain's end: put the zeros back
    if (kRight && x0 + CT > n) {
#pragma unroll
        for (int t = 0; t < CT; t++)
            if (x0 + t >= n) r[t] = 0.0;
    }
}

// One wave per (chain, channel).  Element e of chain k is pixel k * cstride +
// e * estride: rows are (cstride w, estride 1), columns (1, w).  Input is the u8
// channel (as v / 255) or an f64 plane; output an f64 plane or the u8 channel of
// (v * post) * 255 truncated, either into the pixmap or (planar) into a per-channel
// plane in chain order, written contiguously (byte stores scattered over columns of
// the pixmap leave partly written cache lines that go to memory many times over).
// Workgroups are numbered so that each XCD runs a contiguous range of chains
// (neighbouring columns share cache lines).
template <int CT, int K>
__global__ __launch_bounds__(64) void k_iir_chains(const uint8_t *in_u8, const double *in_f64, uint8_t *out_u8,
                                                   double *out_f64, uint32_t n, uint32_t nchains, uint32_t estride,
                                                   uint32_t cstride, uint32_t plane, double dnu, int steps,
                                                   double post, int planar) {
    using Lay = ChainLayout<CT, K>;
    __shared__ double buf[Lay::kDoubles];
    constexpr int L = 64;
    const int j = (int)threadIdx.x;
    const uint32_t N = nchains * 4u, G = blockIdx.x;
    const uint32_t q = N / 8u, rem = N % 8u, xcd = G % 8u;
    const uint32_t id = xcd * q + min(xcd, rem) + G / 8u;
    const uint32_t k = id >> 2, c = id & 3u;
    // 32-bit element indices: the on-chip form takes sides of at most kChainMax, so
    // 4 planes of w * h values stay below 2^26
    const uint32_t base = k * cstride;
    double r[CT];
    // the zero chunks before and after the chain
    for (int i = j; i < Lay::A * Lay::P; i += L) {
        buf[i] = 0.0;
        buf[(Lay::A + L) * Lay::P + i] = 0.0;
    }
    // the chain in lane-interleaved order (element j + L t), every load in flight at
    // once (indices clamped into the chain, so no load sits behind a branch), then
    // through LDS into each lane's chunk
    if (in_f64) {
        const uint32_t cb = c * plane + base;
#pragma unroll
        for (int t = 0; t < CT; t++) {
            const uint32_t e = min((uint32_t)(j + L * t), n - 1u);
            r[t] = in_f64[cb + e * estride];
        }
    } else {
        uint32_t raw[CT];
#pragma unroll
        for (int t = 0; t < CT; t++) {
            const uint32_t e = min((uint32_t)(j + L * t), n - 1u);
            raw[t] = in_u8[(base + e * estride) * 4u + c];
        }
#pragma unroll
        for (int t = 0; t < CT; t++) r[t] = kInv255.v[raw[t]];
    }
#pragma unroll
    for (int t = 0; t < CT; t++) {
        const uint32_t e = (uint32_t)(j + L * t);
        buf[Lay::pos(e)] = e < n ? r[t] : 0.0;
    }
    __syncthreads();
    const int x0 = j * CT;
    const uint32_t pb = (uint32_t)(Lay::A + j) * (uint32_t)Lay::P;
#pragma unroll
    for (int t = 0; t < CT; t++) r[t] = buf[pb + t];
    for (int s = 0; s < steps; s++) {
        chain_sweep<CT, K, true>(buf, pb, r, (int)n, x0, dnu);
        chain_sweep<CT, K, false>(buf, pb, r, (int)n, x0, dnu);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < CT; t++) buf[pb + t] = r[t];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < CT; t++) r[t] = buf[Lay::pos((uint32_t)(j + L * t))];
#pragma unroll
    for (int t = 0; t < CT; t++) {
        const uint32_t e = (uint32_t)(j + L * t);
        if (e < n) {
            const uint32_t p = base + e * estride;
            if (out_f64) out_f64[c * plane + p] = r[t];
            else if (planar) out_u8[c * plane + k * n + e] = (uint8_t)as_u8d(r[t] * post * 255.0);
            else out_u8[p * 4u + c] = (uint8_t)as_u8d(r[t] * post * 255.0);
        }
    }
}

// Per-channel u8 planes in column order (plane c, column x, row y at c * w * h +
// x * h + y) -> the RGBA pixmap, through a 32 x 32 pixel tile in LDS.
__global__ __launch_bounds__(256) void k_planes_to_rgba(const uint8_t *planes, uint32_t *px, uint32_t w, uint32_t h) {
    __shared__ uint8_t tile[4][32][33];
    const uint32_t x0 = blockIdx.x * 32u, y0 = blockIdx.y * 32u, tx = threadIdx.x, ty = threadIdx.y;
    const size_t plane = (size_t)w * h;
    for (uint32_t j = ty; j < 32u; j += 8u) {
        const uint32_t x = x0 + j, y = y0 + tx;
        if (x < w && y < h)
            for (int c = 0; c < 4; c++) tile[c][j][tx] = planes[c * plane + (size_t)x * h + y];
    }
    __syncthreads();
    for (uint32_t j = ty; j < 32u; j += 8u) {
        const uint32_t y = y0 + j, x = x0 + tx;
        if (x < w && y < h)
            px[(size_t)y * w + x] = (uint32_t)tile[0][tx][j] | ((uint32_t)tile[1][tx][j] << 8) |
                                    ((uint32_t)tile[2][tx][j] << 16) | ((uint32_t)tile[3][tx][j] << 24);
    }
}

}  // namespace

namespace rvk {

// gen_coefficients (iir_blur.rs:142-146)
static void coefficients(double sigma, double *lambda, double *dnu) {
    double l = (sigma * sigma) / (2.0 * 4.0);
    *lambda = l;
    *dnu = (1.0 + 2.0 * l - std::sqrt(1.0 + 4.0 * l)) / (2.0 * l);
}

// Warm-up length of a speculated chunk: enough steps for dnu^K to fall below 2^-56
// (an error in the speculated start is then below an ulp of any value that is not
// itself vanishing), rounded up to 16, 24, 32 or 48: 16 for sigma up to about 0.9,
// 24 up to about 1.5, 32 below 2 (the largest sigma the IIR form gets from the filter
// graph), 48 beyond.  Only a cost: the fixpoint makes every chunk exact for any K.
static int warmup_steps(double dnu) {
    const double k = std::ceil(56.0 * std::log(2.0) / std::log(1.0 / dnu));
    return k <= 16.0 ? 16 : k <= 24.0 ? 24 : k <= 32.0 ? 32 : 48;
}

template <int CT>
static void launch_chains_ct(hipStream_t st, int K, const uint8_t *in_u8, const double *in_f64, uint8_t *out_u8,
                             double *out_f64, uint32_t n, uint32_t nchains, uint32_t estride, uint32_t cstride,
                             uint32_t plane, double dnu, int steps, double post, int planar) {
    const dim3 grid(nchains * 4u), blk(64);
#define RVK_CHAINS(KK)                                                                                          \
    hipLaunchKernelGGL((k_iir_chains<CT, KK>), grid, blk, 0, st, in_u8, in_f64, out_u8, out_f64, n, nchains,    \
                       estride, cstride, plane, dnu, steps, post, planar)
    if (K == 16) RVK_CHAINS(16);
    else if (K == 24) RVK_CHAINS(24);
    else if (K == 32) RVK_CHAINS(32);
    else RVK_CHAINS(48);
#undef RVK_CHAINS
}

// one wave per chain, CT values a lane: the smallest instantiated CT with 64 CT >= n
static void launch_chains(hipStream_t st, const uint8_t *in_u8, const double *in_f64, uint8_t *out_u8,
                          double *out_f64, uint32_t n, uint32_t nchains, uint32_t estride, uint32_t cstride,
                          uint32_t plane, double dnu, int steps, double post, int planar) {
    const int K = warmup_steps(dnu);
    const uint32_t cs = (n + 63u) / 64u;
#define RVK_CT(CT) \
    launch_chains_ct<CT>(st, K, in_u8, in_f64, out_u8, out_f64, n, nchains, estride, cstride, plane, dnu, steps, post, planar)
    if (cs <= 8u) RVK_CT(8);
    else if (cs <= 16u) RVK_CT(16);
    else if (cs <= 24u) RVK_CT(24);
    else if (cs <= 32u) RVK_CT(32);
    else if (cs <= 48u) RVK_CT(48);
    else RVK_CT(64);
#undef RVK_CT
}

// Both sides at most kChainMax: rows (u8 in, f64 plane out) then columns (f64 in, u8
// out), or one of them u8 to u8, or (both sigmas off) only the conversion round trip.
static rvk_status iir_on_chip(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, bool rows, bool cols,
                              double dx, double dy, double post) {
    const size_t n = (size_t)w * h;
    const uint32_t n32 = (uint32_t)n;
    // algorithmic bytes of the whole blur: the u8 image read once and written once
    const double pxb = 4.0 * (double)n;
    if (!cols) {  // rows only, or (both sigmas off) the conversion round trip
        timed_launch(ctx, "iir_rows", 2.0 * pxb, [&] {
            launch_chains(ctx->stream, px, nullptr, px, nullptr, w, h, 1u, w, n32, dx, rows ? 4 : 0, post, 0);
        });
        RVK_HIP_CHECK(hipGetLastError());
        return RVK_OK;
    }
    // the column pass writes per-channel planes in column order, then they are
    // interleaved back into the pixmap
    void *pa = nullptr, *pu = nullptr;
    RVK_TRY(scratch_alloc(ctx, n * 4, &pu));
    if (rows) {
        const rvk_status st = scratch_alloc(ctx, n * 4 * sizeof(double), &pa);
        if (st != RVK_OK) { scratch_release(ctx, pu); return st; }
    }
    // Columns longer than 2048 read the row pass's planes transposed (column-major,
    // contiguous chains): reading them across rows touches a cache line per value
    // and, with many such columns in flight, fetched 5.7 times the planes at 4096^2;
    // below that the transpose costs about what it saves.
    const bool transposed = rows && h > 2048u;
    void *pt = nullptr;
    if (transposed) {
        const rvk_status st = scratch_alloc(ctx, n * 4 * sizeof(double), &pt);
        if (st != RVK_OK) { scratch_release(ctx, pa); scratch_release(ctx, pu); return st; }
    }
    uint8_t *cplanes = (uint8_t *)pu;
    double *planes = (double *)pa, *tplanes = (double *)pt;
    if (rows) {
        timed_launch(ctx, "iir_rows", pxb, [&] {
            launch_chains(ctx->stream, px, nullptr, nullptr, planes, w, h, 1u, w, n32, dx, 4, 1.0, 0);
        });
        if (transposed) {
            // row-major planes (h rows of w) read as column-major ones of h x w
            timed_launch(ctx, "iir_transpose", 0.0, [&] {
                hipLaunchKernelGGL(k_transpose_planes, dim3(div_up(h, T), div_up(w, T), 4), dim3(T, 8), 0,
                                   ctx->stream, (const double *)planes, tplanes, h, w);
            });
            timed_launch(ctx, "iir_cols", 0.0, [&] {
                launch_chains(ctx->stream, nullptr, tplanes, cplanes, nullptr, h, w, 1u, h, n32, dy, 4, post, 1);
            });
        } else {
            timed_launch(ctx, "iir_cols", 0.0, [&] {
                launch_chains(ctx->stream, nullptr, planes, cplanes, nullptr, h, w, w, 1u, n32, dy, 4, post, 1);
            });
        }
    } else {
        timed_launch(ctx, "iir_cols", pxb, [&] {
            launch_chains(ctx->stream, px, nullptr, cplanes, nullptr, h, w, w, 1u, n32, dy, 4, post, 1);
        });
    }
    timed_launch(ctx, "iir_interleave", pxb, [&] {
        hipLaunchKernelGGL(k_planes_to_rgba, dim3(div_up(w, 32), div_up(h, 32)), dim3(32, 8), 0, ctx->stream,
                           (const uint8_t *)cplanes, (uint32_t *)px, w, h);
    });
    const hipError_t e = hipGetLastError();
    if (pt) scratch_release(ctx, pt);
    if (pa) scratch_release(ctx, pa);
    scratch_release(ctx, pu);
    RVK_HIP_CHECK(e);
    return RVK_OK;
}

rvk_status launch_iir_blur(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    double lx = 1.0, dx = 1.0, ly = 1.0, dy = 1.0;
    if (sx > 0.0) coefficients(sx, &lx, &dx);
    if (sy > 0.0) coefficients(sy, &ly, &dy);
    // post_scale = (sqrt(dnu_x dnu_y) / sqrt(lambda_x lambda_y))^8 by repeated squaring (iir_blur.rs:137-139)
    double base = std::sqrt(dx * dy) / std::sqrt(lx * ly);
    double post = base * base;
    post = post * post;
    post = post * post;

    size_t n = (size_t)w * h;
    // on chip: both sides within kChainMax and each coefficient in (0, 1) (it is for
    // every finite positive sigma; the zero-start argument needs dnu * 0 == 0)
    const bool coef_ok = (!(sx > 0.0) || (dx > 0.0 && dx < 1.0)) && (!(sy > 0.0) || (dy > 0.0 && dy < 1.0));
    if (w <= kChainMax && h <= kChainMax && coef_ok) return iir_on_chip(ctx, px, w, h, sx > 0.0, sy > 0.0, dx, dy, post);
    size_t bytes = n * 4 * sizeof(double);
    void *pa = nullptr, *pb = nullptr;
    RVK_TRY(scratch_alloc(ctx, bytes, &pa));
    rvk_status st = scratch_alloc(ctx, bytes, &pb);
    if (st != RVK_OK) { scratch_release(ctx, pa); return st; }
    double *colmajor = (double *)pa, *rowmajor = (double *)pb;
    dim3 tblk(T, 8);
    dim3 tgrid(div_up(w, T), div_up(h, T));
    // algorithmic bytes of the whole blur: the u8 image read once (charged to the
    // conversion in) and written once (charged to the conversion out); the f64
    // planes are this implementation's intermediates and are not credited
    const double pxb = 4.0 * (double)n;
    timed_launch(ctx, "iir_to_f64", pxb, [&] {
        hipLaunchKernelGGL(k_to_f64_colmajor, tgrid, tblk, 0, ctx->stream, (const uint32_t *)px, colmajor, w, h);
    });
    if (sx > 0.0) {
        timed_launch(ctx, "iir_rows", 0.0, [&] {
            hipLaunchKernelGGL(k_iir_rows, dim3(div_up(h, 64), 4), dim3(64), 0, ctx->stream, colmajor, w, h, dx);
        });
    }
    if (sy > 0.0) {
        timed_launch(ctx, "iir_transpose", 0.0, [&] {
            hipLaunchKernelGGL(k_transpose_planes, dim3(div_up(w, T), div_up(h, T), 4), tblk, 0, ctx->stream,
                               (const double *)colmajor, rowmajor, w, h);
        });
        timed_launch(ctx, "iir_cols", 0.0, [&] {
            hipLaunchKernelGGL(k_iir_cols, dim3(div_up(w, 64), 4), dim3(64), 0, ctx->stream, rowmajor, w, h, dy);
        });
        size_t g = (n + 255) / 256;
        if (g > 65535u * 4u) g = 65535u * 4u;
        timed_launch(ctx, "iir_from_f64", pxb, [&] {
            hipLaunchKernelGGL(k_from_f64_rowmajor, dim3((unsigned)g), dim3(256), 0, ctx->stream,
                               (const double *)rowmajor, (uint32_t *)px, n, post);
        });
    } else {
        timed_launch(ctx, "iir_from_f64", pxb, [&] {
            hipLaunchKernelGGL(k_from_f64_colmajor, tgrid, tblk, 0, ctx->stream, (const double *)colmajor,
                               (uint32_t *)px, w, h, post);
        });
    }
    hipError_t e = hipGetLastError();
    scratch_release(ctx, pa);
    scratch_release(ctx, pb);
    RVK_HIP_CHECK(e);
    return RVK_OK;
}

}  // namespace rvk
