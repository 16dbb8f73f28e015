// The differentiable f32 form of the pixel pipeline: a forward renderer and
// hand-written adjoint kernels for the scene model below (the C4 form of the
// batched/differentiable workloads).  Everything is f32, premultiplied RGBA, one
// thread per pixel.
//
// Model (the reference's operations restated in f32, resvg 0.48.1):
//   * path i covers pixel p with c_i(p) = coverage_u8 / 255, the anti-aliased
//     coverage of its outline (tiny-skia fill_path, path.rs:26-75), computed by the
//     u8 rasterizer and fixed here (no gradient flows to the geometry);
//   * its paint is a pad-spread linear gradient (path.rs:118-177) along
//     (x0, y0) -> (x1, y1) with 4 stops (offset, unpremultiplied RGBA), times a
//     per-path RGB tint, times a per-path opacity (the group opacity of
//     render.rs:133); the premultiplied source is (rgb·tint·A, A), A = c·opacity·a;
//   * paths composite with SourceOver in order onto a transparent canvas S;
//   * the filter chain is feGaussianBlur (a separable Gaussian of radius
//     ceil(3σ), at most kMaxRadius, zero outside the image) -> feColorMatrix (4x5
//     on the demultiplied colour, clamped, premultiplied; color_matrix.rs:11-110)
//     -> composite over SourceGraphic (filter/mod.rs:731-793): out = F + (1 - F_a)·S.
//
// Parameters (f32, device): per path kDiffPathParams values (tint r, g, b;
// opacity; 4 stop colours r, g, b, a; 4 stop offsets), then σ, then the 20 matrix
// entries.  Gradients come in the same layout.
//
// Checkpointing: the forward stores the canvas before every kSeg-th path and the
// blur intermediates; the backward recomputes each segment of kSeg paths per pixel
// in registers and walks it in reverse.  Parameter gradients are summed per block
// into partial rows and reduced by a second kernel, so they are deterministic.
#include "rvk_internal.h"

namespace {

constexpr int kSeg = 8;          // paths per checkpoint segment
constexpr int kThreads = 1024;   // pixels per block (one thread each)
constexpr int kWaves = kThreads / 64;
constexpr int kMaxRadius = 48;   // 3 * the largest σ the model accepts (16)
constexpr int kPP = RVK_DIFF_PATH_PARAMS;
constexpr int kGlobalParams = 1 + 20;  // σ, matrix

struct Dims {
    uint32_t w, h, n;        // width, height, paths
    uint32_t nblk;           // blocks of the per-pixel kernels
};

// ---- paint ------------------------------------------------------------------
struct PathEval {
    float rgb[3], a;   // unpremultiplied stop colour at the pixel
    float A;           // c * opacity * a
    int k;             // stop interval: colour = lerp(stop k, stop k + 1, u)
    float u, t;
};

__device__ __forceinline__ PathEval eval_path(const float *pp, const float4 ends, float c, float px, float py) {
    PathEval e;
    const float dx = ends.z - ends.x, dy = ends.w - ends.y;
    const float dd = dx * dx + dy * dy;
    float t = dd > 0.0f ? ((px - ends.x) * dx + (py - ends.y) * dy) / dd : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float *col = pp + 4;   // 4 stops x rgba
    const float *off = pp + 20;  // 4 offsets
    int k;
    float u;
    if (t <= off[0]) {
        k = 0; u = 0.0f;
    } else if (t >= off[3]) {
        k = 2; u = 1.0f;
    } else {
        k = t < off[1] ? 0 : (t < off[2] ? 1 : 2);
        const float span = off[k + 1] - off[k];
        u = span > 0.0f ? (t - off[k]) / span : 0.0f;
    }
    for (int ch = 0; ch < 3; ch++) e.rgb[ch] = col[4 * k + ch] + u * (col[4 * (k + 1) + ch] - col[4 * k + ch]);
    e.a = col[4 * k + 3] + u * (col[4 * (k + 1) + 3] - col[4 * k + 3]);
    e.A = c * pp[3] * e.a;
    e.k = k;
    e.u = u;
    e.t = t;
    return e;
}

__device__ __forceinline__ float4 over(const float *pp, const PathEval &e, float4 s) {
    const float A = e.A, ia = 1.0f - A;
    return make_float4(e.rgb[0] * pp[0] * A + ia * s.x, e.rgb[1] * pp[1] * A + ia * s.y,
                       e.rgb[2] * pp[2] * A + ia * s.z, A + ia * s.w);
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Sums 32 values over the 64 lanes of a wave with a halving butterfly: at each
// step a lane keeps half of its values (the upper half when its lane bit is set)
// and adds the partner's copy of that half, so 16 + 8 + 4 + 2 + 1 + 1 exchanges
// replace 32 full reductions.  Returns, in every lane, the wave sum of value
// (lane & 31).
template <int N>
__device__ __forceinline__ void halve(const float (&v)[2 * N], float (&o)[N], int lane) {
    const bool up = (lane & N) != 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const float keep = up ? v[N + i] : v[i], send = up ? v[i] : v[N + i];
        o[i] = keep + __shfl_xor(send, N, 64);
    }
}
__device__ __forceinline__ float wave_sum32(const float (&v)[32], int lane) {
    float a[16], b[8], c[4], d[2], e[1];
    halve<16>(v, a, lane);
    halve<8>(a, b, lane);
    halve<4>(b, c, lane);
    halve<2>(c, d, lane);
    halve<1>(d, e, lane);
    return e[0] + __shfl_xor(e[0], 32, 64);
}

// ---- forward: paths -----------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_diff_paths_fwd(Dims d, const uint8_t *__restrict__ cov,
                                                              const float4 *__restrict__ ends,
                                                              const float *__restrict__ params, float4 *canvas,
                                                              float4 *ck) {
    const size_t np = (size_t)d.w * d.h;
    const size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (p >= np) return;
    const float px = (float)(p % d.w) + 0.5f, py = (float)(p / d.w) + 0.5f;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < d.n; i++) {
        if (i % kSeg == 0) ck[(size_t)(i / kSeg) * np + p] = s;
        const float c = (float)cov[(size_t)i * np + p] * (1.0f / 255.0f);
        if (c == 0.0f) continue;
        const float *pp = params + (size_t)i * kPP;
        s = over(pp, eval_path(pp, ends[i], c, px, py), s);
    }
    canvas[p] = s;
}

// ---- blur weights from σ (one block) ------------------------------------------
// wts[0] = R (as a float), wts[1 .. 2R+1] = w_k, wts[2R+2 .. 4R+2] = dw_k/dσ for
// k = -R .. R, with w_k = g_k / Σg, g_k = exp(-k² / (2σ²)), σ clamped to [0.05, 16];
// dw_k/dσ = w_k (k² - Σ_j w_j j²) / σ³.
__global__ void k_diff_blur_weights(const float *params, uint32_t n, float *wts) {
    __shared__ float g[2 * kMaxRadius + 1];
    __shared__ float sums[2];
    const float raw = params[(size_t)n * kPP];
    const float sigma = fminf(fmaxf(raw, 0.05f), 16.0f);
    const float live = raw > 0.05f && raw < 16.0f ? 1.0f : 0.0f;  // no σ gradient where it is clamped
    const int R = min((int)ceilf(3.0f * sigma), kMaxRadius);
    const int t = threadIdx.x;
    if (t <= 2 * R) {
        const float k = (float)(t - R);
        g[t] = expf(-k * k / (2.0f * sigma * sigma));
    }
    __syncthreads();
    if (t == 0) {
        float z = 0.0f, m2 = 0.0f;
        for (int j = 0; j <= 2 * R; j++) {
            const float k = (float)(j - R);
            z += g[j];
            m2 += g[j] * k * k;
        }
        sums[0] = z;
        sums[1] = m2 / z;  // E[k²]
        wts[0] = (float)R;
    }
    __syncthreads();
    if (t <= 2 * R) {
        const float k = (float)(t - R);
        const float w = g[t] / sums[0];
        wts[1 + t] = w;
        wts[2 + 2 * R + t] = live * w * (k * k - sums[1]) / (sigma * sigma * sigma);
    }
}

// ---- blur passes: out = conv(in, w) along x (dir 0) or y (dir 1), zero padding.
// With `grad` the pass also returns, per block, Σ_p dot(gdot[p], conv(in, w')[p])
// into part[blockIdx.x] (the σ derivative term of this pass).
__global__ __launch_bounds__(kThreads) void k_diff_blur(Dims d, int dir, const float4 *__restrict__ in,
                                                         float4 *__restrict__ out, const float *__restrict__ wts,
                                                         const float4 *__restrict__ gdot, float *part) {
    __shared__ float red[kWaves];
    const size_t np = (size_t)d.w * d.h;
    const size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x;
    const int R = (int)wts[0];
    const float *w = wts + 1, *wd = wts + 2 + 2 * R;
    float dot = 0.0f;
    if (p < np) {
        const int x = (int)(p % d.w), y = (int)(p / d.w);
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f), accd = acc;
        for (int k = -R; k <= R; k++) {
            const int xx = dir == 0 ? x + k : x, yy = dir == 0 ? y : y + k;
            if (xx < 0 || xx >= (int)d.w || yy < 0 || yy >= (int)d.h) continue;
            const float4 v = in[(size_t)yy * d.w + xx];
            const float wk = w[k + R];
            acc.x += wk * v.x; acc.y += wk * v.y; acc.z += wk * v.z; acc.w += wk * v.w;
            if (gdot) {
                const float wdk = wd[k + R];
                accd.x += wdk * v.x; accd.y += wdk * v.y; accd.z += wdk * v.z; accd.w += wdk * v.w;
            }
        }
        if (out) out[p] = acc;
        if (gdot) {
            const float4 g = gdot[p];
            dot = g.x * accd.x + g.y * accd.y + g.z * accd.z + g.w * accd.w;
        }
    }
    if (gdot) {
        dot = wave_sum(dot);
        if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = dot;
        __syncthreads();
        if (threadIdx.x == 0) {
            float s = 0.0f;
            for (int i = 0; i < kWaves; i++) s += red[i];
            part[blockIdx.x] = s;
        }
    }
}

// ---- colour matrix on the demultiplied colour ----------------------------------
struct CmEval {
    float v[4];  // demultiplied input (r, g, b, a)
    float o[4];  // matrix output before clamping
    float q[4];  // clamped
};
__device__ __forceinline__ CmEval cm_eval(const float *m, float4 b) {
    CmEval e;
    const float a = b.w;
    e.v[0] = a > 0.0f ? b.x / a : 0.0f;
    e.v[1] = a > 0.0f ? b.y / a : 0.0f;
    e.v[2] = a > 0.0f ? b.z / a : 0.0f;
    e.v[3] = a;
    for (int c = 0; c < 4; c++) {
        const float *r = m + 5 * c;
        e.o[c] = r[0] * e.v[0] + r[1] * e.v[1] + r[2] * e.v[2] + r[3] * e.v[3] + r[4];
        e.q[c] = fminf(fmaxf(e.o[c], 0.0f), 1.0f);
    }
    return e;
}

// final = F + (1 - F_a) S with F = premultiply(clamp(M v(B)))
__global__ __launch_bounds__(kThreads) void k_diff_cm_fwd(Dims d, const float4 *__restrict__ B,
                                                          const float4 *__restrict__ S,
                                                          const float *__restrict__ params, float4 *out) {
    const size_t np = (size_t)d.w * d.h;
    const size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (p >= np) return;
    const float *m = params + (size_t)d.n * kPP + 1;
    const CmEval e = cm_eval(m, B[p]);
    const float fa = e.q[3];
    const float4 s = S[p];
    out[p] = make_float4(e.q[0] * fa + (1.0f - fa) * s.x, e.q[1] * fa + (1.0f - fa) * s.y,
                         e.q[2] * fa + (1.0f - fa) * s.z, fa + (1.0f - fa) * s.w);
}

// Backward of the colour matrix and the final composite: from dL/dout, writes
// dL/dB and dL/dS (the composite's direct term) and the block's matrix partials
// into part[blockIdx.x * 20 ..].
__global__ __launch_bounds__(kThreads) void k_diff_cm_bwd(Dims d, const float4 *__restrict__ B,
                                                          const float4 *__restrict__ S,
                                                          const float *__restrict__ params,
                                                          const float4 *__restrict__ gout, float4 *gB, float4 *gS,
                                                          float *part) {
    __shared__ float red[kWaves][20];
    const size_t np = (size_t)d.w * d.h;
    const size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x;
    const float *m = params + (size_t)d.n * kPP + 1;
    float gm[20];
    for (int j = 0; j < 20; j++) gm[j] = 0.0f;
    if (p < np) {
        const float4 b = B[p], s = S[p], g = gout[p];
        const CmEval e = cm_eval(m, b);
        const float fa = e.q[3];
        // out_c = F_c + (1 - F_a) s_c ; F_c = q_c q_a (c < 3), F_a = q_a
        gS[p] = make_float4(g.x * (1.0f - fa), g.y * (1.0f - fa), g.z * (1.0f - fa), g.w * (1.0f - fa));
        const float gfa = g.w - (g.x * s.x + g.y * s.y + g.z * s.z + g.w * s.w);  // includes dF_a/dF_a = 1
        float gq[4] = {g.x * fa, g.y * fa, g.z * fa, gfa + g.x * e.q[0] + g.y * e.q[1] + g.z * e.q[2]};
        float go[4];
        for (int c = 0; c < 4; c++) go[c] = (e.o[c] > 0.0f && e.o[c] < 1.0f) ? gq[c] : 0.0f;
        float gv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int c = 0; c < 4; c++) {
            for (int j = 0; j < 4; j++) {
                gm[5 * c + j] = go[c] * e.v[j];
                gv[j] += go[c] * m[5 * c + j];
            }
            gm[5 * c + 4] = go[c];
        }
        const float a = b.w;
        float4 gb;
        if (a > 0.0f) {
            gb.x = gv[0] / a;
            gb.y = gv[1] / a;
            gb.z = gv[2] / a;
            gb.w = gv[3] - (gv[0] * b.x + gv[1] * b.y + gv[2] * b.z) / (a * a);
        } else {
            gb = make_float4(0.0f, 0.0f, 0.0f, gv[3]);
        }
        gB[p] = gb;
    }
    for (int j = 0; j < 20; j++) {
        const float v = wave_sum(gm[j]);
        if (threadIdx.x % 64 == 0) red[threadIdx.x / 64][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < 20) {
        float s = 0.0f;
        for (int i = 0; i < kWaves; i++) s += red[i][threadIdx.x];
        part[(size_t)blockIdx.x * 20 + threadIdx.x] = s;
    }
}

// ---- backward: paths ----------------------------------------------------------
// dL/dS arrives as gS (direct term) + gS2 (through the blur).  Segments are walked
// from the last; each is recomputed from its checkpoint, then reversed.  The block's
// sums of the per-path gradients go to part[(blockIdx.x * n + i) * kPP ..].
__global__ __launch_bounds__(kThreads) void k_diff_paths_bwd(Dims d, const uint8_t *__restrict__ cov,
                                                              const float4 *__restrict__ ends,
                                                              const float *__restrict__ params,
                                                              const float4 *__restrict__ ck,
                                                              const float4 *__restrict__ gS,
                                                              const float4 *__restrict__ gS2, float *part) {
    __shared__ float red[kWaves][kSeg][kPP];  // per wave, the segment's path sums
    const size_t np = (size_t)d.w * d.h;
    const size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x;
    const bool live = p < np;
    const int lane = (int)(threadIdx.x % 64), wave = (int)(threadIdx.x / 64);
    const float px = live ? (float)(p % d.w) + 0.5f : 0.0f, py = live ? (float)(p / d.w) + 0.5f : 0.0f;
    float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (live) {
        const float4 a = gS[p], b = gS2[p];
        g = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    }
    const uint32_t nseg = (d.n + kSeg - 1) / kSeg;
    for (int32_t sg = (int32_t)nseg - 1; sg >= 0; sg--) {
        const uint32_t i0 = (uint32_t)sg * kSeg;
        float4 canv[kSeg + 1];
        canv[0] = live ? ck[(size_t)sg * np + p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int j = 0; j < kSeg; j++) {
            canv[j + 1] = canv[j];
            const uint32_t i = i0 + j;
            if (live && i < d.n) {
                const float c = (float)cov[(size_t)i * np + p] * (1.0f / 255.0f);
                if (c != 0.0f) {
                    const float *pp = params + (size_t)i * kPP;
                    canv[j + 1] = over(pp, eval_path(pp, ends[i], c, px, py), canv[j]);
                }
            }
        }
#pragma unroll
        for (int j = kSeg - 1; j >= 0; j--) {
            const uint32_t i = i0 + j;
            if (i >= d.n) continue;  // uniform
            float gp[32];  // kPP gradients, padded to 32 for the butterfly
#pragma unroll
            for (int q = 0; q < 32; q++) gp[q] = 0.0f;
            const float c = live ? (float)cov[(size_t)i * np + p] * (1.0f / 255.0f) : 0.0f;
            if (c != 0.0f) {
                const float *pp = params + (size_t)i * kPP;
                const PathEval e = eval_path(pp, ends[i], c, px, py);
                const float4 sb = canv[j];
                const float A = e.A;
                // after = (rgb·tint·A, A) + (1 - A)·before
                const float gc[3] = {g.x, g.y, g.z};
                float gA = g.w - (g.x * sb.x + g.y * sb.y + g.z * sb.z + g.w * sb.w);
                float grgb[3];
                for (int ch = 0; ch < 3; ch++) {
                    gA += gc[ch] * e.rgb[ch] * pp[ch];
                    gp[ch] = gc[ch] * A * e.rgb[ch];   // tint
                    grgb[ch] = gc[ch] * A * pp[ch];    // stop colour (interpolated)
                }
                gp[3] = gA * c * e.a;                  // opacity
                const float ga = gA * c * pp[3];       // stop alpha (interpolated)
                const float gcol[4] = {grgb[0], grgb[1], grgb[2], ga};
                const float *col = pp + 4;
                const float *off = pp + 20;
                float gu = 0.0f;
                for (int ch = 0; ch < 4; ch++) {
                    gp[4 + 4 * e.k + ch] += gcol[ch] * (1.0f - e.u);
                    gp[4 + 4 * (e.k + 1) + ch] += gcol[ch] * e.u;
                    gu += gcol[ch] * (col[4 * (e.k + 1) + ch] - col[4 * e.k + ch]);
                }
                // u = (t - o_k) / (o_k+1 - o_k) inside the stops (constant outside)
                if (e.t > off[0] && e.t < off[3]) {
                    const float span = off[e.k + 1] - off[e.k];
                    if (span > 0.0f) {
                        gp[20 + e.k] += gu * (e.t - off[e.k + 1]) / (span * span);
                        gp[20 + e.k + 1] += gu * -(e.t - off[e.k]) / (span * span);
                    }
                }
                g = make_float4(g.x * (1.0f - A), g.y * (1.0f - A), g.z * (1.0f - A), g.w * (1.0f - A));
            }
            // this wave's sums of the path's gradient
            const bool any = __ballot(c != 0.0f) != 0;
            const float v = any ? wave_sum32(gp, lane) : 0.0f;
            if (lane < kPP) red[wave][j][lane] = v;
        }
        // block sums of the segment's paths
        __syncthreads();
        for (int q = threadIdx.x; q < kSeg * kPP; q += kThreads) {
            const int j = q / kPP, k = q % kPP;
            if (i0 + j < d.n) {
                float s = 0.0f;
                for (int w = 0; w < kWaves; w++) s += red[w][j][k];
                part[((size_t)blockIdx.x * d.n + i0 + j) * kPP + k] = s;
            }
        }
        __syncthreads();
    }
}

// Sums `rows` partial rows of `cols` values into out[col]: a block of 64 columns x
// 16 row lanes, each lane summing every 16th row, then the lanes in a fixed order.
__global__ __launch_bounds__(1024) void k_diff_reduce(const float *part, uint32_t rows, uint32_t cols, float *out) {
    __shared__ float acc[16][65];
    const uint32_t tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    const uint32_t c = blockIdx.x * 64 + tx;
    float s = 0.0f;
    if (c < cols)
        for (uint32_t r = ty; r < rows; r += 16) s += part[(size_t)r * cols + c];
    acc[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && c < cols) {
        float t = 0.0f;
        for (int k = 0; k < 16; k++) t += acc[k][tx];
        out[c] = t;
    }
}

// Workspace layout (bytes, 256-aligned): canvas S | T | B | gB | gS | gS2 | weights |
// checkpoints | partial rows (paths) | partial rows (matrix) | σ partials (2 rows).
struct Layout {
    size_t S, T, B, gB, gS, gS2, wts, ck, part_p, part_m, part_s, total;
};
Layout layout(uint32_t w, uint32_t h, uint32_t n) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t img = al((size_t)w * h * sizeof(float4));
    const size_t nblk = ((size_t)w * h + kThreads - 1) / kThreads;
    Layout L;
    L.S = 0;
    L.T = L.S + img;
    L.B = L.T + img;
    L.gB = L.B + img;
    L.gS = L.gB + img;
    L.gS2 = L.gS + img;
    L.wts = L.gS2 + img;
    L.ck = L.wts + al((4 * kMaxRadius + 3) * sizeof(float));
    L.part_p = L.ck + al((size_t)((n + kSeg - 1) / kSeg) * w * h * sizeof(float4));
    L.part_m = L.part_p + al(nblk * n * kPP * sizeof(float));
    L.part_s = L.part_m + al(nblk * 20 * sizeof(float));
    L.total = L.part_s + al(2 * nblk * sizeof(float));
    return L;
}

bool scene_ok(const rvk_diff_scene *s) {
    return s && s->width > 0 && s->height > 0 && s->path_count > 0 && s->dev_coverage && s->dev_gradient_ends &&
           (uint64_t)s->width * s->height <= (1ull << 30);
}

}  // namespace

extern "C" {

rvk_status rvk_diff_workspace_size(uint32_t width, uint32_t height, uint32_t path_count, size_t *bytes) {
    if (!bytes || width == 0 || height == 0 || path_count == 0) return RVK_ERR_INVALID_ARGUMENT;
    *bytes = layout(width, height, path_count).total;
    return RVK_OK;
}

rvk_status rvk_diff_forward(rvk_context *ctx, const rvk_diff_scene *scene, const float *dev_params,
                            void *dev_workspace, float *dev_out) {
    if (!ctx || !scene_ok(scene) || !dev_params || !dev_workspace || !dev_out) return RVK_ERR_INVALID_ARGUMENT;
    const uint32_t w = scene->width, h = scene->height, n = scene->path_count;
    const Layout L = layout(w, h, n);
    uint8_t *ws = (uint8_t *)dev_workspace;
    const Dims d{w, h, n, (uint32_t)(((size_t)w * h + kThreads - 1) / kThreads)};
    float4 *S = (float4 *)(ws + L.S), *T = (float4 *)(ws + L.T), *B = (float4 *)(ws + L.B);
    float *wts = (float *)(ws + L.wts);
    const double px_bytes = (double)w * h;
    rvk::timed_launch(ctx, "diff_paths_fwd", px_bytes * (16.0 + n), [&] {
        hipLaunchKernelGGL(k_diff_paths_fwd, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, scene->dev_coverage,
                           (const float4 *)scene->dev_gradient_ends, dev_params, S, (float4 *)(ws + L.ck));
    });
    rvk::timed_launch(ctx, "diff_blur_weights", 0.0, [&] {
        hipLaunchKernelGGL(k_diff_blur_weights, dim3(1), dim3(2 * kMaxRadius + 1), 0, ctx->stream, dev_params, n,
                           wts);
    });
    rvk::timed_launch(ctx, "diff_blur", px_bytes * 32.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, 0, (const float4 *)S, T,
                           (const float *)wts, (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_blur", px_bytes * 32.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, 1, (const float4 *)T, B,
                           (const float *)wts, (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_cm_fwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_cm_fwd, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, (const float4 *)B,
                           (const float4 *)S, dev_params, (float4 *)dev_out);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

rvk_status rvk_diff_backward(rvk_context *ctx, const rvk_diff_scene *scene, const float *dev_params,
                             void *dev_workspace, const float *dev_grad_out, float *dev_grad_params) {
    if (!ctx || !scene_ok(scene) || !dev_params || !dev_workspace || !dev_grad_out || !dev_grad_params)
        return RVK_ERR_INVALID_ARGUMENT;
    const uint32_t w = scene->width, h = scene->height, n = scene->path_count;
    const Layout L = layout(w, h, n);
    uint8_t *ws = (uint8_t *)dev_workspace;
    const Dims d{w, h, n, (uint32_t)(((size_t)w * h + kThreads - 1) / kThreads)};
    const float4 *S = (const float4 *)(ws + L.S), *T = (const float4 *)(ws + L.T), *B = (const float4 *)(ws + L.B);
    float4 *gB = (float4 *)(ws + L.gB), *gS = (float4 *)(ws + L.gS), *gS2 = (float4 *)(ws + L.gS2);
    const float *wts = (const float *)(ws + L.wts);
    float *part_p = (float *)(ws + L.part_p), *part_m = (float *)(ws + L.part_m), *part_s = (float *)(ws + L.part_s);
    const double px_bytes = (double)w * h;
    rvk::timed_launch(ctx, "diff_cm_bwd", px_bytes * 80.0, [&] {
        hipLaunchKernelGGL(k_diff_cm_bwd, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, B, S, dev_params,
                           (const float4 *)dev_grad_out, gB, gS, part_m);
    });
    // B = V(T): dL/dT = V(gB) (into gS2 for now); σ term Σ gB · V'(T)
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, 1, (const float4 *)gB,
                           gS2, wts, (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, 1, T, (float4 *)nullptr,
                           wts, (const float4 *)gB, part_s);
    });
    // T = H(S): dL/dS += H(dL/dT) (into gB, free now); σ term Σ dL/dT · H'(S)
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, 0, S, (float4 *)nullptr,
                           wts, (const float4 *)gS2, part_s + d.nblk);
    });
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 32.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, 0, (const float4 *)gS2,
                           gB, wts, (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_paths_bwd", px_bytes * (48.0 + 2.0 * n), [&] {
        hipLaunchKernelGGL(k_diff_paths_bwd, dim3(d.nblk), dim3(kThreads), 0, ctx->stream, d, scene->dev_coverage,
                           (const float4 *)scene->dev_gradient_ends, dev_params,
                           (const float4 *)(ws + L.ck), (const float4 *)gS, (const float4 *)gB, part_p);
    });
    const uint32_t np_params = n * kPP;
    rvk::timed_launch(ctx, "diff_reduce", 0.0, [&] {
        hipLaunchKernelGGL(k_diff_reduce, dim3((np_params + 63) / 64), dim3(1024), 0, ctx->stream,
                           (const float *)part_p, d.nblk, np_params, dev_grad_params);
        // σ: both passes' block sums; matrix: 20 columns of block sums
        hipLaunchKernelGGL(k_diff_reduce, dim3(1), dim3(1024), 0, ctx->stream, (const float *)part_s, 2u * d.nblk,
                           1u, dev_grad_params + np_params);
        hipLaunchKernelGGL(k_diff_reduce, dim3(1), dim3(1024), 0, ctx->stream, (const float *)part_m, d.nblk, 20u,
                           dev_grad_params + np_params + 1);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

}  // extern "C"
