// The differentiable f32 form of the pixel pipeline: a forward renderer and
// hand-written adjoint kernels for the scene model below, for one scene or a stack
// of equally shaped scenes.  Everything is f32, premultiplied RGBA, one thread per
// pixel.
//
// Model (the reference's operations restated in f32, resvg 0.48.1):
//   * path i covers pixel p with c_i(p): either coverage_u8 / 255, the anti-aliased
//     coverage of its outline (tiny-skia fill_path, path.rs:26-75) computed by the
//     u8 rasterizer and fixed (no gradient to the geometry), or, in the analytic
//     mode, min(1, |∫ winding|) over the pixel of the path's flattened cubics mapped
//     by its affine transform, whose gradient flows to the control points and the
//     transform through the boundary term;
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

#include <type_traits>

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
    uint32_t k;              // cubics per path (analytic mode), else 0
    size_t ws;               // workspace bytes per scene
    float qscale;            // analytic mode: fixed-point scale of the area accumulation
};

// A batch holds scenes of one shape, stacked: scene blockIdx.y finds its parameters,
// geometry, gradient ends, coverage planes, workspace and images at blockIdx.y times
// the per-scene size of each (a null pointer stays null).
// (char arithmetic, not an integer round trip, so that the compiler keeps the
// pointer in the global address space: scalar loads for uniform values)
template <class T> __device__ __forceinline__ T *scene_at(T *p, size_t bytes) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(p) + (size_t)blockIdx.y * bytes);
}
// the same for a pointer that may be null (an optional buffer)
template <class T> __device__ __forceinline__ T *scene_of(T *p, size_t bytes) {
    return p ? scene_at(p, bytes) : p;
}
__device__ __forceinline__ size_t params_bytes(const Dims &d) {
    return ((size_t)d.n * kPP + kGlobalParams) * sizeof(float);
}
__device__ __forceinline__ size_t geom_bytes(const Dims &d) {
    return (size_t)d.n * (6 + 2 * (3 * d.k + 1)) * sizeof(float);
}
__device__ __forceinline__ size_t image_bytes(const Dims &d) { return (size_t)d.w * d.h * sizeof(float4); }

// Where the coverage of path i at pixel p comes from: u8 planes (coverage / 255), or
// the winding-area planes of the analytic mode (coverage min(1, |area|); row stride
// w + 2, see k_diff_accum).  In the analytic mode the backward also writes
// dL/d(area) per path and pixel into gW.
struct Cov {
    const uint8_t *u8;
    const float *area;
    float *gW;
};
__device__ __forceinline__ size_t area_index(const Dims &d, uint32_t i, size_t p) {
    return ((size_t)i * d.h + p / d.w) * (d.w + 2) + p % d.w;
}
__device__ __forceinline__ Cov scene_cov(const Cov &cv, const Dims &d) {
    return Cov{scene_of(cv.u8, (size_t)d.n * d.w * d.h), scene_of(cv.area, d.ws), scene_of(cv.gW, d.ws)};
}
__device__ __forceinline__ float cov_at(const Cov &cv, const Dims &d, uint32_t i, size_t p) {
    if (cv.u8) return (float)cv.u8[(size_t)i * d.w * d.h + p] * (1.0f / 255.0f);
    return fminf(1.0f, fabsf(cv.area[area_index(d, i, p)]));
}

// ---- paint ------------------------------------------------------------------
struct PathEval {
    float rgb[3], a;   // unpremultiplied stop colour at the pixel
    float A;           // c * opacity * a
    float ws[4];       // colour = sum over stops s of ws[s] * stop s
    int k;             // stop interval: between stop k and stop k + 1, at u
    float u, t, lo, hi;  // lo, hi: offsets of stops k, k + 1
};

// The interpolation is written as per-stop weights so that every access to the
// path's parameters has a compile-time index (the values are uniform over the wave).
__device__ __forceinline__ PathEval eval_path(const float *pp, const float4 ends, float c, float px, float py) {
    PathEval e;
    const float dx = ends.z - ends.x, dy = ends.w - ends.y;
    const float dd = dx * dx + dy * dy;
    float t = dd > 0.0f ? ((px - ends.x) * dx + (py - ends.y) * dy) / dd : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float *col = pp + 4;   // 4 stops x rgba
    const float *off = pp + 20;  // 4 offsets
    // before the first stop: stop 0 (k = 0, u = 0); after the last: stop 3 (k = 2, u = 1)
    const bool before = t <= off[0], after = !before && t >= off[3];
    const int k = before ? 0 : (after ? 2 : (t < off[1] ? 0 : (t < off[2] ? 1 : 2)));
    const float lo = k == 0 ? off[0] : (k == 1 ? off[1] : off[2]);
    const float hi = k == 0 ? off[1] : (k == 1 ? off[2] : off[3]);
    const float span = hi - lo;
    const float u = before ? 0.0f : (after ? 1.0f : (span > 0.0f ? (t - lo) / span : 0.0f));
    const float v = 1.0f - u;
    e.ws[0] = k == 0 ? v : 0.0f;
    e.ws[1] = k == 1 ? v : (k == 0 ? u : 0.0f);
    e.ws[2] = k == 2 ? v : (k == 1 ? u : 0.0f);
    e.ws[3] = k == 2 ? u : 0.0f;
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
        const float q = e.ws[0] * col[ch] + e.ws[1] * col[4 + ch] + e.ws[2] * col[8 + ch] + e.ws[3] * col[12 + ch];
        if (ch < 3) e.rgb[ch] = q;
        else e.a = q;
    }
    e.A = c * pp[3] * e.a;
    e.k = k;
    e.u = u;
    e.t = t;
    e.lo = lo;
    e.hi = hi;
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

// Sums 32 values over the 64 lanes of a wave with a halving butterfly: at each step
// a lane keeps half of its values and adds a partner lane's copy of that half, so
// 16 + 8 + 4 + 2 + 1 exchanges and one final one replace 32 full reductions.  The
// exchanges stay in registers (no LDS): v_permlane16_swap pairs the 16-lane rows
// (values i and 16 + i), then DPP row rotations by 1, 2, 4 and 8 pair, within each
// row, the lanes with that lane bit clear with those with it set (a bijection that
// keeps the lower lane bits, which already select the values a lane holds), and
// v_permlane32_swap adds the two halves of the wave.  Returns, in every lane, the
// wave sum of value (lane & 31).
template <int N>
__device__ __forceinline__ float row_ror(float x) {
    return __uint_as_float(__builtin_amdgcn_update_dpp(0u, __float_as_uint(x), 0x120 + N, 0xf, 0xf, false));
}
template <int B, int M>  // lane bit B in {1, 2, 4, 8}: M values in, M / 2 out
__device__ __forceinline__ void halve_row(const float (&v)[M], float (&o)[M / 2], bool up) {
#pragma unroll
    for (int k = 0; k < M / 2; k++) {
        const float lo = v[2 * k] + row_ror<B>(v[2 * k]);
        const float hi = v[2 * k + 1] + row_ror<B>(v[2 * k + 1]);
        o[k] = up ? hi : lo;
    }
}
__device__ __forceinline__ float wave_sum32(const float (&v)[32], int lane) {
    float a[16], b[8], c[4], d[2], e[1];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        // rows 0 and 2 end with value i, rows 1 and 3 with value 16 + i
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[i]), __float_as_uint(v[16 + i]), false, false);
        a[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    halve_row<1, 16>(a, b, (lane & 1) != 0);
    halve_row<2, 8>(b, c, (lane & 2) != 0);
    halve_row<4, 4>(c, d, (lane & 4) != 0);
    halve_row<8, 2>(d, e, (lane & 8) != 0);
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(e[0]), __float_as_uint(e[0]), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// ---- forward: paths -----------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_diff_paths_fwd(Dims d, Cov cv,
                                                              const float4 *__restrict__ ends,
                                                              const float *__restrict__ params, float4 *canvas,
                                                              float4 *ck) {
    cv = scene_cov(cv, d);
    ends = scene_at(ends, (size_t)d.n * sizeof(float4));
    params = scene_at(params, params_bytes(d));
    canvas = scene_at(canvas, d.ws);
    ck = scene_at(ck, d.ws);
    const size_t np = (size_t)d.w * d.h;
    const size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (p >= np) return;
    const float px = (float)(p % d.w) + 0.5f, py = (float)(p / d.w) + 0.5f;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < d.n; i++) {
        if (i % kSeg == 0) ck[(size_t)(i / kSeg) * np + p] = s;
        const float c = cov_at(cv, d, i, p);
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
__global__ void k_diff_blur_weights(Dims d, const float *params, float *wts) {
    __shared__ float g[2 * kMaxRadius + 1];
    __shared__ float sums[2];
    params = scene_at(params, params_bytes(d));
    wts = scene_at(wts, d.ws);
    const float raw = params[(size_t)d.n * kPP];
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
    in = scene_at(in, d.ws);
    out = scene_of(out, d.ws);
    wts = scene_at(wts, d.ws);
    gdot = scene_of(gdot, d.ws);
    part = scene_of(part, d.ws);
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
// The colour is demultiplied by max(a, 1/255): exact for every alpha a u8 pixmap can
// hold (its non-zero alphas are >= 1/255), and below that a continuous ramp to zero,
// so that alphas which are rounding residue (the blur's far tails, cancelled winding
// sums) cannot turn into full-strength colours.
constexpr float kAlphaFloor = 1.0f / 255.0f;
__device__ __forceinline__ CmEval cm_eval(const float *m, float4 b) {
    CmEval e;
    const float a = b.w, r = 1.0f / fmaxf(a, kAlphaFloor);
    e.v[0] = b.x * r;
    e.v[1] = b.y * r;
    e.v[2] = b.z * r;
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
    B = scene_at(B, d.ws);
    S = scene_at(S, d.ws);
    params = scene_at(params, params_bytes(d));
    out = scene_at(out, image_bytes(d));
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
    B = scene_at(B, d.ws);
    S = scene_at(S, d.ws);
    params = scene_at(params, params_bytes(d));
    gout = scene_at(gout, image_bytes(d));
    gB = scene_at(gB, d.ws);
    gS = scene_at(gS, d.ws);
    part = scene_at(part, d.ws);
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
        const float a = b.w, r = 1.0f / fmaxf(a, kAlphaFloor);
        float4 gb;
        gb.x = gv[0] * r;
        gb.y = gv[1] * r;
        gb.z = gv[2] * r;
        // v_c = b_c / a depends on a only above the floor
        gb.w = a > kAlphaFloor ? gv[3] - (gv[0] * e.v[0] + gv[1] * e.v[1] + gv[2] * e.v[2]) * r : gv[3];
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
__global__ __launch_bounds__(kThreads) void k_diff_paths_bwd(Dims d, Cov cv,
                                                              const float4 *__restrict__ ends,
                                                              const float *__restrict__ params,
                                                              const float4 *__restrict__ ck,
                                                              const float4 *__restrict__ gS,
                                                              const float4 *__restrict__ gS2, float *part) {
    __shared__ float red[kWaves][kSeg][kPP];  // per wave, the segment's path sums
    cv = scene_cov(cv, d);
    ends = scene_at(ends, (size_t)d.n * sizeof(float4));
    params = scene_at(params, params_bytes(d));
    ck = scene_at(ck, d.ws);
    gS = scene_at(gS, d.ws);
    gS2 = scene_at(gS2, d.ws);
    part = scene_at(part, d.ws);
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
        float cov[kSeg], area[kSeg];  // the segment's coverage (and signed area) at this pixel, read once
        canv[0] = live ? ck[(size_t)sg * np + p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int j = 0; j < kSeg; j++) {
            canv[j + 1] = canv[j];
            const uint32_t i = i0 + j;
            cov[j] = 0.0f;
            area[j] = 0.0f;
            if (live && i < d.n) {
                if (cv.u8) {
                    cov[j] = (float)cv.u8[(size_t)i * np + p] * (1.0f / 255.0f);
                } else {
                    area[j] = cv.area[area_index(d, i, p)];
                    cov[j] = fminf(1.0f, fabsf(area[j]));
                }
                const float c = cov[j];
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
            const float c = cov[j];
            float gcov = 0.0f;  // dL/d(coverage)
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
                gcov = gA * pp[3] * e.a;
                const float ga = gA * c * pp[3];       // stop alpha (interpolated)
                const float gcol[4] = {grgb[0], grgb[1], grgb[2], ga};
                const float *col = pp + 4;
                const float *off = pp + 20;
                // stop colours: d colour / d stop s = ws[s]; du: the colour's slope
                // along u is stop k+1 - stop k, dotted with gcol
                float gd[4];
#pragma unroll
                for (int st = 0; st < 4; st++) {
                    gd[st] = 0.0f;
#pragma unroll
                    for (int ch = 0; ch < 4; ch++) {
                        gp[4 + 4 * st + ch] = gcol[ch] * e.ws[st];
                        gd[st] += gcol[ch] * col[4 * st + ch];
                    }
                }
                const float gu = e.k == 0 ? gd[1] - gd[0] : (e.k == 1 ? gd[2] - gd[1] : gd[3] - gd[2]);
                // u = (t - lo) / (hi - lo) inside the stops (constant outside)
                const float span = e.hi - e.lo;
                if (e.t > off[0] && e.t < off[3] && span > 0.0f) {
                    const float ga_lo = gu * (e.t - e.hi) / (span * span);
                    const float ga_hi = gu * -(e.t - e.lo) / (span * span);
                    gp[20] = e.k == 0 ? ga_lo : 0.0f;
                    gp[21] = e.k == 1 ? ga_lo : (e.k == 0 ? ga_hi : 0.0f);
                    gp[22] = e.k == 2 ? ga_lo : (e.k == 1 ? ga_hi : 0.0f);
                    gp[23] = e.k == 2 ? ga_hi : 0.0f;
                }
                g = make_float4(g.x * (1.0f - A), g.y * (1.0f - A), g.z * (1.0f - A), g.w * (1.0f - A));
            }
            if (cv.gW && live) {
                // coverage = min(1, |area|): d/d(area) = sign(area) inside (0, 1), else 0
                const float ar = area[j];
                const float da = fabsf(ar) < 1.0f ? (ar > 0.0f ? 1.0f : (ar < 0.0f ? -1.0f : 0.0f)) : 0.0f;
                cv.gW[(size_t)i * np + p] = gcov * da;
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

// ---- analytic coverage (geometry mode) ---------------------------------------
// A path is a closed chain of k cubics: user-space points P0 .. P3k (geom floats
// after the path's 6 affine values sx, ky, kx, sy, tx, ty; device = (sx x + kx y +
// tx, ky x + sy y + ty)), each cubic flattened into kFlat segments at uniform t,
// and the polygon closed by a line from P3k back to P0.  Vertex j of path i:
// vert[i * nv + j], nv = k * kFlat + 1.
constexpr int kFlat = 16;

__device__ __forceinline__ void bern(float t, float b[4]) {
    const float u = 1.0f - t;
    b[0] = u * u * u;
    b[1] = 3.0f * u * u * t;
    b[2] = 3.0f * u * t * t;
    b[3] = t * t * t;
}

__global__ void k_diff_flatten(Dims d, const float *__restrict__ geom, float2 *vert) {
    geom = scene_at(geom, geom_bytes(d));
    vert = scene_at(vert, d.ws);
    const uint32_t n = d.n, k = d.k;
    const uint32_t nv = k * kFlat + 1;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * nv) return;
    const uint32_t i = gid / nv, j = gid % nv;
    const float *g = geom + (size_t)i * (6 + 2 * (3 * k + 1));
    const uint32_t c = j / kFlat < k ? j / kFlat : k - 1;  // the last vertex is t = 1 of the last cubic
    const float t = (float)(j - c * kFlat) / (float)kFlat;
    float b[4];
    bern(t, b);
    float x = 0.0f, y = 0.0f;
    for (int m = 0; m < 4; m++) {
        x += b[m] * g[6 + 2 * (3 * c + m)];
        y += b[m] * g[6 + 2 * (3 * c + m) + 1];
    }
    vert[gid] = make_float2(g[0] * x + g[2] * y + g[4], g[1] * x + g[3] * y + g[5]);
}

__global__ void k_diff_zero(float *p, size_t n, size_t stride) {
    p = scene_at(p, stride);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// One straight piece of an edge within a pixel row (x in [0, w], dd its signed share
// of the row's height): adds the differences whose prefix sum along the row is the
// piece's contribution to each pixel's winding area (the trapezoid right of the
// piece, split by the pixel columns).  The differences are added in fixed point
// (round(v * q), integer atomics), so the sums do not depend on the order in which
// edges arrive.
__device__ __forceinline__ void add_q(int *cell, float v, float q) { atomicAdd(cell, __float2int_rn(v * q)); }
__device__ __forceinline__ void accum_piece(int *row, float xa, float xb, float dd, float q) {
    const float x0 = fminf(xa, xb), x1 = fmaxf(xa, xb);
    const float x0f = floorf(x0);
    const int x0i = (int)x0f;
    const float x1c = ceilf(x1);
    if (x1c <= x0f + 1.0f) {  // within one column
        const float xm = 0.5f * (xa + xb) - x0f;
        add_q(row + x0i, dd * (1.0f - xm), q);
        add_q(row + x0i + 1, dd * xm, q);
        return;
    }
    const float sl = 1.0f / (x1 - x0);
    const float fx0 = x0 - x0f;
    const float a0 = 0.5f * sl * (1.0f - fx0) * (1.0f - fx0);
    const float fx1 = x1 - x1c + 1.0f;
    const float am = 0.5f * sl * fx1 * fx1;
    const int x1i = (int)x1c;
    add_q(row + x0i, dd * a0, q);
    if (x1i == x0i + 2) {
        add_q(row + x0i + 1, dd * (1.0f - a0 - am), q);
    } else {
        const float a1 = sl * (1.5f - fx0);
        add_q(row + x0i + 1, dd * (a1 - a0), q);
        for (int xi = x0i + 2; xi < x1i - 1; xi++) add_q(row + xi, dd * sl, q);
        const float a2 = a1 + (float)(x1i - x0i - 3) * sl;
        add_q(row + x1i - 1, dd * (1.0f - a2 - am), q);
    }
    add_q(row + x1i, dd * am, q);
}

// Signed area accumulation of one polygon edge into its path's plane (row stride
// w + 2): per pixel row the edge spans, the row's share of the edge's winding is
// split between the cells its x range touches so that the prefix sum along the row
// is the winding-weighted area of each pixel (the exact area coverage of the
// polygon's winding number).  The edge's stretch in a row is cut where it crosses
// x = 0 and x = w: a piece left of the canvas is moved onto x = 0 (cell 0, so it
// changes the winding of every pixel of the row alike), a piece right of it lands
// in cell w, beyond the prefix, so the planes are exact whatever lies outside.  The
// planes hold fixed-point integers (scale d.qscale) until k_diff_prefix.
__global__ void k_diff_accum(Dims d, const float2 *__restrict__ vert, int *area) {
    vert = scene_at(vert, d.ws);
    area = scene_at(area, d.ws);
    const uint32_t nv = d.k * kFlat + 1;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= d.n * nv) return;
    const uint32_t i = gid / nv, j = gid % nv;
    float2 a = vert[gid], b = vert[(size_t)i * nv + (j + 1 == nv ? 0 : j + 1)];
    if (a.y == b.y) return;
    float dir = 1.0f;
    if (a.y > b.y) {
        const float2 t = a;
        a = b;
        b = t;
        dir = -1.0f;
    }
    const float W = (float)d.w;
    const float dxdy = (b.x - a.x) / (b.y - a.y);
    int *plane = area + (size_t)i * d.h * (d.w + 2);
    const int y0 = max(0, (int)floorf(a.y)), y1 = min((int)d.h, (int)ceilf(b.y));
    for (int y = y0; y < y1; y++) {
        const float ya = fmaxf(a.y, (float)y), yb = fminf(b.y, (float)(y + 1));
        const float dy = yb - ya;
        if (!(dy > 0.0f)) continue;
        const float xa = a.x + dxdy * (ya - a.y), xb = a.x + dxdy * (yb - a.y);
        int *row = plane + (size_t)y * (d.w + 2);
        const float dd = dy * dir;
        // cut points (fractions of the stretch) at x = 0 and x = w, in order
        float c0 = 2.0f, c1 = 2.0f;
        if (xb != xa) {
            const float u = (0.0f - xa) / (xb - xa), v = (W - xa) / (xb - xa);
            const float lo = fminf(u, v), hi = fmaxf(u, v);
            if (lo > 0.0f && lo < 1.0f) c0 = lo;
            if (hi > 0.0f && hi < 1.0f) {
                if (c0 > 1.0f) c0 = hi;
                else c1 = hi;
            }
        }
        const float cuts[4] = {0.0f, c0 < 1.0f ? c0 : 1.0f, c1 < 1.0f ? c1 : 1.0f, 1.0f};
        for (int q = 0; q < 3; q++) {
            const float t0 = cuts[q], t1 = cuts[q + 1];
            if (!(t1 > t0)) continue;
            const float pa = fminf(fmaxf(xa + t0 * (xb - xa), 0.0f), W);
            const float pb = fminf(fmaxf(xa + t1 * (xb - xa), 0.0f), W);
            accum_piece(row, pa, pb, dd * (t1 - t0), d.qscale);
        }
    }
}

// Prefix sums along each row of the area planes (one wave per row), in place: the
// fixed-point differences are summed in wrapping integer arithmetic (exact, and
// every prefix, a pixel's winding area, is within range) and written back as floats.
__global__ void k_diff_prefix(Dims d, float *area) {
    area = scene_at(area, d.ws);
    const uint32_t row = blockIdx.x;  // over n * h
    const int lane = threadIdx.x;
    float *r = area + (size_t)row * (d.w + 2);
    const int *ri = (const int *)r;
    const float inv = 1.0f / d.qscale;
    uint32_t carry = 0;
    // four 64-column chunks at a time: their loads are issued together and their
    // scans are independent chains
    for (uint32_t x0 = 0; x0 < d.w; x0 += 256) {
        uint32_t v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t x = x0 + 64 * c + lane;
            v[c] = x < d.w ? (uint32_t)ri[x] : 0u;
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t t = __shfl_up(v[c], o, 64);
                if (lane >= o) v[c] += t;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t x = x0 + 64 * c + lane;
            if (x < d.w) r[x] = (float)(int)(carry + v[c]) * inv;
            carry += __shfl(v[c], 63, 64);
        }
    }
}

// dL/d(area) through one edge: the boundary term.  Moving the edge point x(s) =
// a + s (b - a) by δx changes the winding area of the pixel it crosses by
// δx · (-(b-a).y, (b-a).x) ds, so per pixel crossed on s in [s0, s1]:
// dA/da = n ∫(1 - s) ds, dA/db = n ∫ s ds.  The pixels are walked in order along
// the edge; pixels outside the canvas are skipped.  Writes (dL/da, dL/db) per edge.
__global__ void k_diff_boundary(Dims d, const float2 *__restrict__ vert, const float *__restrict__ gW,
                                float4 *egrad) {
    vert = scene_at(vert, d.ws);
    gW = scene_at(gW, d.ws);
    egrad = scene_at(egrad, d.ws);
    const uint32_t nv = d.k * kFlat + 1;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= d.n * nv) return;
    const uint32_t i = gid / nv, j = gid % nv;
    const float2 a = vert[gid], b = vert[(size_t)i * nv + (j + 1 == nv ? 0 : j + 1)];
    const float tx = b.x - a.x, ty = b.y - a.y;
    float wa = 0.0f, wb = 0.0f;
    // the part of the edge inside the canvas, s in [s0, s1] (Liang-Barsky): the walk
    // below then crosses at most w + h + 2 pixels whatever the edge's length, and its
    // step count is capped at w + h + 3 even for non-finite end points (reads stay
    // behind the bounds check)
    float s0 = 0.0f, s1 = 1.0f;
    bool inside = tx != 0.0f || ty != 0.0f;
    if (tx != 0.0f) {
        const float u = -a.x / tx, v = ((float)d.w - a.x) / tx;
        s0 = fmaxf(s0, fminf(u, v));
        s1 = fminf(s1, fmaxf(u, v));
    } else {
        inside = inside && a.x >= 0.0f && a.x < (float)d.w;
    }
    if (ty != 0.0f) {
        const float u = -a.y / ty, v = ((float)d.h - a.y) / ty;
        s0 = fmaxf(s0, fminf(u, v));
        s1 = fminf(s1, fmaxf(u, v));
    } else {
        inside = inside && a.y >= 0.0f && a.y < (float)d.h;
    }
    if (inside && s0 < s1) {
        const float *g = gW + (size_t)i * d.w * d.h;
        const float px = fminf(fmaxf(a.x + s0 * tx, 0.0f), (float)d.w - 1.0f);
        const float py = fminf(fmaxf(a.y + s0 * ty, 0.0f), (float)d.h - 1.0f);
        int ix = (int)floorf(px), iy = (int)floorf(py);
        const int sx = tx > 0.0f ? 1 : -1, sy = ty > 0.0f ? 1 : -1;
        const float inf = 3.0e38f;
        float tmx = tx != 0.0f ? ((float)(ix + (sx > 0 ? 1 : 0)) - a.x) / tx : inf;
        float tmy = ty != 0.0f ? ((float)(iy + (sy > 0 ? 1 : 0)) - a.y) / ty : inf;
        const float tdx = tx != 0.0f ? fabsf(1.0f / tx) : inf, tdy = ty != 0.0f ? fabsf(1.0f / ty) : inf;
        float s = s0;
        const int steps = (int)fminf((s1 - s0) * (fabsf(tx) + fabsf(ty)), (float)(d.w + d.h)) + 3;
        for (int q = 0; q < steps && s < s1; q++) {
            const float sn = fminf(fminf(tmx, tmy), s1);
            if (ix >= 0 && ix < (int)d.w && iy >= 0 && iy < (int)d.h) {
                const float G = g[(size_t)iy * d.w + ix];
                const float i1 = sn - s, i2 = 0.5f * (sn * sn - s * s);
                wa += G * (i1 - i2);
                wb += G * i2;
            }
            s = sn;
            if (tmx < tmy) {
                ix += sx;
                tmx += tdx;
            } else {
                iy += sy;
                tmy += tdy;
            }
        }
    }
    // n = (-t.y, t.x)
    egrad[gid] = make_float4(-ty * wa, tx * wa, -ty * wb, tx * wb);
}

// Per path (one thread): vertex gradients from the edge gradients, through the
// flattening (Bernstein weights) to the device control points, then through the
// affine to the user-space points and the affine's six values.
__global__ void k_diff_geom_chain(Dims d, const float *__restrict__ geom, const float4 *__restrict__ egrad,
                                  float *ggeom) {
    geom = scene_at(geom, geom_bytes(d));
    egrad = scene_at(egrad, d.ws);
    ggeom = scene_at(ggeom, geom_bytes(d));
    const uint32_t n = d.n, k = d.k;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t nv = k * kFlat + 1, np = 3 * k + 1;
    const float *g = geom + (size_t)i * (6 + 2 * np);
    float *o = ggeom + (size_t)i * (6 + 2 * np);
    for (uint32_t q = 0; q < 6 + 2 * np; q++) o[q] = 0.0f;
    const float4 *e = egrad + (size_t)i * nv;
    for (uint32_t j = 0; j < nv; j++) {
        // vertex j starts edge j and ends edge j - 1 (edge nv - 1 closes to vertex 0)
        const float4 ea = e[j], eb = e[j == 0 ? nv - 1 : j - 1];
        const float gx = ea.x + eb.z, gy = ea.y + eb.w;
        const uint32_t c = j / kFlat < k ? j / kFlat : k - 1;
        const float t = (float)(j - c * kFlat) / (float)kFlat;
        float b[4];
        bern(t, b);
        for (int m = 0; m < 4; m++) {
            const uint32_t pi = 3 * c + m;
            const float ux = g[6 + 2 * pi], uy = g[6 + 2 * pi + 1];
            const float dx = b[m] * gx, dy = b[m] * gy;  // dL/d(device point)
            // device = (sx x + kx y + tx, ky x + sy y + ty)
            o[6 + 2 * pi] += g[0] * dx + g[1] * dy;
            o[6 + 2 * pi + 1] += g[2] * dx + g[3] * dy;
            o[0] += dx * ux;
            o[1] += dy * ux;
            o[2] += dx * uy;
            o[3] += dy * uy;
            o[4] += dx;
            o[5] += dy;
        }
    }
}

// Sums `rows` partial rows of `cols` values into out[col]: a block of 64 columns x
// 16 row lanes, each lane summing every 16th row, then the lanes in a fixed order.
__global__ __launch_bounds__(1024) void k_diff_reduce(const float *part, uint32_t rows, uint32_t cols, float *out,
                                                      size_t part_stride, size_t out_stride) {
    part = scene_at(part, part_stride);
    out = scene_at(out, out_stride);
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
    size_t S, T, B, gB, gS, gS2, wts, ck, part_p, part_m, part_s, area, gW, vert, egrad, total;
};
Layout layout(uint32_t w, uint32_t h, uint32_t n, uint32_t k) {
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
    L.area = L.part_s + al(2 * nblk * sizeof(float));
    const size_t nv = (size_t)k * kFlat + 1;
    L.gW = L.area + (k ? al((size_t)n * h * (w + 2) * sizeof(float)) : 0);
    L.vert = L.gW + (k ? al((size_t)n * w * h * sizeof(float)) : 0);
    L.egrad = L.vert + (k ? al((size_t)n * nv * sizeof(float2)) : 0);
    L.total = L.egrad + (k ? al((size_t)n * nv * sizeof(float4)) : 0);
    return L;
}

// Fixed-point scale of the area accumulation: a pixel's winding area is at most half
// the polygon's edge count (k * kFlat + 1) in magnitude, so with q = 2^(30 - ceil(log2
// of that + 1)) every prefix fits a 32-bit integer; at most 2^24 (a resolution of
// 6e-8, f32's near 1).  8-cubic paths: 2^23.
float area_qscale(uint32_t k) {
    const uint64_t edges = (uint64_t)k * kFlat + 1;
    int bits = 0;
    while ((1ull << bits) < edges / 2 + 1) bits++;
    const int sh = 30 - bits < 24 ? 30 - bits : 24;
    return std::ldexp(1.0f, sh);
}

// The kernels index vertices (path_count x (16 cubics + 1)) and prefix rows
// (path_count x height) with 32-bit grid and thread indices.
bool scene_ok(const rvk_diff_scene *s, const float *geom) {
    return s && s->width > 0 && s->height > 0 && s->path_count > 0 && s->dev_gradient_ends &&
           (s->cubics ? geom != nullptr && s->cubics <= 4096 : s->dev_coverage != nullptr) &&
           (uint64_t)s->width * s->height <= (1ull << 30) &&
           (uint64_t)s->path_count * ((uint64_t)s->cubics * kFlat + 1) < (1ull << 31) &&
           (uint64_t)s->path_count * s->height < (1ull << 31);
}

}  // namespace

extern "C" {

rvk_status rvk_diff_workspace_size(uint32_t width, uint32_t height, uint32_t path_count, uint32_t cubics,
                                   size_t *bytes) {
    if (!bytes || width == 0 || height == 0 || path_count == 0 || cubics > 4096) return RVK_ERR_INVALID_ARGUMENT;
    *bytes = layout(width, height, path_count, cubics).total;
    return RVK_OK;
}

rvk_status rvk_diff_forward_batch(rvk_context *ctx, const rvk_diff_scene *scene, uint32_t batch,
                                  const float *dev_params, const float *dev_geom, void *dev_workspace,
                                  float *dev_out) {
    if (!ctx || !scene_ok(scene, dev_geom) || !dev_params || !dev_workspace || !dev_out || batch == 0 ||
        batch > 65535)
        return RVK_ERR_INVALID_ARGUMENT;
    const uint32_t w = scene->width, h = scene->height, n = scene->path_count, k = scene->cubics;
    const Layout L = layout(w, h, n, k);
    uint8_t *ws = (uint8_t *)dev_workspace;
    const Dims d{w, h, n, (uint32_t)(((size_t)w * h + kThreads - 1) / kThreads), k, L.total, area_qscale(k)};
    float4 *S = (float4 *)(ws + L.S), *T = (float4 *)(ws + L.T), *B = (float4 *)(ws + L.B);
    float *wts = (float *)(ws + L.wts);
    const double px_bytes = (double)w * h * batch;
    Cov cv{scene->dev_coverage, nullptr, nullptr};
    if (k) {
        // analytic coverage: flatten, accumulate winding area, prefix along rows
        float *area = (float *)(ws + L.area);
        float2 *vert = (float2 *)(ws + L.vert);
        const size_t nv = (size_t)k * kFlat + 1, cells = (size_t)n * h * (w + 2);
        const unsigned ge = (unsigned)((n * nv + 255) / 256);
        const unsigned gz = (unsigned)((cells + 255) / 256 < 65535 ? (cells + 255) / 256 : 65535);
        rvk::timed_launch(ctx, "diff_coverage", (double)cells * 8.0 * batch, [&] {
            hipLaunchKernelGGL(k_diff_zero, dim3(gz, batch), dim3(256), 0, ctx->stream, area, cells, L.total);
            hipLaunchKernelGGL(k_diff_flatten, dim3(ge, batch), dim3(256), 0, ctx->stream, d, dev_geom, vert);
            hipLaunchKernelGGL(k_diff_accum, dim3(ge, batch), dim3(256), 0, ctx->stream, d, (const float2 *)vert,
                               (int *)area);
            hipLaunchKernelGGL(k_diff_prefix, dim3(n * h, batch), dim3(64), 0, ctx->stream, d, area);
        });
        cv = Cov{nullptr, area, nullptr};
    }
    rvk::timed_launch(ctx, "diff_paths_fwd", px_bytes * (16.0 + n), [&] {
        hipLaunchKernelGGL(k_diff_paths_fwd, dim3(d.nblk, batch), dim3(kThreads), 0, ctx->stream, d, cv,
                           (const float4 *)scene->dev_gradient_ends, dev_params, S, (float4 *)(ws + L.ck));
    });
    rvk::timed_launch(ctx, "diff_blur_weights", 0.0, [&] {
        hipLaunchKernelGGL(k_diff_blur_weights, dim3(1, batch), dim3(2 * kMaxRadius + 1), 0, ctx->stream, d,
                           dev_params, wts);
    });
    rvk::timed_launch(ctx, "diff_blur", px_bytes * 32.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk, batch), dim3(kThreads), 0, ctx->stream, d, 0, (const float4 *)S,
                           T, (const float *)wts, (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_blur", px_bytes * 32.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, dim3(d.nblk, batch), dim3(kThreads), 0, ctx->stream, d, 1, (const float4 *)T,
                           B, (const float *)wts, (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_cm_fwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_cm_fwd, dim3(d.nblk, batch), dim3(kThreads), 0, ctx->stream, d, (const float4 *)B,
                           (const float4 *)S, dev_params, (float4 *)dev_out);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

rvk_status rvk_diff_forward(rvk_context *ctx, const rvk_diff_scene *scene, const float *dev_params,
                            const float *dev_geom, void *dev_workspace, float *dev_out) {
    return rvk_diff_forward_batch(ctx, scene, 1, dev_params, dev_geom, dev_workspace, dev_out);
}

rvk_status rvk_diff_backward_batch(rvk_context *ctx, const rvk_diff_scene *scene, uint32_t batch,
                                   const float *dev_params, const float *dev_geom, void *dev_workspace,
                                   const float *dev_grad_out, float *dev_grad_params, float *dev_grad_geom) {
    if (!ctx || !scene_ok(scene, dev_geom) || !dev_params || !dev_workspace || !dev_grad_out || !dev_grad_params ||
        (scene->cubics && !dev_grad_geom) || batch == 0 || batch > 65535)
        return RVK_ERR_INVALID_ARGUMENT;
    const uint32_t w = scene->width, h = scene->height, n = scene->path_count, k = scene->cubics;
    const Layout L = layout(w, h, n, k);
    uint8_t *ws = (uint8_t *)dev_workspace;
    Cov cv{scene->dev_coverage, nullptr, nullptr};
    if (k) cv = Cov{nullptr, (const float *)(ws + L.area), (float *)(ws + L.gW)};
    const Dims d{w, h, n, (uint32_t)(((size_t)w * h + kThreads - 1) / kThreads), k, L.total, area_qscale(k)};
    const float4 *S = (const float4 *)(ws + L.S), *T = (const float4 *)(ws + L.T), *B = (const float4 *)(ws + L.B);
    float4 *gB = (float4 *)(ws + L.gB), *gS = (float4 *)(ws + L.gS), *gS2 = (float4 *)(ws + L.gS2);
    const float *wts = (const float *)(ws + L.wts);
    float *part_p = (float *)(ws + L.part_p), *part_m = (float *)(ws + L.part_m), *part_s = (float *)(ws + L.part_s);
    const double px_bytes = (double)w * h * batch;
    const dim3 grid(d.nblk, batch);
    rvk::timed_launch(ctx, "diff_cm_bwd", px_bytes * 80.0, [&] {
        hipLaunchKernelGGL(k_diff_cm_bwd, grid, dim3(kThreads), 0, ctx->stream, d, B, S, dev_params,
                           (const float4 *)dev_grad_out, gB, gS, part_m);
    });
    // B = V(T): dL/dT = V(gB) (into gS2 for now); σ term Σ gB · V'(T)
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, grid, dim3(kThreads), 0, ctx->stream, d, 1, (const float4 *)gB, gS2, wts,
                           (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, grid, dim3(kThreads), 0, ctx->stream, d, 1, T, (float4 *)nullptr, wts,
                           (const float4 *)gB, part_s);
    });
    // T = H(S): dL/dS += H(dL/dT) (into gB, free now); σ term Σ dL/dT · H'(S)
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 48.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, grid, dim3(kThreads), 0, ctx->stream, d, 0, S, (float4 *)nullptr, wts,
                           (const float4 *)gS2, part_s + d.nblk);
    });
    rvk::timed_launch(ctx, "diff_blur_bwd", px_bytes * 32.0, [&] {
        hipLaunchKernelGGL(k_diff_blur, grid, dim3(kThreads), 0, ctx->stream, d, 0, (const float4 *)gS2, gB, wts,
                           (const float4 *)nullptr, (float *)nullptr);
    });
    rvk::timed_launch(ctx, "diff_paths_bwd", px_bytes * (48.0 + 2.0 * n), [&] {
        hipLaunchKernelGGL(k_diff_paths_bwd, grid, dim3(kThreads), 0, ctx->stream, d, cv,
                           (const float4 *)scene->dev_gradient_ends, dev_params, (const float4 *)(ws + L.ck),
                           (const float4 *)gS, (const float4 *)gB, part_p);
    });
    const uint32_t np_params = n * kPP;
    const size_t pstride = ((size_t)np_params + kGlobalParams) * sizeof(float);
    rvk::timed_launch(ctx, "diff_reduce", 0.0, [&] {
        hipLaunchKernelGGL(k_diff_reduce, dim3((np_params + 63) / 64, batch), dim3(1024), 0, ctx->stream,
                           (const float *)part_p, d.nblk, np_params, dev_grad_params, L.total, pstride);
        // σ: both passes' block sums; matrix: 20 columns of block sums
        hipLaunchKernelGGL(k_diff_reduce, dim3(1, batch), dim3(1024), 0, ctx->stream, (const float *)part_s,
                           2u * d.nblk, 1u, dev_grad_params + np_params, L.total, pstride);
        hipLaunchKernelGGL(k_diff_reduce, dim3(1, batch), dim3(1024), 0, ctx->stream, (const float *)part_m, d.nblk,
                           20u, dev_grad_params + np_params + 1, L.total, pstride);
    });
    if (k) {
        const size_t nv = (size_t)k * kFlat + 1;
        const unsigned ge = (unsigned)((n * nv + 255) / 256);
        float4 *egrad = (float4 *)(ws + L.egrad);
        rvk::timed_launch(ctx, "diff_boundary", (double)n * w * h * 4.0 * batch, [&] {
            hipLaunchKernelGGL(k_diff_boundary, dim3(ge, batch), dim3(256), 0, ctx->stream, d,
                               (const float2 *)(ws + L.vert), (const float *)(ws + L.gW), egrad);
            hipLaunchKernelGGL(k_diff_geom_chain, dim3((n + 63) / 64, batch), dim3(64), 0, ctx->stream, d, dev_geom,
                               (const float4 *)egrad, dev_grad_geom);
        });
    }
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

rvk_status rvk_diff_backward(rvk_context *ctx, const rvk_diff_scene *scene, const float *dev_params,
                             const float *dev_geom, void *dev_workspace, const float *dev_grad_out,
                             float *dev_grad_params, float *dev_grad_geom) {
    return rvk_diff_backward_batch(ctx, scene, 1, dev_params, dev_geom, dev_workspace, dev_grad_out,
                                   dev_grad_params, dev_grad_geom);
}

}  // extern "C"
