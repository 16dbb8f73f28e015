// feTurbulence (crates/resvg/src/filter/turbulence.rs).  Write-only and FP64-VALU
// bound: per pixel and channel an octave loop of Perlin noise2 in f64.  The
// Park–Miller seeded lattice and gradient tables (turbulence.rs:93-140) are built on
// the host and staged into LDS per workgroup (34.9 KB); every f64 operation follows
// the reference's order, so the output is bit-identical.
#include "rvk_internal.h"
#include "device_math.h"

using namespace rvkd;

namespace {

constexpr int32_t kPerlinN = 0x1000, kBM = 0xff;
constexpr int kBLen = 0x100 + 0x100 + 2;

struct TurbArgs {
    double ox, oy, sx, sy, bfx, bfy;
    uint32_t octaves, stitch, fractal, premultiply;
};

__device__ __forceinline__ double s_curve(double t) { return t * t * (3.0 - 2.0 * t); }
__device__ __forceinline__ double lerp(double t, double a, double b) { return a + t * (b - a); }

__device__ __forceinline__ bool approx_zero_d(double a) {
    if (a == 0.0) return true;
    if (signbit(a)) return false;
    uint64_t u = (uint64_t)__double_as_longlong(a);
    return u <= 4u;
}

struct StitchInfo { int32_t width, height, wrap_x, wrap_y; };

__device__ double noise2(const int32_t *lat, const double *grad, int c, double x, double y, const StitchInfo *st) {
    double t = x + (double)kPerlinN;
    int32_t bx0 = as_i32d(t);
    int32_t bx1 = bx0 + 1;
    double rx0 = t - (double)as_i64d(t);
    double rx1 = rx0 - 1.0;
    t = y + (double)kPerlinN;
    int32_t by0 = as_i32d(t);
    int32_t by1 = by0 + 1;
    double ry0 = t - (double)as_i64d(t);
    double ry1 = ry0 - 1.0;
    if (st) {
        if (bx0 >= st->wrap_x) bx0 -= st->width;
        if (bx1 >= st->wrap_x) bx1 -= st->width;
        if (by0 >= st->wrap_y) by0 -= st->height;
        if (by1 >= st->wrap_y) by1 -= st->height;
    }
    bx0 &= kBM; bx1 &= kBM; by0 &= kBM; by1 &= kBM;
    int32_t i = lat[bx0], j = lat[bx1];
    int32_t b00 = lat[i + by0], b10 = lat[j + by0], b01 = lat[i + by1], b11 = lat[j + by1];
    double sx = s_curve(rx0), sy = s_curve(ry0);
    const double *g = grad + (size_t)c * kBLen * 2;
    double u = rx0 * g[2 * b00] + ry0 * g[2 * b00 + 1];
    double v = rx1 * g[2 * b10] + ry0 * g[2 * b10 + 1];
    double a = lerp(sx, u, v);
    u = rx0 * g[2 * b01] + ry1 * g[2 * b01 + 1];
    v = rx1 * g[2 * b11] + ry1 * g[2 * b11 + 1];
    double b = lerp(sx, u, v);
    return lerp(sy, a, b);
}

__global__ __launch_bounds__(256) void k_turbulence(uint32_t *dst, uint32_t w, uint32_t h, TurbArgs a,
                                                    const int32_t *g_lat, const double *g_grad) {
    __shared__ int32_t lat[kBLen];
    __shared__ double grad[4 * kBLen * 2];
    for (int i = threadIdx.x; i < kBLen; i += blockDim.x) lat[i] = g_lat[i];
    for (int i = threadIdx.x; i < 4 * kBLen * 2; i += blockDim.x) grad[i] = g_grad[i];
    __syncthreads();
    size_t n = (size_t)w * h;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(idx % w), y = (uint32_t)(idx / w);
        uint32_t out = 0;
        for (int c = 0; c < 4; c++) {
            double px = ((double)x + a.ox) / a.sx, py = ((double)y + a.oy) / a.sy;
            double bfx = a.bfx, bfy = a.bfy;
            StitchInfo st{}, *stp = nullptr;
            if (a.stitch) {  // turbulence.rs:155-180, tile = (x, y, w, h)
                double tw = (double)w, th = (double)h;
                if (!approx_zero_d(bfx)) {
                    double lo = floor(tw * bfx) / tw, hi = ceil(tw * bfx) / tw;
                    bfx = (bfx / lo < hi / bfx) ? lo : hi;
                }
                if (!approx_zero_d(bfy)) {
                    double lo = floor(th * bfy) / th, hi = ceil(th * bfy) / th;
                    bfy = (bfy / lo < hi / bfy) ? lo : hi;
                }
                st.width = as_i32d(tw * bfx + 0.5);
                st.height = as_i32d(th * bfy + 0.5);
                st.wrap_x = as_i32d((double)x * bfx + (double)kPerlinN + (double)st.width);
                st.wrap_y = as_i32d((double)y * bfy + (double)kPerlinN + (double)st.height);
                stp = &st;
            }
            double sum = 0.0, ratio = 1.0;
            px *= bfx;
            py *= bfy;
            for (uint32_t o = 0; o < a.octaves; o++) {
                double nv = noise2(lat, grad, c, px, py, stp);
                sum += (a.fractal ? nv : fabs(nv)) / ratio;
                px *= 2.0;
                py *= 2.0;
                ratio *= 2.0;
                if (stp) {
                    st.width *= 2;
                    st.wrap_x = 2 * st.wrap_x - kPerlinN;
                    st.height *= 2;
                    st.wrap_y = 2 * st.wrap_y - kPerlinN;
                }
            }
            double v = a.fractal ? (sum * 255.0 + 255.0) / 2.0 : sum * 255.0;
            out |= as_u8(f32_bound(0.0f, (float)v, 255.0f) + 0.5f) << (8 * c);
        }
        if (a.premultiply) out = premul_px(out);  // multiply_alpha (filter/mod.rs:1015)
        dst[idx] = out;
    }
}

}  // namespace

namespace rvk {

// Park–Miller random (turbulence.rs:287-294)
static int32_t pm_random(int32_t seed) {
    const int32_t RAND_M = 2147483647, RAND_A = 16807, RAND_Q = 127773, RAND_R = 2836;
    int32_t r = RAND_A * (seed % RAND_Q) - RAND_R * (seed / RAND_Q);
    if (r <= 0) r += RAND_M;
    return r;
}

// init (turbulence.rs:93-140)
static void turb_tables(int32_t seed, std::vector<int32_t> &lat, std::vector<double> &grad) {
    const int32_t RAND_M = 2147483647, BS = 0x100;
    lat.assign(kBLen, 0);
    grad.assign(4 * kBLen * 2, 0.0);
    if (seed <= 0) seed = -(seed % (RAND_M - 1)) + 1;
    if (seed > RAND_M - 1) seed = RAND_M - 1;
    auto G = [&](int k, int i, int j) -> double & { return grad[((size_t)k * kBLen + i) * 2 + j]; };
    for (int k = 0; k < 4; k++) {
        for (int i = 0; i < BS; i++) {
            lat[i] = i;
            for (int j = 0; j < 2; j++) {
                seed = pm_random(seed);
                G(k, i, j) = (double)((seed % (BS + BS)) - BS) / (double)BS;
            }
            double s = std::sqrt(G(k, i, 0) * G(k, i, 0) + G(k, i, 1) * G(k, i, 1));
            G(k, i, 0) /= s;
            G(k, i, 1) /= s;
        }
    }
    for (int i = BS - 1; i >= 1; i--) {
        int32_t k = lat[i];
        seed = pm_random(seed);
        int32_t j = seed % BS;
        lat[i] = lat[j];
        lat[j] = k;
    }
    for (int i = 0; i < BS + 2; i++) {
        lat[BS + i] = lat[i];
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < 2; j++) G(k, BS + i, j) = G(k, i, j);
    }
}

rvk_status launch_turbulence(rvk_context *ctx, uint8_t *dst, uint32_t w, uint32_t h, double ox, double oy,
                             double sx, double sy, double bfx, double bfy, uint32_t octaves, int32_t seed,
                             uint32_t stitch, uint32_t fractal, int premultiply) {
    std::vector<int32_t> lat;
    std::vector<double> grad;
    turb_tables(seed, lat, grad);
    size_t lat_bytes = lat.size() * 4, grad_bytes = grad.size() * 8;
    void *buf = nullptr;
    RVK_TRY(scratch_alloc(ctx, lat_bytes + 64 + grad_bytes, &buf));
    int32_t *dlat = (int32_t *)buf;
    double *dgrad = (double *)((uint8_t *)buf + ((lat_bytes + 63) / 64) * 64);
    // the tables are copied into the pinned staging ring, so they may die with this call
    rvk_status ust = upload_async(ctx, dlat, lat.data(), lat_bytes);
    if (ust == RVK_OK) ust = upload_async(ctx, dgrad, grad.data(), grad_bytes);
    if (ust != RVK_OK) { scratch_release(ctx, buf); return ust; }
    hipError_t e;
    TurbArgs a{ox, oy, sx, sy, bfx, bfy, octaves, stitch ? 1u : 0u, fractal ? 1u : 0u, premultiply ? 1u : 0u};
    size_t n = (size_t)w * h;
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;  // grid-stride: amortise the LDS table load
    timed_launch(ctx, "turbulence", 4.0 * (double)n, [&] {  // writes 4 bytes per pixel
        hipLaunchKernelGGL(k_turbulence, dim3((unsigned)g), dim3(256), 0, ctx->stream, (uint32_t *)dst, w, h, a,
                           (const int32_t *)dlat, (const double *)dgrad);
    });
    e = hipGetLastError();
    scratch_release(ctx, buf);
    RVK_HIP_CHECK(e);
    return RVK_OK;
}

}  // namespace rvk
