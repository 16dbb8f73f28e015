// Neighbourhood kernels: feConvolveMatrix (crates/resvg/src/filter/convolve_matrix.rs)
// and diffuse/specular lighting (crates/resvg/src/filter/lighting.rs).  Both read a
// small window per output pixel; the window is served by L1/L2 from a one-pixel-per-
// lane layout (adjacent lanes touch adjacent pixels).  Tap order and f32 operation
// order follow the reference exactly, so the results are bit-identical.
#include "rvk_internal.h"
#include "device_math.h"

using namespace rvkd;

namespace {

constexpr int kMaxKernel = 64;  // convolve kernel entries kept in the argument block

struct ConvArgs {
    uint32_t columns, rows, tx, ty, edge, preserve_alpha;
    float divisor, bias;
    float k[kMaxKernel];  // already flipped: k[oy*columns+ox] = data[(rows-oy-1)*columns + (columns-ox-1)]
};

__constant__ float c_inv255_dummy;  // keeps a constant segment even when unused

__global__ void k_convolve(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint32_t w, uint32_t h,
                           ConvArgs a, const float *__restrict__ big_k) {
    size_t n = (size_t)w * h;
    const float *kk = big_k ? big_k : a.k;
    const int32_t wmax = (int32_t)w - 1, hmax = (int32_t)h - 1;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int32_t x = (int32_t)(i % w), y = (int32_t)(i / w);
        float nr = 0.0f, ng = 0.0f, nb = 0.0f, na = 0.0f;
        for (uint32_t oy = 0; oy < a.rows; oy++) {
            for (uint32_t ox = 0; ox < a.columns; ox++) {
                int32_t tx = x - (int32_t)a.tx + (int32_t)ox;
                int32_t ty = y - (int32_t)a.ty + (int32_t)oy;
                if (a.edge == RVK_EDGE_NONE) {
                    if (tx < 0 || tx > wmax || ty < 0 || ty > hmax) continue;
                } else if (a.edge == RVK_EDGE_DUPLICATE) {
                    tx = max(0, min(wmax, tx));
                    ty = max(0, min(hmax, ty));
                } else {
                    // while tx < 0 { tx += w }; tx %= w  ==  Euclidean remainder
                    tx %= (int32_t)w; if (tx < 0) tx += (int32_t)w;
                    ty %= (int32_t)h; if (ty < 0) ty += (int32_t)h;
                }
                float kv = kk[oy * a.columns + ox];
                uint32_t p = src[(size_t)ty * w + tx];
                nr += (float)ch(p, 0) / 255.0f * kv;
                ng += (float)ch(p, 1) / 255.0f * kv;
                nb += (float)ch(p, 2) / 255.0f * kv;
                if (!a.preserve_alpha) na += (float)ch(p, 3) / 255.0f * kv;
            }
        }
        if (a.preserve_alpha) na = (float)ch(src[i], 3) / 255.0f;
        else na = na / a.divisor + a.bias;
        float bna = f32_bound(0.0f, na, 1.0f);
        auto calc = [&](float v) {
            v = v / a.divisor + a.bias * na;
            v = a.preserve_alpha ? f32_bound(0.0f, v, 1.0f) * bna : f32_bound(0.0f, v, bna);
            return as_u8(v * 255.0f + 0.5f);
        };
        dst[i] = pack(calc(nr), calc(ng), calc(nb), as_u8(bna * 255.0f + 0.5f));
    }
}

// ---------------------------------------------------------------------------
// Lighting
// ---------------------------------------------------------------------------
struct LightArgs {
    uint32_t specular, kind, has_cone;
    float surface_scale, constant, spec_exp;
    float r, g, b;                 // lighting colour as f32 of its u8 values
    uint32_t cr, cg, cb;           // lighting colour u8
    float lx, ly, lz;              // distant: fixed light vector; point/spot: position
    float dx, dy, dz;              // spot: normalized direction (host-computed, lighting.rs:313-316)
    float spot_exp, cos_cone;      // spot: exponent, cos(to_radians(cone)) (host libm)
    uint32_t exp_is_one;           // specular_exponent.approx_eq_ulps(1.0, 4)
};

struct V3 { float x, y, z; };
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }

__device__ __forceinline__ int32_t A(const uint32_t *s, uint32_t w, uint32_t x, uint32_t y) {
    return (int32_t)(s[(size_t)y * w + x] >> 24);
}

// Sobel normals with border factors (lighting.rs:340-476): returns (fx, fy, nx, ny)
// where (nx, ny) is the negated kernel response.
__device__ void sobel(const uint32_t *s, uint32_t w, uint32_t h, uint32_t x, uint32_t y, float &fx, float &fy,
                      float &nx, float &ny) {
    const float F12 = 1.0f / 2.0f, F13 = 1.0f / 3.0f, F14 = 1.0f / 4.0f, F23 = 2.0f / 3.0f;
    const uint32_t r = w - 1, b = h - 1;
    int32_t kx, ky;
    if (y == 0) {
        if (x == 0) {
            int32_t c = A(s, w, 0, 0), rr = A(s, w, 1, 0), bo = A(s, w, 0, 1), br = A(s, w, 1, 1);
            fx = F23; fy = F23; kx = -2 * c + 2 * rr - bo + br; ky = -2 * c - rr + 2 * bo + br;
        } else if (x == r) {
            int32_t l = A(s, w, r - 1, 0), c = A(s, w, r, 0), bl = A(s, w, r - 1, 1), bo = A(s, w, r, 1);
            fx = F23; fy = F23; kx = -2 * l + 2 * c - bl + bo; ky = -l - 2 * c + bl + 2 * bo;
        } else {
            int32_t l = A(s, w, x - 1, 0), c = A(s, w, x, 0), rr = A(s, w, x + 1, 0), bl = A(s, w, x - 1, 1),
                    bo = A(s, w, x, 1), br = A(s, w, x + 1, 1);
            fx = F13; fy = F12; kx = -2 * l + 2 * rr - bl + br; ky = -l - 2 * c - rr + bl + 2 * bo + br;
        }
    } else if (y == b) {
        if (x == 0) {
            int32_t t = A(s, w, 0, b - 1), tr = A(s, w, 1, b - 1), c = A(s, w, 0, b), rr = A(s, w, 1, b);
            fx = F23; fy = F23; kx = -t + tr - 2 * c + 2 * rr; ky = -2 * t - tr + 2 * c + rr;
        } else if (x == r) {
            int32_t tl = A(s, w, r - 1, b - 1), t = A(s, w, r, b - 1), l = A(s, w, r - 1, b), c = A(s, w, r, b);
            fx = F23; fy = F23; kx = -tl + t - 2 * l + 2 * c; ky = -tl - 2 * t + l + 2 * c;
        } else {
            int32_t tl = A(s, w, x - 1, b - 1), t = A(s, w, x, b - 1), tr = A(s, w, x + 1, b - 1),
                    l = A(s, w, x - 1, b), c = A(s, w, x, b), rr = A(s, w, x + 1, b);
            fx = F13; fy = F12; kx = -tl + tr - 2 * l + 2 * rr; ky = -tl - 2 * t - tr + l + 2 * c + rr;
        }
    } else if (x == 0) {
        int32_t t = A(s, w, 0, y - 1), tr = A(s, w, 1, y - 1), c = A(s, w, 0, y), rr = A(s, w, 1, y),
                bo = A(s, w, 0, y + 1), br = A(s, w, 1, y + 1);
        fx = F12; fy = F13; kx = -t + tr - 2 * c + 2 * rr - bo + br; ky = -2 * t - tr + 2 * bo + br;
    } else if (x == r) {
        int32_t tl = A(s, w, r - 1, y - 1), t = A(s, w, r, y - 1), l = A(s, w, r - 1, y), c = A(s, w, r, y),
                bl = A(s, w, r - 1, y + 1), bo = A(s, w, r, y + 1);
        fx = F12; fy = F13; kx = -tl + t - 2 * l + 2 * c - bl + bo; ky = -tl - 2 * t + bl + 2 * bo;
    } else {
        int32_t tl = A(s, w, x - 1, y - 1), t = A(s, w, x, y - 1), tr = A(s, w, x + 1, y - 1), l = A(s, w, x - 1, y),
                rr = A(s, w, x + 1, y), bl = A(s, w, x - 1, y + 1), bo = A(s, w, x, y + 1),
                br = A(s, w, x + 1, y + 1);
        fx = F14; fy = F14; kx = -tl + tr - 2 * l + 2 * rr - bl + br; ky = -tl - 2 * t - tr + bl + 2 * bo + br;
    }
    nx = (float)(-kx);
    ny = (float)(-ky);
}

__device__ __forceinline__ uint32_t lcompute(float c, float f) { return as_u8(f32_bound(0.0f, c * f, 255.0f) + 0.5f); }

__global__ void k_lighting(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint32_t w, uint32_t h,
                           LightArgs a) {
    size_t n = (size_t)w * h;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(i % w), y = (uint32_t)(i / w);
        float fx, fy, nx, ny;
        sobel(src, w, h, x, y, fx, fy, nx, ny);
        V3 lv{a.lx, a.ly, a.lz};
        if (a.kind != RVK_LIGHT_DISTANT) {
            float nz = (float)(src[i] >> 24) / 255.0f * a.surface_scale;
            V3 v{a.lx - (float)x, a.ly - (float)y, a.lz - nz};
            float l = len3(v);
            lv = approx_zero_ulps(l, 4) ? v : V3{v.x / l, v.y / l, v.z / l};
        }
        float cr = a.r, cg = a.g, cb = a.b;
        if (a.kind == RVK_LIGHT_SPOT) {
            float mlds = -dot3(lv, V3{a.dx, a.dy, a.dz});
            bool black = mlds <= 0.0f || (a.has_cone && mlds < a.cos_cone);
            if (black) { cr = cg = cb = 0.0f; }
            else {
                float f = powf(mlds, a.spot_exp);
                cr = (float)lcompute((float)a.cr, f);
                cg = (float)lcompute((float)a.cg, f);
                cb = (float)lcompute((float)a.cb, f);
            }
        }
        bool nzero = approx_zero_ulps(nx, 4) && approx_zero_ulps(ny, 4);
        float factor;
        if (!a.specular) {
            float k;
            if (nzero) k = lv.z;
            else {
                float s = a.surface_scale / 255.0f;
                float px_ = nx * s, py_ = ny * s;
                px_ *= fx; py_ *= fy;
                V3 nn{px_, py_, 1.0f};
                k = dot3(nn, lv) / len3(nn);
            }
            factor = a.constant * k;
        } else {
            V3 hv{lv.x + 0.0f, lv.y + 0.0f, lv.z + 1.0f};
            float hl = len3(hv);
            if (approx_zero_ulps(hl, 4)) factor = 0.0f;
            else {
                float ndh;
                if (nzero) ndh = hv.z / hl;
                else {
                    float s = a.surface_scale / 255.0f;
                    float px_ = nx * s, py_ = ny * s;
                    px_ *= fx; py_ *= fy;
                    V3 nn{px_, py_, 1.0f};
                    ndh = dot3(nn, hv) / len3(nn) / hl;
                }
                float k = a.exp_is_one ? ndh : powf(ndh, a.spec_exp);
                factor = a.constant * k;
            }
        }
        uint32_t r = lcompute(cr, factor), g = lcompute(cg, factor), b = lcompute(cb, factor);
        uint32_t al = a.specular ? max(max(r, g), b) : 255u;
        dst[i] = pack(r, g, b, al);
    }
}

inline unsigned grid_for(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 65535u * 4u) g = 65535u * 4u;
    return (unsigned)(g ? g : 1);
}

}  // namespace

namespace rvk {

rvk_status launch_convolve(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p,
                           const float *values) {
    const auto &cm = p->u.convolve;
    if (cm.columns == 0 || cm.rows == 0 || cm.target_x >= cm.columns || cm.target_y >= cm.rows ||
        cm.edge_mode > RVK_EDGE_WRAP || !values || cm.divisor == 0.0f)
        return RVK_ERR_INVALID_ARGUMENT;
    ConvArgs a{};
    a.columns = cm.columns; a.rows = cm.rows; a.tx = cm.target_x; a.ty = cm.target_y;
    a.edge = cm.edge_mode; a.preserve_alpha = cm.preserve_alpha;
    a.divisor = cm.divisor; a.bias = cm.bias;
    size_t nk = (size_t)cm.columns * cm.rows;
    std::vector<float> flipped(nk);
    const float *data = values + cm.data_offset;
    for (uint32_t oy = 0; oy < cm.rows; oy++)
        for (uint32_t ox = 0; ox < cm.columns; ox++)
            flipped[oy * cm.columns + ox] = data[(cm.rows - oy - 1) * cm.columns + (cm.columns - ox - 1)];
    void *bigk = nullptr;
    if (nk <= (size_t)kMaxKernel) std::memcpy(a.k, flipped.data(), nk * sizeof(float));
    else {
        RVK_TRY(scratch_alloc(ctx, nk * sizeof(float), &bigk));
        rvk_status ust = upload_async(ctx, bigk, flipped.data(), nk * sizeof(float));
        if (ust != RVK_OK) { scratch_release(ctx, bigk); return ust; }
    }
    void *tmp = nullptr;
    size_t bytes = (size_t)w * h * 4;
    rvk_status st = scratch_alloc(ctx, bytes, &tmp);
    if (st == RVK_OK) {
        timed_launch(ctx, "convolve_matrix", 2.0 * (double)bytes, [&] {
            hipLaunchKernelGGL(k_convolve, dim3(grid_for((size_t)w * h)), dim3(256), 0, ctx->stream,
                               (const uint32_t *)px, (uint32_t *)tmp, w, h, a, (const float *)bigk);
        });
        if (hipGetLastError() != hipSuccess) st = RVK_ERR_HIP;
        if (st == RVK_OK) st = launch_copy(ctx, px, (const uint8_t *)tmp, bytes);
        scratch_release(ctx, tmp);
    }
    if (bigk) scratch_release(ctx, bigk);
    return st;
}

rvk_status launch_lighting(rvk_context *ctx, const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                           uint32_t specular, float surface_scale, float constant, float spec_exp,
                           const uint8_t color[4], const rvk_light_source *light) {
    if (light->kind > RVK_LIGHT_SPOT) return RVK_ERR_INVALID_ARGUMENT;
    if (w < 3 || h < 3) return RVK_OK;  // lighting.rs:236-238
    LightArgs a{};
    a.specular = specular ? 1u : 0u;
    a.kind = light->kind;
    a.surface_scale = surface_scale;
    a.constant = constant;
    a.spec_exp = spec_exp;
    a.r = (float)color[0]; a.g = (float)color[1]; a.b = (float)color[2];
    a.cr = color[0]; a.cg = color[1]; a.cb = color[2];
    // strict-num approx_eq_ulps(1.0, 4)
    {
        uint32_t ub, uo;
        float one = 1.0f;
        std::memcpy(&ub, &spec_exp, 4);
        std::memcpy(&uo, &one, 4);
        int32_t d = (int32_t)(ub - uo);
        a.exp_is_one = (spec_exp == 1.0f) || (!std::signbit(spec_exp) && d >= -4 && d <= 4);
    }
    const float rads_per_deg = 3.14159274101257324f / 180.0f;
    if (light->kind == RVK_LIGHT_DISTANT) {
        // fixed light vector (lighting.rs:244-255), host libm cos/sin like the reference
        float az = light->azimuth * rads_per_deg, el = light->elevation * rads_per_deg;
        a.lx = std::cos(az) * std::cos(el);
        a.ly = std::sin(az) * std::cos(el);
        a.lz = std::sin(el);
    } else {
        a.lx = light->x; a.ly = light->y; a.lz = light->z;
    }
    if (light->kind == RVK_LIGHT_SPOT) {
        float dx = light->points_at_x - light->x, dy = light->points_at_y - light->y,
              dz = light->points_at_z - light->z;
        float l = std::sqrt(dx * dx + dy * dy + dz * dz);
        uint32_t ul;
        std::memcpy(&ul, &l, 4);
        bool lz = (l == 0.0f) || (!std::signbit(l) && ul <= 4u);
        if (!lz) { dx /= l; dy /= l; dz /= l; }
        a.dx = dx; a.dy = dy; a.dz = dz;
        a.spot_exp = light->specular_exponent;
        a.has_cone = light->has_limiting_cone_angle;
        a.cos_cone = std::cos(light->limiting_cone_angle * rads_per_deg);
    }
    timed_launch(ctx, "lighting", 8.0 * (double)w * h, [&] {  // source read + destination write
        hipLaunchKernelGGL(k_lighting, dim3(grid_for((size_t)w * h)), dim3(256), 0, ctx->stream,
                           (const uint32_t *)src, (uint32_t *)dst, w, h, a);
    });
    RVK_HIP_CHECK(hipGetLastError());
    return RVK_OK;
}

}  // namespace rvk
