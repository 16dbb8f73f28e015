// IIR Gaussian blur (crates/resvg/src/filter/iir_blur.rs).  The reference runs, per
// channel, 4 steps of a causal then an anti-causal first-order recurrence along every
// row and then along every column, in f64, and truncates (v*255) to u8.  The chains
// are kept sequential (bit-exact order of f64 operations); parallelism comes from
// running every (row, channel) and (column, channel) chain in its own lane, and each
// lane streams its chain a block of values ahead of the recurrence.  The f64 planes
// are laid out so that lanes of a wave touch consecutive addresses at every step:
// column-major for the row sweeps, row-major for the column sweeps.
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

}  // namespace

namespace rvk {

// gen_coefficients (iir_blur.rs:142-146)
static void coefficients(double sigma, double *lambda, double *dnu) {
    double l = (sigma * sigma) / (2.0 * 4.0);
    *lambda = l;
    *dnu = (1.0 + 2.0 * l - std::sqrt(1.0 + 4.0 * l)) / (2.0 * l);
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
