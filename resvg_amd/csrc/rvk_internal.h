// Internal declarations of the gfx950 pixel pipeline library (librvk.so).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/resvg_amd.h"

#define RVK_HIP_CHECK(expr)                                         \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) {                                     \
            rvk::set_last_error(hipGetErrorString(e_), #expr);      \
            return RVK_ERR_HIP;                                     \
        }                                                           \
    } while (0)

#define RVK_TRY(expr)                         \
    do {                                      \
        rvk_status s_ = (expr);               \
        if (s_ != RVK_OK) return s_;          \
    } while (0)

// One profiled launch: events around it on the context stream.
struct rvk_prof_rec {
    const char *name;
    hipEvent_t start, stop;
    double bytes;
};

// A pinned host staging buffer; `done` is recorded after the copy that reads it.
struct rvk_staging {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};

struct rvk_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    rvk_stats stats{};
    bool profiling = false;
    std::vector<rvk_prof_rec> prof;
    std::vector<hipEvent_t> event_pool;
    rvk_staging staging[3];  // ring used by upload_async
    uint32_t staging_next = 0;
    // Simple size-bucketed cache of device scratch buffers (layers, filter images).
    std::vector<std::pair<size_t, void *>> free_list;
    std::vector<std::pair<size_t, void *>> live;
    std::mutex mu;
    // Sub-contexts with their own streams, scratch caches and staging rings on the
    // same device; rvk_render_batch spreads documents over them.
    std::vector<rvk_context *> lanes;
    // A second stream (default priority, created on first use) that a render's
    // coverage work runs on, so that it overlaps the previous flush's compositing on
    // `stream`; every render joins it back into `stream` before it returns.
    hipStream_t side = nullptr;
    bool is_lane = false;  // a batch lane (renders on its own stream only)
};

namespace rvk {

void set_last_error(const char *what, const char *where);

// Scratch device memory owned by the context (returned to its cache on release).
rvk_status scratch_alloc(rvk_context *ctx, size_t bytes, void **out);
void scratch_release(rvk_context *ctx, void *p);

// Copies host memory to the device on the context stream and waits for it.
rvk_status upload_sync(rvk_context *ctx, void *dev_dst, const void *host_src, size_t bytes);
// A pinned staging buffer of at least `bytes` from the context's ring (waits until
// the copy that last used it has finished); fill it, then call upload_staged.
rvk_status staging_acquire(rvk_context *ctx, size_t bytes, void **host, uint32_t *slot);
// Enqueues the copy of `bytes` of staging slot `slot` to `dev_dst` (no host wait).
rvk_status upload_staged(rvk_context *ctx, uint32_t slot, void *dev_dst, size_t bytes);
// upload_staged on `stream` instead of the context stream.
rvk_status upload_staged_on(rvk_context *ctx, uint32_t slot, void *dev_dst, size_t bytes, hipStream_t stream);
// The context's side stream (see rvk_context::side), created on first use.
rvk_status side_stream(rvk_context *ctx, hipStream_t *out);
// Copies host memory to the device through the staging ring without waiting.
rvk_status upload_async(rvk_context *ctx, void *dev_dst, const void *host_src, size_t bytes);

// Makes `n` lanes available (created on first use with non-blocking streams of the
// default priority), resets their statistics and mirrors the profiling switch.
rvk_status lanes_prepare(rvk_context *ctx, uint32_t n);
// Adds the statistics of the first `n` lanes to the context's and moves their
// profiled launches into the context's records.
void lanes_collect(rvk_context *ctx, uint32_t n);

static inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Profiling hooks around one kernel launch (no-ops unless rvk_profile_begin ran).
size_t prof_start(rvk_context *ctx, const char *name, double bytes);
void prof_stop(rvk_context *ctx, size_t slot);

// Launches `fn` (a lambda holding one hipLaunchKernelGGL) on the context stream,
// counting it and, while profiling, timing it with HIP events.
template <class F>
static inline void timed_launch(rvk_context *ctx, const char *name, double bytes, F &&fn) {
    const size_t slot = prof_start(ctx, name, bytes);
    fn();
    prof_stop(ctx, slot);
    ctx->stats.kernel_launches++;
}

// Upper bound on one pixmap side: tiny-skia's Pixmap size limit is i32::MAX / 4 bytes per row.
constexpr uint32_t kMaxSide = 1u << 16;

static inline bool valid_dims(uint32_t w, uint32_t h) {
    return w > 0 && h > 0 && w <= kMaxSide && h <= kMaxSide;
}

// ---- kernel launchers (defined in the .hip files) ----
rvk_status launch_into_cs(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, int to_linear);
rvk_status launch_demultiply(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h);
rvk_status launch_multiply(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h);
rvk_status launch_box_blur(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy);
// rvk_box_blur_noalloc: the back buffer is a file-scope device array of this size
constexpr size_t kNoAllocBoxBlurBytes = (size_t)64 << 20;
rvk_status noalloc_init(int device);  // looks the array up on `device` (current)
rvk_status launch_box_blur_noalloc(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy);
rvk_status launch_iir_blur(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy);
rvk_status launch_morphology(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, uint32_t op, float rx,
                             float ry);
// src -> dst, distinct buffers
rvk_status launch_morphology_to(rvk_context *ctx, const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                uint32_t op, float rx, float ry);
rvk_status launch_convolve(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p,
                           const float *values);
rvk_status launch_color_matrix(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p,
                               int fuse_alpha);
rvk_status launch_component_transfer(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h,
                                     const rvk_primitive *p, const float *values, int fuse_alpha);
rvk_status launch_arithmetic(rvk_context *ctx, const uint8_t *s1, const uint8_t *s2, uint8_t *dst, size_t n,
                             float k1, float k2, float k3, float k4);
rvk_status launch_turbulence(rvk_context *ctx, uint8_t *dst, uint32_t w, uint32_t h, double ox, double oy,
                             double sx, double sy, double bfx, double bfy, uint32_t octaves, int32_t seed,
                             uint32_t stitch, uint32_t fractal, int premultiply);
rvk_status launch_lighting(rvk_context *ctx, const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                           uint32_t specular, float surface_scale, float constant, float spec_exp,
                           const uint8_t color[4], const rvk_light_source *light);
rvk_status launch_displacement(rvk_context *ctx, const uint8_t *src, const uint8_t *map, uint8_t *dst,
                               uint32_t w, uint32_t h, float sx, float sy, float scale, uint32_t xch,
                               uint32_t ych);
rvk_status launch_draw_pixmap(rvk_context *ctx, uint8_t *dst, uint32_t dw, uint32_t dh, int32_t x, int32_t y,
                              const uint8_t *src, uint32_t sw, uint32_t sh, float opacity, uint32_t mode);
rvk_status launch_fill(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const uint8_t premul[4]);
rvk_status launch_fill_sub(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, int64_t x0, int64_t y0,
                           int64_t x1, int64_t y1, uint32_t value);
rvk_status launch_copy(rvk_context *ctx, uint8_t *dst, const uint8_t *src, size_t bytes);
// Fills a w x h image with the periodic repeat of the tw x th tile at (tx0, ty0)
// of a sw x sh source, phase (ox, oy): feTile.
rvk_status launch_tile(rvk_context *ctx, uint8_t *dst, uint32_t w, uint32_t h, const uint8_t *src, uint32_t sw,
                       uint32_t sh, uint32_t tx0, uint32_t ty0, uint32_t tw, uint32_t th, uint32_t ox, uint32_t oy);
rvk_status launch_zero_rgb(rvk_context *ctx, uint8_t *px, size_t n);
// px = div255(px * m) per channel, m from `src` (0 alpha, 1 inverted alpha, 2 luminance)
rvk_status launch_apply_mask(rvk_context *ctx, uint8_t *px, const uint8_t *src, size_t n, uint32_t mode);
// tiny_skia::Mask operations on u8 planes of n bytes (mask 4-byte, pixmap 16-byte aligned)
rvk_status launch_mask_from_pixmap(rvk_context *ctx, const uint8_t *px, uint8_t *mask, size_t n, uint32_t type);
rvk_status launch_mask_invert(rvk_context *ctx, uint8_t *mask, size_t n);
rvk_status launch_apply_mask_u8(rvk_context *ctx, uint8_t *px, const uint8_t *mask, size_t n);
rvk_status launch_fill_u8(rvk_context *ctx, uint8_t *p, size_t n, uint8_t value);

}  // namespace rvk
