// Context, stream and device-memory management of librvk.
#include "rvk_internal.h"
#include <algorithm>
#include <chrono>
#include <cstdio>

namespace rvk {

static thread_local std::string g_last_error;

void set_last_error(const char *what, const char *where) {
    g_last_error = std::string(what ? what : "") + " at " + (where ? where : "");
    if (std::getenv("RVK_VERBOSE")) std::fprintf(stderr, "[rvk] HIP error: %s\n", g_last_error.c_str());
}

size_t prof_start(rvk_context *ctx, const char *name, double bytes) {
    if (!ctx->profiling) return SIZE_MAX;
    hipEvent_t ev[2];
    for (auto &e : ev) {
        if (!ctx->event_pool.empty()) {
            e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            return SIZE_MAX;
        }
    }
    if (hipEventRecord(ev[0], ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        ctx->event_pool.push_back(ev[0]);
        ctx->event_pool.push_back(ev[1]);
        return SIZE_MAX;
    }
    ctx->prof.push_back(rvk_prof_rec{name, ev[0], ev[1], bytes});
    return ctx->prof.size() - 1;
}

void prof_stop(rvk_context *ctx, size_t slot) {
    if (slot == SIZE_MAX || slot >= ctx->prof.size()) return;
    if (hipEventRecord(ctx->prof[slot].stop, ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        ctx->prof[slot].name = nullptr;  // dropped from the summary
    }
}

rvk_status scratch_alloc(rvk_context *ctx, size_t bytes, void **out) {
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~(size_t)255;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        // best fit among cached buffers no more than 2x the request
        size_t best = SIZE_MAX, bi = 0;
        for (size_t i = 0; i < ctx->free_list.size(); i++) {
            size_t s = ctx->free_list[i].first;
            if (s >= bytes && s <= 2 * bytes && s < best) { best = s; bi = i; }
        }
        if (best != SIZE_MAX) {
            *out = ctx->free_list[bi].second;
            ctx->live.push_back(ctx->free_list[bi]);
            ctx->free_list.erase(ctx->free_list.begin() + (long)bi);
            return RVK_OK;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // drop the cache and retry once; the stream must be idle before freeing
        std::lock_guard<std::mutex> g(ctx->mu);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &f : ctx->free_list) (void)hipFree(f.second);
        ctx->free_list.clear();
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return RVK_ERR_OUT_OF_MEMORY;
        }
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->live.emplace_back(bytes, p);
    ctx->stats.scratch_bytes += bytes;
    *out = p;
    return RVK_OK;
}

// Buffers go back to the context cache at once: all users of a context share one
// stream, so a later user is ordered after every earlier kernel that touched it.
void scratch_release(rvk_context *ctx, void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(ctx->mu);
    for (size_t i = 0; i < ctx->live.size(); i++) {
        if (ctx->live[i].second == p) {
            ctx->free_list.push_back(ctx->live[i]);
            ctx->live.erase(ctx->live.begin() + (long)i);
            return;
        }
    }
}

rvk_status lanes_prepare(rvk_context *ctx, uint32_t n) {
    while (ctx->lanes.size() < n) {
        rvk_context *l = new rvk_context();
        l->device = ctx->device;
        if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            delete l;
            return RVK_ERR_HIP;
        }
        l->own_stream = true;
        l->is_lane = true;
        ctx->lanes.push_back(l);
    }
    for (uint32_t i = 0; i < n; i++) {
        ctx->lanes[i]->stats = rvk_stats{};
        ctx->lanes[i]->profiling = ctx->profiling;
    }
    return RVK_OK;
}

void lanes_collect(rvk_context *ctx, uint32_t n) {
    for (uint32_t i = 0; i < n && i < ctx->lanes.size(); i++) {
        rvk_context *l = ctx->lanes[i];
        rvk_stats &a = ctx->stats;
        const rvk_stats &b = l->stats;
        a.kernel_launches += b.kernel_launches;
        a.layers += b.layers;
        a.paths += b.paths;
        a.filters += b.filters;
        a.scratch_bytes += b.scratch_bytes;
        a.edge_pieces += b.edge_pieces;
        a.band_entries += b.band_entries;
        a.upload_bytes += b.upload_bytes;
        a.flushes += b.flushes;
        a.host_geom_ms += b.host_geom_ms;  // host_wait_ms of a lane is its own thread's
        for (auto &r : l->prof) ctx->prof.push_back(r);
        l->prof.clear();
        l->profiling = false;
    }
}

rvk_status upload_sync(rvk_context *ctx, void *dev_dst, const void *host_src, size_t bytes) {
    RVK_HIP_CHECK(hipMemcpyAsync(dev_dst, host_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    RVK_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return RVK_OK;
}

rvk_status staging_acquire(rvk_context *ctx, size_t bytes, void **host, uint32_t *slot) {
    const uint32_t k = ctx->staging_next;
    ctx->staging_next = (k + 1) % 3;
    rvk_staging &sb = ctx->staging[k];
    if (!sb.done) RVK_HIP_CHECK(hipEventCreateWithFlags(&sb.done, hipEventDisableTiming));
    if (sb.pending) {
        const auto t0 = std::chrono::steady_clock::now();
        RVK_HIP_CHECK(hipEventSynchronize(sb.done));
        ctx->stats.host_wait_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        sb.pending = false;
    }
    if (sb.cap < bytes) {
        if (sb.host) RVK_HIP_CHECK(hipHostFree(sb.host));
        sb.host = nullptr;
        sb.cap = 0;
        size_t cap = 1u << 20;
        while (cap < bytes) cap <<= 1;
        if (hipHostMalloc(&sb.host, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            sb.host = nullptr;
            return RVK_ERR_OUT_OF_MEMORY;
        }
        sb.cap = cap;
    }
    *host = sb.host;
    *slot = k;
    return RVK_OK;
}

rvk_status upload_staged(rvk_context *ctx, uint32_t slot, void *dev_dst, size_t bytes) {
    return upload_staged_on(ctx, slot, dev_dst, bytes, ctx->stream);
}

rvk_status upload_staged_on(rvk_context *ctx, uint32_t slot, void *dev_dst, size_t bytes, hipStream_t stream) {
    rvk_staging &sb = ctx->staging[slot];
    RVK_HIP_CHECK(hipMemcpyAsync(dev_dst, sb.host, bytes, hipMemcpyHostToDevice, stream));
    RVK_HIP_CHECK(hipEventRecord(sb.done, stream));
    sb.pending = true;
    return RVK_OK;
}

rvk_status side_stream(rvk_context *ctx, hipStream_t *out) {
    if (!ctx->side) RVK_HIP_CHECK(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    *out = ctx->side;
    return RVK_OK;
}

rvk_status upload_async(rvk_context *ctx, void *dev_dst, const void *host_src, size_t bytes) {
    void *h = nullptr;
    uint32_t slot = 0;
    RVK_TRY(staging_acquire(ctx, bytes, &h, &slot));
    std::memcpy(h, host_src, bytes);
    return upload_staged(ctx, slot, dev_dst, bytes);
}

}  // namespace rvk

extern "C" {

const char *rvk_status_string(rvk_status s) {
    switch (s) {
        case RVK_OK: return "ok";
        case RVK_ERR_INVALID_ARGUMENT: return "invalid argument";
        case RVK_ERR_OUT_OF_MEMORY: return "out of device memory";
        case RVK_ERR_HIP: return "HIP runtime error";
        case RVK_ERR_INVALID_REGION: return "invalid filter region";
        case RVK_ERR_UNSUPPORTED: return "unsupported feature";
        case RVK_ERR_NO_DEVICE: return "no HIP device";
        case RVK_ERR_TOO_LARGE: return "input larger than the entry point's fixed limit";
        default: return "unknown status";
    }
}

rvk_status rvk_context_create(int device, rvk_context **out) {
    if (!out) return RVK_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return RVK_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) return RVK_ERR_INVALID_ARGUMENT;
    RVK_HIP_CHECK(hipSetDevice(device));
    // the no-allocation entry point's device array: its address on this device
    RVK_TRY(rvk::noalloc_init(device));
    rvk_context *ctx = new rvk_context();
    ctx->device = device;
    if (hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return RVK_ERR_HIP;
    }
    ctx->own_stream = true;
    *out = ctx;
    return RVK_OK;
}

rvk_status rvk_context_destroy(rvk_context *ctx) {
    if (!ctx) return RVK_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    for (rvk_context *l : ctx->lanes) (void)rvk_context_destroy(l);
    ctx->lanes.clear();
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->prof) { (void)hipEventDestroy(r.start); (void)hipEventDestroy(r.stop); }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto &sb : ctx->staging) {
        if (sb.host) (void)hipHostFree(sb.host);
        if (sb.done) (void)hipEventDestroy(sb.done);
    }
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    for (auto &f : ctx->free_list) (void)hipFree(f.second);
    for (auto &f : ctx->live) (void)hipFree(f.second);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RVK_OK;
}

rvk_status rvk_context_set_stream(rvk_context *ctx, void *hip_stream) {
    if (!ctx) return RVK_ERR_INVALID_ARGUMENT;
    RVK_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    ctx->stream = (hipStream_t)hip_stream;
    ctx->own_stream = false;
    return RVK_OK;
}

void *rvk_context_stream(rvk_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

rvk_status rvk_synchronize(rvk_context *ctx) {
    if (!ctx) return RVK_ERR_INVALID_ARGUMENT;
    RVK_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return RVK_OK;
}

rvk_status rvk_device_alloc(rvk_context *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return RVK_ERR_INVALID_ARGUMENT;
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return RVK_ERR_OUT_OF_MEMORY;
    }
    return RVK_OK;
}

rvk_status rvk_device_free(rvk_context *ctx, void *p) {
    if (!ctx) return RVK_ERR_INVALID_ARGUMENT;
    RVK_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    RVK_HIP_CHECK(hipFree(p));
    return RVK_OK;
}

rvk_status rvk_upload(rvk_context *ctx, void *dev_dst, const void *host_src, size_t bytes) {
    if (!ctx || (!dev_dst && bytes) || (!host_src && bytes)) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::upload_sync(ctx, dev_dst, host_src, bytes);
}

rvk_status rvk_download(rvk_context *ctx, void *host_dst, const void *dev_src, size_t bytes) {
    if (!ctx || (!dev_src && bytes) || (!host_dst && bytes)) return RVK_ERR_INVALID_ARGUMENT;
    RVK_HIP_CHECK(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    RVK_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return RVK_OK;
}

rvk_status rvk_get_stats(rvk_context *ctx, rvk_stats *out) {
    if (!ctx || !out) return RVK_ERR_INVALID_ARGUMENT;
    *out = ctx->stats;
    return RVK_OK;
}

rvk_status rvk_profile_begin(rvk_context *ctx) {
    if (!ctx) return RVK_ERR_INVALID_ARGUMENT;
    ctx->profiling = true;
    return RVK_OK;
}

rvk_status rvk_profile_end(rvk_context *ctx, rvk_kernel_stat *out, uint32_t capacity, uint32_t *count) {
    if (!ctx || (!out && capacity) || !count) return RVK_ERR_INVALID_ARGUMENT;
    ctx->profiling = false;
    RVK_HIP_CHECK(hipSetDevice(ctx->device));
    RVK_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<rvk_kernel_stat> agg;
    rvk_status st = RVK_OK;
    for (auto &r : ctx->prof) {
        float ms = 0.0f;
        if (r.name && st == RVK_OK) {
            hipError_t e = hipEventElapsedTime(&ms, r.start, r.stop);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                rvk::set_last_error(hipGetErrorString(e), "hipEventElapsedTime");
                st = RVK_ERR_HIP;
            } else {
                auto it = std::find_if(agg.begin(), agg.end(),
                                       [&](const rvk_kernel_stat &k) { return std::strcmp(k.name, r.name) == 0; });
                if (it == agg.end()) {
                    rvk_kernel_stat k{};
                    std::snprintf(k.name, sizeof(k.name), "%s", r.name);
                    agg.push_back(k);
                    it = agg.end() - 1;
                }
                it->launches++;
                it->total_ms += ms;
                it->bytes += r.bytes;
            }
        }
        ctx->event_pool.push_back(r.start);
        ctx->event_pool.push_back(r.stop);
    }
    ctx->prof.clear();
    std::sort(agg.begin(), agg.end(),
              [](const rvk_kernel_stat &a, const rvk_kernel_stat &b) { return a.total_ms > b.total_ms; });
    *count = (uint32_t)std::min<size_t>(agg.size(), capacity);
    for (uint32_t i = 0; i < *count; i++) out[i] = agg[i];
    return st;
}

}  // extern "C"
