// resvg::render / resvg_render on the device (crates/resvg/src/lib.rs:34-43,
// crates/c-api/lib.rs:874-897).
//
// The render walk restates render.rs:10-143 and path.rs:6-24 on the host; pixel
// work is recorded into per-target draw lists and executed by two kernels per
// flush (coverage, then fused shade/composite).  A group that isolates only for
// opacity or blend mode and holds a single path is not materialised: its layer
// pixel is formed in registers and composited like draw_pixmap would (the layer
// is transparent wherever the path has no coverage, and every group blend mode
// leaves the destination unchanged under a transparent source).  Groups with
// filters, or with several children, get a real layer, the filter graph, and a
// draw_pixmap into their parent.
#include "rvk_internal.h"
#include "host_geom.h"
#include "hairline.h"
#include "raster.h"
#include "stroker.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>

using namespace rvkh;
using namespace rvkr;

namespace rvk {
rvk_status launch_expand(rvk_context *ctx, const EdgeRec *edges, uint32_t nedges, const EdgeCov *cov,
                         PieceDev *pieces, uint32_t *npieces, int32_t *overflow, double algo_bytes);
rvk_status launch_coverage(rvk_context *ctx, const CovJob *jobs, uint32_t njobs, const uint2 *bands128,
                           uint32_t n128, const uint2 *bands256, uint32_t n256, const uint2 *bands512,
                           uint32_t n512, const EdgeCov *edges, const uint2 *windows, const PieceDev *pieces,
                           const uint32_t *npieces, uint16_t *table, uint8_t *arena, int32_t *overflow,
                           double algo_bytes);
rvk_status launch_draw_items(rvk_context *ctx, uint8_t *target, uint32_t w, uint32_t h, int32_t x0, int32_t y0,
                             int32_t x1, int32_t y1, const DrawItem *items, const DrawCull *cull, uint32_t nitems,
                             const uint16_t *table, const uint8_t *arena, const ShaderParams *shaders,
                             const uint32_t *hair_tiles, const uint32_t *hair_bins, const HairLine *hair_lines,
                             bool patterns, double algo_bytes);
}

namespace rvk {
// What a render walks: the root group's children (resvg::render), one node
// (resvg::render_node), the fills of the root group's children drawn with
// BlendMode::Clear (a clip path's children, clip.rs:30-70), or the children of one
// group (feImage content, filter/mod.rs:878).
enum class Walk { kRoot, kNode, kClearFills, kChildren };
rvk_status render_walk(rvk_context *ctx, const rvk_scene *scene, rvkh::Ts ts, uint32_t w, uint32_t h, uint8_t *px,
                       uint64_t *overflow_rows, Walk walk, rvk_node_ref node, const rvkh::IRect *max_bbox = nullptr);
}  // namespace rvk

namespace {

// Coverage bytes per flush (1 GiB: with a flush's coverage overlapping the previous
// flush's compositing, the 4096x4096 full-filter document renders in 48 flushes at
// 152.7 ms, against 162-156 ms at 0.5, 2 and 4 GiB and 221 ms at 256 MiB; the same
// pixels).  RVK_FLUSH_BYTES overrides it, which tests use to drive many flushes
// through small documents.
uint64_t arena_budget() {
    if (const char *e = std::getenv("RVK_FLUSH_BYTES")) {
        const long long v = std::atoll(e);
        if (v > 0) return (uint64_t)v;
    }
    return 1ull << 30;
}

struct DrawList {
    uint8_t *px = nullptr;
    uint32_t w = 0, h = 0;
    std::vector<DrawItem> items;
    std::vector<JobPrep> jobs;
    std::vector<HairPrep> hair;       // hairline strokes
    std::vector<ShaderParams> shaders;
    uint64_t n_edges = 0, n_windows = 0, piece_slots = 0;
    uint64_t n_groups[3] = {0, 0, 0};  // bands per coverage kernel variant
    uint64_t n_table = 0;              // tile-table entries
    uint64_t n_hair_tiles = 0, n_hair_bins = 0, n_hair_lines = 0;
    uint64_t arena = 0;
    int32_t bx0 = INT32_MAX, by0 = INT32_MAX, bx1 = INT32_MIN, by1 = INT32_MIN;
};

uint32_t convert_blend_mode(uint32_t m) {  // render.rs:145-164
    static const uint32_t map[RVK_MIX_COUNT] = {
        RVK_BLEND_SOURCE_OVER, RVK_BLEND_MULTIPLY, RVK_BLEND_SCREEN, RVK_BLEND_OVERLAY, RVK_BLEND_DARKEN,
        RVK_BLEND_LIGHTEN, RVK_BLEND_COLOR_DODGE, RVK_BLEND_COLOR_BURN, RVK_BLEND_HARD_LIGHT, RVK_BLEND_SOFT_LIGHT,
        RVK_BLEND_DIFFERENCE, RVK_BLEND_EXCLUSION, RVK_BLEND_HUE, RVK_BLEND_SATURATION, RVK_BLEND_COLOR,
        RVK_BLEND_LUMINOSITY};
    return m < RVK_MIX_COUNT ? map[m] : RVK_BLEND_SOURCE_OVER;
}

bool int_rect(int64_t x, int64_t y, int64_t w, int64_t h, IRect &out) {
    if (w <= 0 || h <= 0 || x < INT32_MIN || y < INT32_MIN || x + w > INT32_MAX || y + h > INT32_MAX) return false;
    out = IRect{(int32_t)x, (int32_t)y, (int32_t)w, (int32_t)h};
    return true;
}

IRect max_filter_bbox(uint32_t w, uint32_t h) {  // lib.rs:86-97
    int64_t x = -2 * (int64_t)std::min<uint32_t>(w, INT32_MAX), y = -2 * (int64_t)std::min<uint32_t>(h, INT32_MAX);
    IRect r;
    if (int_rect(x, y, std::min<int64_t>(5 * (int64_t)w, UINT32_MAX), std::min<int64_t>(5 * (int64_t)h, UINT32_MAX), r))
        return r;
    return IRect{INT32_MIN / 2, INT32_MIN / 2, INT32_MAX / 2 - INT32_MIN / 2, INT32_MAX / 2 - INT32_MIN / 2};
}

struct PathGeom {
    JobPrep fill, stroke;
    bool hair = false;               // the stroke is an anti-aliased hairline
    float hair_opacity = 1.0f;       // folded into its paint
    HairPrep hair_job;
};

// Device-space geometry of one path drawn into a pw x ph pixmap whose visible
// pixels are `clip` and whose origin sits at target pixel (lx, ly): tiny-skia's
// fill_path preconditions and edges, and (unless `fill_only`) the stroke outline's
// edges or, for a stroke at most one device pixel wide, its hairline segments, each
// prepared as a job.
void compute_geom(const rvk_scene &sc, const rvk_path &p, Ts ts, uint32_t pw, uint32_t ph, bool fill_only,
                  IRect clip, int32_t lx, int32_t ly, PathGeom &g) {
    auto finish = [&](std::vector<Seg> &segs, uint32_t rule, JobPrep &out) {
        if (!segs_drawable(segs) || bounds_nearly_zero(segs)) return;
        EdgeSet es;
        float mnx = INFINITY, mxx = -INFINITY;
        for (const Seg &s : segs)
            for (int i = 0; i < s.n; i++) { mnx = std::min(mnx, s.p[i].x); mxx = std::max(mxx, s.p[i].x); }
        if (segs_inside_pixmap(segs, p.anti_alias != 0, pw, ph)) {
            build_edges(segs, p.anti_alias ? 2 : 0, es.edges, es.caps);
        } else {
            std::vector<Seg> clipped;
            clip_segments(segs, pw, ph, clipped);
            build_edges(clipped, p.anti_alias ? 2 : 0, es.edges, es.caps, true);
        }
        es.mnx = mnx;
        es.mxx = mxx;
        es.ok = !es.edges.empty();
        prep_job(es, rule, p.anti_alias != 0, clip, lx, ly, out);
    };
    if (!p.visible) return;
    if (p.fill.kind != RVK_PAINT_NONE && p.bounds.w != 0.0f && p.bounds.h != 0.0f) {
        std::vector<Seg> segs;
        fill_segments(sc, p, ts, segs);
        finish(segs, p.fill_rule, g.fill);
    }
    if (!fill_only && p.stroke.kind != RVK_PAINT_NONE && p.stroke_width >= 0.0f) {
        float cov = 1.0f;
        if (rvks::treat_as_hairline(p.stroke_width, p.anti_alias != 0, ts.sx, ts.ky, ts.kx, ts.sy, cov)) {
            g.hair = true;
            g.hair_opacity = rvks::hairline_paint_opacity(cov);
            std::vector<HairLine> lines;
            hairline_lines(sc, p, ts, pw, ph, lines);
            prep_hair(std::move(lines), clip, g.hair_job);
            return;
        }
        std::vector<Seg> segs;
        stroke_segments(sc, p, ts, segs);
        finish(segs, RVK_FILL_NONZERO, g.stroke);
    }
}

unsigned host_threads() {
    static const unsigned n = [] {
        if (const char *e = std::getenv("RVK_HOST_THREADS")) {
            const int v = std::atoi(e);
            if (v > 0) return (unsigned)std::min(v, 64);
        }
        const unsigned hc = std::thread::hardware_concurrency();
        return std::max(1u, std::min(hc ? hc : 1u, 16u));
    }();
    return n;
}

// Host worker threads shared by every context: parallel_for(n, fn) runs fn(0..n-1)
// on the calling thread and the workers.  One parallel_for runs at a time; a call
// made while another is running (another context on another thread) runs serially.
class Pool {
  public:
    Pool() {
        const unsigned n = host_threads();
        for (unsigned k = 1; k < n; k++) threads_.emplace_back([this] { worker(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    void run(size_t n, const std::function<void(size_t)> &fn) {
        std::unique_lock<std::mutex> busy(run_mu_, std::try_to_lock);
        if (n < 2 || threads_.empty() || !busy.owns_lock()) {
            for (size_t i = 0; i < n; i++) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            active_ = (unsigned)threads_.size();
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return active_ == 0; });
        fn_ = nullptr;
    }

  private:
    void work() {
        for (size_t i; (i = next_.fetch_add(1)) < n_;) (*fn_)(i);
    }
    void worker() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> g(mu_);
        for (;;) {
            cv_.wait(g, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            g.unlock();
            work();
            g.lock();
            if (--active_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    unsigned active_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

void parallel_for(size_t n, const std::function<void(size_t)> &fn) {
    static Pool pool;
    pool.run(n, fn);
}

struct Renderer {
    rvk_context *ctx;
    const rvk_scene &sc;
    IRect max_bbox;
    rvk_status err = RVK_OK;
    uint64_t overflow_rows = 0;
    int32_t *dev_overflow = nullptr;  // per render: crossings or rows the kernels could not hold
    const uint64_t budget = arena_budget();
    bool parallel_pack = true;  // flush packs large draw lists on the host thread pool

    // Flushes alternate between two device blocks.  A flush's upload, curve expansion
    // and coverage run on the context's side stream and its compositing on the
    // context stream, which waits for them; the side stream waits until the draw of
    // the flush before last, the previous reader of the block, is done (or, for a
    // block just taken from the scratch cache, until the context stream's work so far
    // is), so the coverage of a flush overlaps the compositing of the one before.
    void *blk[2] = {nullptr, nullptr};
    size_t blk_cap[2] = {0, 0};
    hipEvent_t blk_free[2] = {nullptr, nullptr};  // recorded on the context stream after a block's draw
    bool blk_used[2] = {false, false};
    hipEvent_t ev_cov = nullptr, ev_main = nullptr, ev_side = nullptr;
    uint64_t nflush = 0;
    bool side_used = false;
    // false: a flush runs on the context stream alone (on a batch's lanes, which
    // already run documents side by side)
    const bool overlap = !ctx->is_lane;

    ~Renderer() { (void)finish_flushes(); }

    // Joins the side stream into the context stream and returns the blocks to the
    // scratch cache (the context stream's later work is ordered after every use).
    rvk_status finish_flushes() {
        rvk_status st = RVK_OK;
        if (side_used) {
            if (hipEventRecord(ev_side, ctx->side) != hipSuccess ||
                hipStreamWaitEvent(ctx->stream, ev_side, 0) != hipSuccess) {
                (void)hipGetLastError();
                (void)hipStreamSynchronize(ctx->side);  // the join failed: wait for the side stream
                st = RVK_ERR_HIP;
            }
            side_used = false;
        }
        for (int b = 0; b < 2; b++) {
            if (blk[b]) rvk::scratch_release(ctx, blk[b]);
            blk[b] = nullptr;
            blk_cap[b] = 0;
            blk_used[b] = false;
        }
        for (hipEvent_t *e : {&blk_free[0], &blk_free[1], &ev_cov, &ev_main, &ev_side})
            if (*e) {
                (void)hipEventDestroy(*e);
                *e = nullptr;
            }
        return st;
    }

    rvk_status flush_events() {
        for (hipEvent_t *e : {&blk_free[0], &blk_free[1], &ev_cov, &ev_main, &ev_side})
            if (!*e) RVK_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
        return RVK_OK;
    }

    // The walk runs twice: kCollect records every path drawn (with its device
    // transform), the geometry is computed in parallel, and kExecute replays the
    // walk consuming it in the same order (kSkip steps over a skipped subtree).
    enum Mode { kCollect, kExecute, kSkip } mode = kExecute;
    struct Task {
        const rvk_scene *scene;
        const rvk_path *path;
        Ts ts;
        uint32_t pw, ph;  // size of the pixmap drawn into
        bool fill_only;
        IRect clip;       // its visible pixels
        int32_t lx, ly;   // the pixmap's origin in the target
    };
    std::vector<Task> tasks;
    std::vector<PathGeom> geoms;
    size_t next_task = 0, ready_end = 0;
    static constexpr size_t kChunk = 256;

    PathGeom take_geom() {
        const size_t k = next_task++;
        if (k >= tasks.size()) return PathGeom{};
        if (k >= ready_end) {
            const auto t0 = std::chrono::steady_clock::now();
            const size_t end = std::min(tasks.size(), k + kChunk);
            parallel_for(end - k, [&](size_t i) {
                const Task &t = tasks[k + i];
                compute_geom(*t.scene, *t.path, t.ts, t.pw, t.ph, t.fill_only, t.clip, t.lx, t.ly, geoms[k + i]);
            });
            ready_end = end;
            ctx->stats.host_geom_ms +=
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return std::move(geoms[k]);
    }

    // Places a prepared coverage job in the draw list; returns its index or -1 when
    // it covers nothing.
    int32_t add_job(DrawList &dl, JobPrep &P) {
        if (P.too_wide) fail(RVK_ERR_UNSUPPORTED);
        if (!P.ok) return -1;
        CovJob &j = P.j;
        j.edge_offset = (uint32_t)dl.n_edges;
        dl.n_edges += j.edge_count;
        j.window_offset = (uint32_t)dl.n_windows;
        dl.n_windows += P.windows.size();
        P.piece_base = dl.piece_slots;
        dl.piece_slots += P.pieces;
        for (int t = 0; t < 3; t++) dl.n_groups[t] += P.tier_count[t];
        // the slot region holds every tile (at most one slot per tile and band); the
        // coverage kernel writes only the partial ones
        const uint64_t tiles = (uint64_t)j.tiles_x * j.tiles_y;
        j.table_offset = (uint32_t)dl.n_table;
        dl.n_table += tiles;
        j.cov_offset = dl.arena;
        dl.arena += (tiles * 64 + 255) & ~255ull;
        dl.jobs.push_back(std::move(P));
        return (int32_t)dl.jobs.size() - 1;
    }

    // Places a prepared hairline job in the draw list; returns its index or -1 when
    // nothing is drawn.
    int32_t add_hair_job(DrawList &dl, HairPrep &P) {
        if (!P.ok) return -1;
        P.j.tile_offset = (uint32_t)dl.n_hair_tiles;
        dl.n_hair_tiles += P.tiles.size();
        P.bin_base = (uint32_t)dl.n_hair_bins;
        dl.n_hair_bins += P.bins.size();
        P.line_base = (uint32_t)dl.n_hair_lines;
        dl.n_hair_lines += P.lines.size();
        dl.hair.push_back(std::move(P));
        return (int32_t)dl.hair.size() - 1;
    }

    // fill / stroke coverage jobs + shaders of one path drawn with `ts` into a
    // pw x ph pixmap whose visible pixels are `clip`.  `fill_only` draws no stroke;
    // `clear` fills with BlendMode::Clear; `psc` is the scene holding the path (the
    // render's scene unless given).
    void path_jobs(DrawList &dl, const rvk_path &p, Ts ts, uint32_t pw, uint32_t ph, IRect clip, DrawItem &it,
                   bool fill_only = false, bool clear = false, const rvk_scene *psc = nullptr) {
        const rvk_scene &S = psc ? *psc : sc;
        it.fill_job = it.stroke_job = -1;
        it.fill_shader = it.stroke_shader = -1;
        it.stroke_hair = 0;
        it.stroke_first = p.paint_order_stroke_first ? 1 : 0;
        if (mode == kCollect) {
            tasks.push_back(Task{&S, &p, ts, pw, ph, fill_only, clip, it.lx, it.ly});
            return;
        }
        if (mode == kSkip) {
            next_task++;
            return;
        }
        PathGeom g = take_geom();
        if (!p.visible) return;
        ShaderParams sp;
        // path::fill_path (path.rs:26-75)
        if (g.fill.ok && paint_shader(S, p.fill, ts, sp)) {
            it.fill_job = add_job(dl, g.fill);
            if (it.fill_job >= 0) {
                sp.clear = clear ? 1 : 0;
                it.fill_shader = (int32_t)dl.shaders.size();
                dl.shaders.push_back(sp);
            }
        }
        if (fill_only) return;
        // path::stroke_path (path.rs:77-116): a hairline, or tiny-skia's stroker outline
        // filled nonzero
        if (g.hair) {
            if (paint_shader(S, p.stroke, ts, sp, g.hair_opacity)) {
                it.stroke_job = add_hair_job(dl, g.hair_job);
                if (it.stroke_job >= 0) {
                    it.stroke_hair = 1;
                    it.stroke_shader = (int32_t)dl.shaders.size();
                    dl.shaders.push_back(sp);
                }
            }
        } else if (g.stroke.ok && paint_shader(S, p.stroke, ts, sp)) {
            it.stroke_job = add_job(dl, g.stroke);
            if (it.stroke_job >= 0) {
                it.stroke_shader = (int32_t)dl.shaders.size();
                dl.shaders.push_back(sp);
            }
        }
    }

    // The shader of a paint; a pattern first has its tile rendered
    // (render_pattern_pixmap, path.rs:179-205) by a walk of its own over the pattern's
    // group with Transform::from_scale(sx, sy) and this walk's max_bbox, into a
    // transparent device pixmap that lives until the end of this render.
    std::vector<void *> pattern_tiles;
    static thread_local inline int pattern_depth = 0;  // pattern renders in progress on this thread

    // A raster image's pixmap on the device, uploaded once per render (the fill of its
    // aux rect path carries kPaintImage and the image index).
    static constexpr uint32_t kPaintImage = 0x100;
    std::vector<void *> image_tiles;

    bool image_shader(uint32_t index, Ts ts, ShaderParams &sp) {
        if (index >= sc.image_count) return false;
        const rvk_image &im = sc.images[index];
        if (!image_tiles[index]) {
            const size_t bytes = (size_t)im.width * im.height * 4;
            void *tile = nullptr;
            if (rvk::scratch_alloc(ctx, bytes, &tile) != RVK_OK) return false;
            pattern_tiles.push_back(tile);
            const rvk_status st = rvk::upload_async(ctx, tile, im.pixels, bytes);
            if (st != RVK_OK) {
                fail(st);
                return false;
            }
            image_tiles[index] = tile;
        }
        return make_image_shader(im, ts, image_tiles[index], sp);
    }

    bool paint_shader(const rvk_scene &S, const rvk_paint &paint, Ts ts, ShaderParams &sp, float extra = 1.0f) {
        if (paint.kind == kPaintImage) return image_shader(paint.root, ts, sp);
        if (paint.kind != RVK_PAINT_PATTERN) return make_shader(S, paint, ts, sp, extra);
        float sx, sy;
        uint32_t w, h;
        if (paint.root >= S.group_count || pattern_depth >= kMaxNesting || !pattern_tile_size(paint, ts, sx, sy, w, h))
            return false;
        void *tile = nullptr;
        if (rvk::scratch_alloc(ctx, (size_t)w * h * 4, &tile) != RVK_OK) return false;  // Pixmap::new failure
        pattern_tiles.push_back(tile);
        rvk_status st = rvk::launch_fill_sub(ctx, (uint8_t *)tile, w, h, 0, 0, w, h, 0u);
        if (st == RVK_OK) {
            pattern_depth++;
            st = rvk::render_walk(ctx, &S, Ts{sx, 0.0f, 0.0f, sy, 0.0f, 0.0f}, w, h, (uint8_t *)tile, nullptr,
                                  rvk::Walk::kChildren, rvk_node_ref{RVK_NODE_GROUP, paint.root}, &max_bbox);
            pattern_depth--;
        }
        if (st != RVK_OK) {
            if (err == RVK_OK) err = st;
            return false;
        }
        return make_pattern_shader(paint, ts, sx, sy, tile, w, h, sp, extra);
    }

    void push_item(DrawList &dl, DrawItem &it) {
        if (it.fill_job < 0 && it.stroke_job < 0) return;
        // the item's target rect: union of its jobs' rects moved to target space
        int32_t x0 = INT32_MAX, y0 = INT32_MAX, x1 = INT32_MIN, y1 = INT32_MIN;
        for (int k = 0; k < 2; k++) {
            const int32_t j = k == 0 ? it.fill_job : it.stroke_job;
            if (j < 0) continue;
            int32_t jx, jy, jw, jh;
            if (k == 1 && it.stroke_hair) {
                const HairJob &hj = dl.hair[j].j;
                jx = hj.x0; jy = hj.y0; jw = hj.w; jh = hj.h;
            } else {
                const CovJob &jb = dl.jobs[j].j;
                jx = jb.x0; jy = jb.y0; jw = jb.w; jh = jb.h;
            }
            x0 = std::min(x0, jx + it.lx); y0 = std::min(y0, jy + it.ly);
            x1 = std::max(x1, jx + jw + it.lx); y1 = std::max(y1, jy + jh + it.ly);
        }
        it.x0 = std::max(x0, 0); it.y0 = std::max(y0, 0);
        it.x1 = std::min(x1, (int32_t)dl.w); it.y1 = std::min(y1, (int32_t)dl.h);
        if (it.x1 <= it.x0 || it.y1 <= it.y0) return;
        auto cov_ref = [&](int32_t j) {
            JobRef r{};
            if (j < 0) return r;
            const CovJob &jb = dl.jobs[j].j;
            r.x0 = (jb.gx0 + it.lx) / kTile;  // target-aligned, so exact
            r.y0 = (jb.gy0 + it.ly) / kTile;
            r.w = jb.tiles_x;
            r.h = jb.tiles_y;
            r.table = jb.table_offset;
            r.cov_offset = jb.cov_offset;
            return r;
        };
        it.fj = cov_ref(it.fill_job);
        if (it.stroke_hair) {
            const HairJob &hj = dl.hair[it.stroke_job].j;
            it.sj = JobRef{};
            it.sj.x0 = hj.x0; it.sj.y0 = hj.y0; it.sj.w = (uint32_t)hj.w; it.sj.h = (uint32_t)hj.h;
            it.sj.stride = hj.tiles_x;
            it.sj.table = hj.tile_offset;
        } else {
            it.sj = cov_ref(it.stroke_job);
        }
        dl.bx0 = std::min(dl.bx0, it.x0); dl.by0 = std::min(dl.by0, it.y0);
        dl.bx1 = std::max(dl.bx1, it.x1); dl.by1 = std::max(dl.by1, it.y1);
        dl.items.push_back(it);
        ctx->stats.paths++;
    }

    // Checks every index the device will read from a packed flush block: job edge,
    // window, table and arena ranges; each edge's piece slots; the band lists; the
    // items' jobs, shaders and job references; the hairline tiles and bins; pattern
    // tiles.  Host-side and off unless RVK_VALIDATE is set; returns the first
    // violation, or an empty string.
    static bool validating() {
        const char *e = std::getenv("RVK_VALIDATE");  // read per flush: tests switch it on and off
        return e && *e && *e != '0';
    }
    std::string validate_block(const DrawList &dl, const uint8_t *host, size_t o_jobs, size_t o_edges, size_t o_ecov,
                               size_t o_win, const size_t o_grp[3], size_t o_ht, size_t o_hb) const {
        char buf[256];
        auto bad = [&](const char *what, uint64_t i, uint64_t v, uint64_t lim) {
            std::snprintf(buf, sizeof buf, "%s %llu: %llu not below %llu", what, (unsigned long long)i,
                          (unsigned long long)v, (unsigned long long)lim);
            return std::string(buf);
        };
        const size_t njobs = dl.jobs.size(), nhair = dl.hair.size();
        const CovJob *jobs = (const CovJob *)(host + o_jobs);
        const EdgeRec *er = (const EdgeRec *)(host + o_edges);
        const EdgeCov *ec = (const EdgeCov *)(host + o_ecov);
        const uint2 *win = (const uint2 *)(host + o_win);
        for (size_t i = 0; i < njobs; i++) {
            const CovJob &j = jobs[i];
            const uint64_t tiles = (uint64_t)j.tiles_x * j.tiles_y;
            if ((uint64_t)j.edge_offset + j.edge_count > dl.n_edges)
                return bad("job edge end", i, (uint64_t)j.edge_offset + j.edge_count, dl.n_edges + 1);
            if ((uint64_t)j.window_offset + j.tiles_y > dl.n_windows)
                return bad("job window end", i, (uint64_t)j.window_offset + j.tiles_y, dl.n_windows + 1);
            if ((uint64_t)j.table_offset + tiles > dl.n_table)
                return bad("job table end", i, (uint64_t)j.table_offset + tiles, dl.n_table + 1);
            if (j.cov_offset + tiles * 64 > dl.arena) return bad("job arena end", i, j.cov_offset + tiles * 64, dl.arena + 1);
            if (j.tiles_x > kMaxTilesX) return bad("job tiles_x", i, j.tiles_x, kMaxTilesX + 1);
            for (uint32_t g = 0; g < j.tiles_y; g++) {
                const uint2 w = win[j.window_offset + g];
                if (w.x > w.y || w.y > j.edge_count) return bad("job window", i, w.y, j.edge_count + 1);
            }
            // piece slots: increasing offsets inside the job's range
            const uint64_t base = dl.jobs[i].piece_base, end = base + dl.jobs[i].pieces;
            if (end > dl.piece_slots) return bad("job piece end", i, end, dl.piece_slots + 1);
            uint64_t prev = base;
            for (uint32_t k = 0; k < j.edge_count; k++) {
                const EdgeRec &e = er[j.edge_offset + k];
                const EdgeCov &c = ec[j.edge_offset + k];
                if (e.piece_offset < prev || e.piece_offset > end) return bad("edge piece offset", k, e.piece_offset, end + 1);
                if (c.piece_offset != e.piece_offset) return bad("edge piece mirror", k, c.piece_offset, e.piece_offset + 1);
                if ((uint32_t)c.id >= j.edge_count) return bad("edge id", k, (uint32_t)c.id, j.edge_count);
                prev = e.piece_offset;
            }
        }
        for (int t = 0; t < 3; t++) {
            const uint2 *b = (const uint2 *)(host + o_grp[t]);
            for (uint64_t k = 0; k < dl.n_groups[t]; k++) {
                if (b[k].x >= njobs) return bad("band job", k, b[k].x, njobs);
                if (b[k].y >= jobs[b[k].x].tiles_y) return bad("band index", k, b[k].y, jobs[b[k].x].tiles_y);
                if (dl.jobs[b[k].x].tier[b[k].y] != t) return bad("band tier", k, dl.jobs[b[k].x].tier[b[k].y], t);
            }
        }
        const uint32_t *ht = (const uint32_t *)(host + o_ht);
        const uint32_t *hb = (const uint32_t *)(host + o_hb);
        for (uint64_t k = 0; k < dl.n_hair_tiles; k++)
            if (ht[k] > dl.n_hair_bins) return bad("hair tile", k, ht[k], dl.n_hair_bins + 1);
        for (uint64_t k = 0; k < dl.n_hair_bins; k++)
            if (hb[k] >= dl.n_hair_lines) return bad("hair bin", k, hb[k], dl.n_hair_lines);
        for (size_t i = 0; i < dl.items.size(); i++) {
            const DrawItem &it = dl.items[i];
            const size_t ns = dl.shaders.size();
            if (it.fill_job >= 0) {
                if ((size_t)it.fill_job >= njobs) return bad("item fill job", i, (uint64_t)it.fill_job, njobs);
                if (it.fill_shader < 0 || (size_t)it.fill_shader >= ns) return bad("item fill shader", i, (uint64_t)it.fill_shader, ns);
                const CovJob &j = jobs[it.fill_job];
                if (it.fj.table != j.table_offset || it.fj.cov_offset != j.cov_offset || it.fj.w != j.tiles_x ||
                    it.fj.h != j.tiles_y)
                    return bad("item fill ref", i, it.fj.table, j.table_offset + 1);
            } else if (it.fj.w != 0 || it.fj.h != 0) {
                return bad("item empty fill grid", i, it.fj.w, 1);
            }
            if (it.stroke_job >= 0) {
                if (it.stroke_shader < 0 || (size_t)it.stroke_shader >= ns)
                    return bad("item stroke shader", i, (uint64_t)it.stroke_shader, ns);
                if (it.stroke_hair) {
                    if ((size_t)it.stroke_job >= nhair) return bad("item hair job", i, (uint64_t)it.stroke_job, nhair);
                    const HairJob &h = dl.hair[it.stroke_job].j;
                    const uint64_t tend = (uint64_t)h.tile_offset + (uint64_t)h.tiles_x * h.tiles_y + 1;
                    if (tend > dl.n_hair_tiles) return bad("item hair tiles", i, tend, dl.n_hair_tiles + 1);
                    if (it.sj.table != h.tile_offset) return bad("item hair ref", i, it.sj.table, h.tile_offset + 1);
                } else {
                    if ((size_t)it.stroke_job >= njobs) return bad("item stroke job", i, (uint64_t)it.stroke_job, njobs);
                    const CovJob &j = jobs[it.stroke_job];
                    if (it.sj.table != j.table_offset || it.sj.cov_offset != j.cov_offset || it.sj.w != j.tiles_x ||
                        it.sj.h != j.tiles_y)
                        return bad("item stroke ref", i, it.sj.table, j.table_offset + 1);
                }
            }
        }
        for (size_t k = 0; k < dl.shaders.size(); k++) {
            const ShaderHead &s = dl.shaders[k];
            if (s.kind == 3 && (s.tile == 0 || s.tile_w == 0 || s.tile_h == 0)) return bad("pattern tile", k, 0, 1);
            if (s.kind == 2 && !s.two_stop && s.nstops > (uint32_t)kMaxStops + 2) return bad("stops", k, s.nstops, kMaxStops + 3);
        }
        return std::string();
    }

    rvk_status flush(DrawList &dl) {
        if (dl.items.empty()) {
            dl = DrawList{dl.px, dl.w, dl.h};
            return RVK_OK;
        }
        // one device block: jobs | edges | ecov | windows | shaders | items | item cull
        // records | hairline tiles, bins, lines | bands per coverage variant ||
        // npieces | pieces | tile tables | arena (the part after || is written by the
        // device).  The metadata is packed straight into a pinned staging buffer (the
        // jobs in parallel, each moved to its place in the block) and copied without a
        // host wait; device buffers are reused only by later work on the same stream.
        const size_t njobs = dl.jobs.size(), nhair = dl.hair.size(), nitems = dl.items.size();
        auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
        const size_t o_jobs = 0;
        const size_t o_edges = al(o_jobs + njobs * sizeof(CovJob));
        const size_t o_ecov = al(o_edges + dl.n_edges * sizeof(EdgeRec));
        const size_t o_win = al(o_ecov + dl.n_edges * sizeof(EdgeCov));
        const size_t o_sh = al(o_win + dl.n_windows * sizeof(uint2));
        const size_t o_it = al(o_sh + dl.shaders.size() * sizeof(ShaderParams));
        const size_t o_cull = al(o_it + nitems * sizeof(DrawItem));
        const size_t o_ht = al(o_cull + nitems * sizeof(DrawCull));
        const size_t o_hb = al(o_ht + dl.n_hair_tiles * 4);
        const size_t o_hl = al(o_hb + dl.n_hair_bins * 4);
        size_t o_grp[3];
        o_grp[0] = al(o_hl + dl.n_hair_lines * sizeof(HairLine));
        o_grp[1] = al(o_grp[0] + dl.n_groups[0] * sizeof(uint2));
        o_grp[2] = al(o_grp[1] + dl.n_groups[1] * sizeof(uint2));
        const size_t meta = al(o_grp[2] + dl.n_groups[2] * sizeof(uint2));
        const size_t o_np = meta;
        const size_t o_pc = al(o_np + dl.n_edges * 4);
        const size_t o_tab = al(o_pc + (size_t)dl.piece_slots * sizeof(PieceDev));
        const size_t o_arena = al(o_tab + (size_t)dl.n_table * 2);
        void *hostv = nullptr;
        uint32_t slot = 0;
        RVK_TRY(rvk::staging_acquire(ctx, meta, &hostv, &slot));
        uint8_t *host = (uint8_t *)hostv;
        for (ShaderParams &sp : dl.shaders) seal_shader(sp);
        if (!dl.shaders.empty()) std::memcpy(host + o_sh, dl.shaders.data(), dl.shaders.size() * sizeof(ShaderParams));
        std::memcpy(host + o_it, dl.items.data(), nitems * sizeof(DrawItem));
        {
            DrawCull *cull = (DrawCull *)(host + o_cull);
            for (size_t i = 0; i < nitems; i++) {
                const DrawItem &it = dl.items[i];
                DrawCull c{};
                c.tx0 = it.x0 / kTile;
                c.ty0 = it.y0 / kTile;
                c.tx1 = (it.x1 + kTile - 1) / kTile;
                c.ty1 = (it.y1 + kTile - 1) / kTile;
                c.fx0 = it.fj.x0; c.fy0 = it.fj.y0; c.fw = it.fj.w; c.fh = it.fj.h; c.ftable = it.fj.table;
                c.hair = it.stroke_hair;
                if (!it.stroke_hair) {
                    c.sx0 = it.sj.x0; c.sy0 = it.sj.y0; c.sw = it.sj.w; c.sh = it.sj.h; c.stable = it.sj.table;
                }
                cull[i] = c;
            }
        }
        // each job's first entry in the band lists
        std::vector<uint64_t> gbase(njobs * 3);
        {
            uint64_t run[3] = {0, 0, 0};
            for (size_t i = 0; i < njobs; i++)
                for (int t = 0; t < 3; t++) {
                    gbase[i * 3 + t] = run[t];
                    run[t] += dl.jobs[i].tier_count[t];
                }
        }
        auto pack = [&](size_t i) {
            if (i < njobs) {
                const JobPrep &P = dl.jobs[i];
                const CovJob &j = P.j;
                std::memcpy(host + o_jobs + i * sizeof(CovJob), &j, sizeof(CovJob));
                EdgeRec *er = (EdgeRec *)(host + o_edges) + j.edge_offset;
                EdgeCov *ec = (EdgeCov *)(host + o_ecov) + j.edge_offset;
                const uint32_t base = (uint32_t)P.piece_base;
                for (uint32_t k = 0; k < j.edge_count; k++) {
                    EdgeRec e = P.edges[k];
                    e.piece_offset += base;
                    er[k] = e;
                    EdgeCov c = P.ecov[k];
                    c.piece_offset += base;
                    ec[k] = c;
                }
                if (!P.windows.empty())
                    std::memcpy((uint2 *)(host + o_win) + j.window_offset, P.windows.data(),
                                P.windows.size() * sizeof(uint2));
                uint2 *gl[3];
                for (int t = 0; t < 3; t++) gl[t] = (uint2 *)(host + o_grp[t]) + gbase[i * 3 + t];
                for (size_t g = 0; g < P.tier.size(); g++) *gl[P.tier[g]]++ = uint2{(uint32_t)i, (uint32_t)g};
            } else {
                const size_t k = i - njobs;
                const HairPrep &P = dl.hair[k];
                uint32_t *tl = (uint32_t *)(host + o_ht) + P.j.tile_offset;
                for (size_t t = 0; t < P.tiles.size(); t++) tl[t] = P.tiles[t] + P.bin_base;
                uint32_t *bl = (uint32_t *)(host + o_hb) + P.bin_base;
                for (size_t b = 0; b < P.bins.size(); b++) bl[b] = P.bins[b] + P.line_base;
                std::memcpy((HairLine *)(host + o_hl) + P.line_base, P.lines.data(), P.lines.size() * sizeof(HairLine));
            }
        };
        if (parallel_pack && dl.n_edges >= 16384) {
            parallel_for(njobs + nhair, pack);
        } else {
            for (size_t i = 0; i < njobs + nhair; i++) pack(i);
        }
        if (validating()) {
            const std::string why = validate_block(dl, host, o_jobs, o_edges, o_ecov, o_win, o_grp, o_ht, o_hb);
            if (!why.empty()) {
                std::fprintf(stderr, "rvk: flush block invalid: %s\n", why.c_str());
                return RVK_ERR_INVALID_ARGUMENT;
            }
            ctx->stats.validated_flushes++;
        }
        // the flush's block (see blk), replaced by a larger one from the scratch cache
        // when it is too small (the old one's last reader is on the context stream)
        const int b = (int)(nflush++ & 1);
        const size_t need = o_arena + dl.arena;
        RVK_TRY(flush_events());
        bool fresh = false;
        if (blk_cap[b] < need) {
            if (blk[b]) rvk::scratch_release(ctx, blk[b]);
            blk[b] = nullptr;
            blk_cap[b] = 0;
            blk_used[b] = false;
            RVK_TRY(rvk::scratch_alloc(ctx, need, &blk[b]));
            blk_cap[b] = need;
            fresh = true;
        }
        uint8_t *d = (uint8_t *)blk[b];
        // algorithmic bytes: expansion reads each edge once and writes its pieces;
        // coverage reads edges, windows and pieces once and writes the tile tables (the
        // partial tiles' 64 bytes each, written by coverage and read by the compositor,
        // are data-dependent and not counted); the compositor is credited what the
        // reference's per-path fills move: 8 B (destination read + write) per pixel of
        // each fill's, stroke's and hairline's rect, and 12 B per pixel of each
        // single-path layer it composites in registers (draw_pixmap)
        double job_tiles = 0.0, job_px = 0.0, layer_px = 0.0;
        for (const JobPrep &P : dl.jobs) {
            job_tiles += (double)P.j.tiles_x * (double)P.j.tiles_y;
            job_px += (double)P.j.w * (double)P.j.h;
        }
        for (const HairPrep &P : dl.hair) job_px += (double)P.j.w * (double)P.j.h;
        for (const DrawItem &it : dl.items)
            if (it.layer) layer_px += (double)(it.x1 - it.x0) * (double)(it.y1 - it.y0);
        const double piece_bytes = (double)dl.piece_slots * sizeof(PieceDev);
        const double exp_bytes = (double)dl.n_edges * (sizeof(EdgeRec) + 4) + piece_bytes;
        const double cov_bytes = 2.0 * job_tiles + (double)dl.n_edges * sizeof(EdgeCov) +
                                 (double)dl.n_windows * sizeof(uint2) + piece_bytes;
        const double draw_bytes = 8.0 * job_px + 12.0 * layer_px;
        ctx->stats.edge_pieces += dl.piece_slots;
        ctx->stats.band_entries += dl.n_windows;
        ctx->stats.upload_bytes += meta;
        hipStream_t side = ctx->stream;
        if (overlap) {
            RVK_TRY(rvk::side_stream(ctx, &side));
            side_used = true;
            if (fresh) {
                RVK_HIP_CHECK(hipEventRecord(ev_main, ctx->stream));
                RVK_HIP_CHECK(hipStreamWaitEvent(side, ev_main, 0));
            } else if (blk_used[b]) {
                RVK_HIP_CHECK(hipStreamWaitEvent(side, blk_free[b], 0));
            }
        }
        rvk_status st = rvk::upload_staged_on(ctx, slot, d, meta, side);
        {
            // expansion and coverage are launched on the side stream
            const hipStream_t main_stream = ctx->stream;
            ctx->stream = side;
            if (st == RVK_OK)
                st = rvk::launch_expand(ctx, (const EdgeRec *)(d + o_edges), (uint32_t)dl.n_edges,
                                        (const EdgeCov *)(d + o_ecov), (PieceDev *)(d + o_pc),
                                        (uint32_t *)(d + o_np), dev_overflow, exp_bytes);
            if (st == RVK_OK)
                st = rvk::launch_coverage(ctx, (const CovJob *)(d + o_jobs), (uint32_t)njobs,
                                          (const uint2 *)(d + o_grp[0]), (uint32_t)dl.n_groups[0],
                                          (const uint2 *)(d + o_grp[1]), (uint32_t)dl.n_groups[1],
                                          (const uint2 *)(d + o_grp[2]), (uint32_t)dl.n_groups[2],
                                          (const EdgeCov *)(d + o_ecov), (const uint2 *)(d + o_win),
                                          (const PieceDev *)(d + o_pc), (const uint32_t *)(d + o_np),
                                          (uint16_t *)(d + o_tab), d + o_arena, dev_overflow, cov_bytes);
            ctx->stream = main_stream;
        }
        if (st == RVK_OK && side != ctx->stream &&
            (hipEventRecord(ev_cov, side) != hipSuccess || hipStreamWaitEvent(ctx->stream, ev_cov, 0) != hipSuccess)) {
            (void)hipGetLastError();
            st = RVK_ERR_HIP;
        }
        if (st == RVK_OK)
            st = rvk::launch_draw_items(ctx, dl.px, dl.w, dl.h, dl.bx0, dl.by0, dl.bx1, dl.by1,
                                        (const DrawItem *)(d + o_it), (const DrawCull *)(d + o_cull),
                                        (uint32_t)nitems, (const uint16_t *)(d + o_tab), d + o_arena,
                                        (const ShaderParams *)(d + o_sh), (const uint32_t *)(d + o_ht),
                                        (const uint32_t *)(d + o_hb), (const HairLine *)(d + o_hl),
                                        std::any_of(dl.shaders.begin(), dl.shaders.end(),
                                                    [](const ShaderParams &q) { return q.kind == 3; }),
                                        draw_bytes);
        if (st == RVK_OK) {
            if (hipEventRecord(blk_free[b], ctx->stream) != hipSuccess) {
                (void)hipGetLastError();
                st = RVK_ERR_HIP;
            } else {
                blk_used[b] = true;
            }
        }
        ctx->stats.flushes++;
        dl = DrawList{dl.px, dl.w, dl.h};
        return st;
    }

    void fail(rvk_status s) {
        if (err == RVK_OK) err = s;
    }

    void render_path(const rvk_path &p, Ts ts, DrawList &dl) {
        if (dl.arena > budget) fail(flush(dl));
        DrawItem it{};
        it.layer = 0;
        it.opacity = 1.0f;
        it.blend_mode = RVK_BLEND_SOURCE_OVER;
        path_jobs(dl, p, ts, dl.w, dl.h, IRect{0, 0, (int32_t)dl.w, (int32_t)dl.h}, it);
        push_item(dl, it);
    }

    // steps over the paths of a group's content, clip path and mask when the group
    // is not drawn
    void skip_nodes(const rvk_group &g, Ts ts, DrawList &dl) {
        const Mode saved = mode;
        mode = kSkip;
        render_nodes(g, ts, dl);
        walk_clip_mask(g, ts, (int32_t)dl.w, (int32_t)dl.h);
        mode = saved;
    }

    // collect / skip passes: the clip path and mask paths in drawing order
    void walk_clip_mask(const rvk_group &g, Ts ts, int32_t w, int32_t h) {
        if (g.clip_path >= 0) fail(apply_clip(g.clip_path, ts, nullptr, w, h, 0));
        if (g.mask >= 0) fail(apply_mask(g.mask, ts, nullptr, w, h, 0));
    }

    // ---- clip paths (crates/resvg/src/clip.rs) ----
    static constexpr int kMaxNesting = 64;  // deeper clip / mask / feImage chains are malformed input
    static thread_local inline int fe_image_depth = 0;  // feImage renders in progress on this thread

    // clip::draw_children (clip.rs:30-70) into `dl`: fills of the visible paths (Clear
    // when `clear`), groups with only their transforms; a group with its own clip path
    // is drawn through clip_group
    void clip_children(const rvk_group &parent, bool clear, Ts ts, DrawList &dl, int depth) {
        if (depth > kMaxNesting) {
            fail(RVK_ERR_INVALID_ARGUMENT);
            return;
        }
        for (uint32_t i = 0; i < parent.child_count && err == RVK_OK; i++) {
            const rvk_node_ref &n = sc.children[parent.child_offset + i];
            if (n.kind == RVK_NODE_IMAGE) continue;  // not clip path content
            if (n.kind == RVK_NODE_PATH) {
                const rvk_path &p = sc.paths[n.index];
                if (!p.visible) continue;
                if (dl.arena > budget) fail(flush(dl));
                DrawItem it{};
                it.layer = 0;
                it.opacity = 1.0f;
                it.blend_mode = RVK_BLEND_SOURCE_OVER;
                path_jobs(dl, p, ts, dl.w, dl.h, IRect{0, 0, (int32_t)dl.w, (int32_t)dl.h}, it, true, clear);
                push_item(dl, it);
            } else {
                const rvk_group &g = sc.groups[n.index];
                const Ts gts = ts_concat(ts, ts_from(g.transform));
                if (g.clip_path >= 0) clip_group(g, gts, dl, depth + 1);
                else clip_children(g, clear, gts, dl, depth + 1);
            }
        }
    }

    // clip_group (clip.rs:72-98): the group's fills on their own canvas, clipped, then
    // Xor-ed into the clip pixmap
    void clip_group(const rvk_group &g, Ts ts, DrawList &dl, int depth) {
        DrawList tdl;
        tdl.w = dl.w;
        tdl.h = dl.h;
        if (mode != kExecute) {
            clip_children(g, false, ts, tdl, depth);
            fail(apply_clip(g.clip_path, ts, nullptr, (int32_t)dl.w, (int32_t)dl.h, depth + 1));
            return;
        }
        void *buf = nullptr;
        rvk_status s = rvk::scratch_alloc(ctx, (size_t)dl.w * dl.h * 4, &buf);
        if (s != RVK_OK) {
            fail(s);
            return;
        }
        tdl.px = (uint8_t *)buf;
        s = rvk::launch_fill_sub(ctx, tdl.px, dl.w, dl.h, 0, 0, dl.w, dl.h, 0u);
        if (s == RVK_OK) {
            clip_children(g, false, ts, tdl, depth);
            s = err != RVK_OK ? err : flush(tdl);
        }
        if (s == RVK_OK) s = apply_clip(g.clip_path, ts, tdl.px, (int32_t)dl.w, (int32_t)dl.h, depth + 1);
        if (s == RVK_OK) s = flush(dl);  // earlier fills of the clip pixmap come first
        if (s == RVK_OK)
            s = rvk::launch_draw_pixmap(ctx, dl.px, dl.w, dl.h, 0, 0, tdl.px, dl.w, dl.h, 1.0f, RVK_BLEND_XOR);
        rvk::scratch_release(ctx, buf);
        fail(s);
    }

    // clip::apply (clip.rs:6-28) on a w x h layer: a black pixmap with the clip's
    // fills cleared out of it, the clip's own clip path applied to the layer, then the
    // layer scaled by the inverted alpha
    rvk_status apply_clip(int32_t idx, Ts ts, uint8_t *layer, int32_t w, int32_t h, int depth) {
        if (idx < 0 || (uint32_t)idx >= sc.clip_path_count || depth > kMaxNesting) return RVK_ERR_INVALID_ARGUMENT;
        const rvk_clip_path &cp = sc.clip_paths[idx];
        if (cp.root >= sc.group_count) return RVK_ERR_INVALID_ARGUMENT;
        const Ts cts = ts_concat(ts, ts_from(cp.transform));
        DrawList cdl;
        cdl.w = (uint32_t)w;
        cdl.h = (uint32_t)h;
        if (mode != kExecute) {
            clip_children(sc.groups[cp.root], true, cts, cdl, depth);
            return cp.clip_path >= 0 ? apply_clip(cp.clip_path, ts, nullptr, w, h, depth + 1) : RVK_OK;
        }
        void *buf = nullptr;
        RVK_TRY(rvk::scratch_alloc(ctx, (size_t)w * h * 4, &buf));
        cdl.px = (uint8_t *)buf;
        rvk_status s = rvk::launch_fill_sub(ctx, cdl.px, w, h, 0, 0, w, h, 0xff000000u);  // Color::BLACK
        if (s == RVK_OK) {
            clip_children(sc.groups[cp.root], true, cts, cdl, depth);
            s = err != RVK_OK ? err : flush(cdl);
        }
        if (s == RVK_OK && cp.clip_path >= 0) s = apply_clip(cp.clip_path, ts, layer, w, h, depth + 1);
        if (s == RVK_OK) s = rvk::launch_apply_mask(ctx, layer, cdl.px, (size_t)w * h, 1u);
        rvk::scratch_release(ctx, buf);
        return s;
    }

    // ---- masks (crates/resvg/src/mask.rs) ----
    // mask::apply (mask.rs:6-46): the mask content rendered into its own pixmap and
    // limited to the mask region (Mask::fill_path of the region rect), the mask's own
    // mask applied to the layer, then the layer scaled by the content's luminance or
    // alpha
    rvk_status apply_mask(int32_t idx, Ts ts, uint8_t *layer, int32_t w, int32_t h, int depth) {
        if (idx < 0 || (uint32_t)idx >= sc.mask_count || depth > kMaxNesting) return RVK_ERR_INVALID_ARGUMENT;
        const rvk_mask &mk = sc.masks[idx];
        if (mk.root >= sc.group_count) return RVK_ERR_INVALID_ARGUMENT;
        const rvk_group &root = sc.groups[mk.root];
        if (root.child_count == 0)  // no children: the whole element is masked out
            return mode == kExecute ? rvk::launch_fill_sub(ctx, layer, w, h, 0, 0, w, h, 0u) : RVK_OK;
        DrawList rdl, mdl;
        rdl.w = mdl.w = (uint32_t)w;
        rdl.h = mdl.h = (uint32_t)h;
        auto region = [&] {
            DrawItem it{};
            it.layer = 0;
            it.opacity = 1.0f;
            it.blend_mode = RVK_BLEND_SOURCE_OVER;
            path_jobs(rdl, aux.paths[idx], ts, rdl.w, rdl.h, IRect{0, 0, w, h}, it, true, false, &aux.scene);
            push_item(rdl, it);
        };
        if (mode != kExecute) {
            region();
            render_nodes(root, ts, mdl);
            return mk.mask >= 0 ? apply_mask(mk.mask, ts, nullptr, w, h, depth + 1) : RVK_OK;
        }
        const size_t bytes = (size_t)w * h * 4;
        void *rbuf = nullptr, *mbuf = nullptr;
        RVK_TRY(rvk::scratch_alloc(ctx, bytes, &rbuf));
        rvk_status s = rvk::scratch_alloc(ctx, bytes, &mbuf);
        if (s != RVK_OK) {
            rvk::scratch_release(ctx, rbuf);
            return s;
        }
        rdl.px = (uint8_t *)rbuf;
        mdl.px = (uint8_t *)mbuf;
        s = rvk::launch_fill_sub(ctx, rdl.px, w, h, 0, 0, w, h, 0u);
        if (s == RVK_OK) s = rvk::launch_fill_sub(ctx, mdl.px, w, h, 0, 0, w, h, 0u);
        // the region's coverage: an opaque white fill (Source) leaves alpha = coverage
        if (s == RVK_OK) {
            region();
            s = err != RVK_OK ? err : flush(rdl);
        }
        if (s == RVK_OK) {
            render_nodes(root, ts, mdl);
            s = err != RVK_OK ? err : flush(mdl);
        }
        if (s == RVK_OK) s = rvk::launch_apply_mask(ctx, mdl.px, rdl.px, (size_t)w * h, 0u);
        if (s == RVK_OK && mk.mask >= 0) s = apply_mask(mk.mask, ts, layer, w, h, depth + 1);
        if (s == RVK_OK)
            s = rvk::launch_apply_mask(ctx, layer, mdl.px, (size_t)w * h, mk.kind == RVK_MASK_ALPHA ? 0u : 2u);
        rvk::scratch_release(ctx, mbuf);
        rvk::scratch_release(ctx, rbuf);
        return s;
    }

    // the region rect of every mask as a path (PathBuilder::from_rect) filled opaque white
    struct Aux {
        std::vector<uint8_t> verbs;
        std::vector<float> points;
        std::vector<rvk_path> paths;
        rvk_scene scene{};
    } aux;

    void build_aux() {
        for (uint32_t i = 0; i < sc.mask_count; i++) {
            const rvk_rect r = sc.masks[i].rect;
            const float l = r.x, t = r.y, rr = r.x + r.w, b = r.y + r.h;
            rvk_path p{};
            p.verb_offset = (uint32_t)aux.verbs.size();
            p.verb_count = 5;
            p.point_offset = (uint32_t)(aux.points.size() / 2);
            p.visible = 1;
            p.anti_alias = 1;
            p.fill.kind = RVK_PAINT_COLOR;
            p.fill.color[0] = p.fill.color[1] = p.fill.color[2] = 255;
            p.fill.opacity = 1.0f;
            p.fill_rule = RVK_FILL_NONZERO;
            p.stroke.kind = RVK_PAINT_NONE;
            p.bounds = r;
            for (uint8_t v : {RVK_VERB_MOVE, RVK_VERB_LINE, RVK_VERB_LINE, RVK_VERB_LINE, RVK_VERB_CLOSE})
                aux.verbs.push_back(v);
            for (float v : {l, t, rr, t, rr, b, l, b}) aux.points.push_back(v);
            aux.paths.push_back(p);
        }
        // every raster image's rect (0, 0, w, h) as a path filled with the image
        // (render_raster's fill_rect), after the mask regions
        for (uint32_t i = 0; i < sc.image_count; i++) {
            const rvk_image &im = sc.images[i];
            const float w = (float)im.width, h = (float)im.height;
            rvk_path p{};
            p.verb_offset = (uint32_t)aux.verbs.size();
            p.verb_count = 5;
            p.point_offset = (uint32_t)(aux.points.size() / 2);
            p.visible = 1;
            p.anti_alias = 1;
            p.fill.kind = kPaintImage;
            p.fill.root = i;
            p.fill.opacity = 1.0f;
            p.fill_rule = RVK_FILL_NONZERO;
            p.stroke.kind = RVK_PAINT_NONE;
            p.bounds = rvk_rect{0.0f, 0.0f, w, h};
            for (uint8_t v : {RVK_VERB_MOVE, RVK_VERB_LINE, RVK_VERB_LINE, RVK_VERB_LINE, RVK_VERB_CLOSE})
                aux.verbs.push_back(v);
            for (float v : {0.0f, 0.0f, w, 0.0f, w, h, 0.0f, h}) aux.points.push_back(v);
            aux.paths.push_back(p);
        }
        image_tiles.assign(sc.image_count, nullptr);
        aux.scene.verbs = aux.verbs.data();
        aux.scene.points = aux.points.data();
        aux.scene.paths = aux.paths.data();
        aux.scene.path_count = (uint32_t)aux.paths.size();
    }

    // image::render (image.rs:4-35): a visible raster image is its rect filled with the
    // image's pixmap
    void render_image(uint32_t index, Ts ts, DrawList &dl) {
        const rvk_image &im = sc.images[index];
        if (!im.visible || im.width == 0 || im.height == 0 || !im.pixels) return;
        if (dl.arena > budget) fail(flush(dl));
        DrawItem it{};
        it.layer = 0;
        it.opacity = 1.0f;
        it.blend_mode = RVK_BLEND_SOURCE_OVER;
        path_jobs(dl, aux.paths[sc.mask_count + index], ts, dl.w, dl.h, IRect{0, 0, (int32_t)dl.w, (int32_t)dl.h}, it,
                  true, false, &aux.scene);
        push_item(dl, it);
    }

    void render_nodes(const rvk_group &g, Ts ts, DrawList &dl) {
        for (uint32_t i = 0; i < g.child_count && err == RVK_OK; i++) {
            const rvk_node_ref &n = sc.children[g.child_offset + i];
            if (n.kind == RVK_NODE_GROUP) render_group(sc.groups[n.index], ts, dl);
            else if (n.kind == RVK_NODE_IMAGE) render_image(n.index, ts, dl);
            else render_path(sc.paths[n.index], ts, dl);
        }
    }

    // render_group (render.rs:49-143)
    void render_group(const rvk_group &g, Ts ts, DrawList &dl) {
        ts = ts_concat(ts, ts_from(g.transform));
        const bool isolate = g.isolate || g.opacity != 1.0f || g.clip_path >= 0 || g.mask >= 0 ||
                             g.filter_count > 0 || g.blend_mode != RVK_MIX_NORMAL;
        if (!isolate) {
            render_nodes(g, ts, dl);
            return;
        }
        FRect bb = rect_transform(g.layer_bbox, ts);
        if (!bb.ok) return;
        const float bw = bb.r - bb.l, bh = bb.b - bb.t;
        IRect ib;
        if (g.filter_count == 0) {
            int64_t x = (int64_t)as_i32(std::floor(bb.l)) - 2, y = (int64_t)as_i32(std::floor(bb.t)) - 2;
            if (x < INT32_MIN || y < INT32_MIN) return;
            if (!int_rect(x, y, (int64_t)as_u32(std::ceil(bw)) + 4, (int64_t)as_u32(std::ceil(bh)) + 4, ib)) return;
            if (!fit_to_rect(ib, max_bbox, ib)) return;
        } else {
            IRect t;
            if (!int_rect(as_i32(std::floor(bb.l)), as_i32(std::floor(bb.t)), as_u32(std::max(std::ceil(bw), 1.0f)),
                          as_u32(std::max(std::ceil(bh), 1.0f)), t))
                return;
            if (!fit_to_rect(t, max_bbox, ib)) return;
        }
        float dx = bb.l, dy = bb.t;
        dx -= bb.l - (float)ib.x;
        dy -= bb.t - (float)ib.y;
        const Ts lts = ts_concat(Ts{1, 0, 0, 1, -dx, -dy}, ts);
        if (mode == kExecute) ctx->stats.layers++;
        // elided layer: one path child, no filters, clip path or mask
        if (g.filter_count == 0 && g.clip_path < 0 && g.mask < 0 && g.child_count == 1 &&
            sc.children[g.child_offset].kind == RVK_NODE_PATH) {
            if (dl.arena > budget) fail(flush(dl));
            const rvk_path &p = sc.paths[sc.children[g.child_offset].index];
            DrawItem it{};
            it.layer = 1;
            it.lx = ib.x;
            it.ly = ib.y;
            it.opacity = g.opacity;
            it.blend_mode = convert_blend_mode(g.blend_mode);
            // layer pixels that land on the target
            int32_t cx0 = std::max(0, -ib.x), cy0 = std::max(0, -ib.y);
            int32_t cx1 = std::min(ib.w, (int32_t)dl.w - ib.x), cy1 = std::min(ib.h, (int32_t)dl.h - ib.y);
            if (cx1 <= cx0 || cy1 <= cy0) return;
            path_jobs(dl, p, lts, (uint32_t)ib.w, (uint32_t)ib.h, IRect{cx0, cy0, cx1 - cx0, cy1 - cy0}, it);
            push_item(dl, it);
            return;
        }
        // materialised layer
        DrawList child;
        child.w = (uint32_t)ib.w;
        child.h = (uint32_t)ib.h;
        if (mode != kExecute) {
            render_nodes(g, lts, child);
            walk_clip_mask(g, lts, ib.w, ib.h);
            return;
        }
        void *buf = nullptr;
        rvk_status s = rvk::scratch_alloc(ctx, (size_t)ib.w * ib.h * 4, &buf);
        if (s != RVK_OK) {  // Pixmap::new failure skips the group (render.rs:108-109)
            skip_nodes(g, lts, child);
            return;
        }
        uint8_t *layer = (uint8_t *)buf;
        child.px = layer;
        s = rvk::launch_fill_sub(ctx, layer, ib.w, ib.h, 0, 0, ib.w, ib.h, 0u);
        if (s == RVK_OK) {
            render_nodes(g, lts, child);
            s = flush(child);
        } else {
            skip_nodes(g, lts, child);
        }
        // feImage content: the children of a scene group drawn by a render of their own
        // into the primitive's pixmap, with max_bbox the pixmap (filter/mod.rs:878-882)
        const rvk::FeImageHook fe_image = [this](uint32_t root, Ts t, uint8_t *img, uint32_t iw, uint32_t ih) {
            if (root >= sc.group_count || fe_image_depth >= kMaxNesting) return (rvk_status)RVK_ERR_INVALID_ARGUMENT;
            const IRect mb{0, 0, (int32_t)iw, (int32_t)ih};
            fe_image_depth++;
            const rvk_status st = rvk::render_walk(ctx, &sc, t, iw, ih, img, nullptr, rvk::Walk::kChildren,
                                                   rvk_node_ref{RVK_NODE_GROUP, root}, &mb);
            fe_image_depth--;
            return st;
        };
        for (uint32_t i = 0; s == RVK_OK && err == RVK_OK && i < g.filter_count; i++)
            s = rvk::filter_apply(ctx, &sc.filters[g.filter_offset + i], &sc.filter_data, lts, layer, ib.w, ib.h,
                                  nullptr, &fe_image);
        if (s == RVK_OK && err == RVK_OK && g.clip_path >= 0) s = apply_clip(g.clip_path, lts, layer, ib.w, ib.h, 0);
        if (s == RVK_OK && err == RVK_OK && g.mask >= 0) s = apply_mask(g.mask, lts, layer, ib.w, ib.h, 0);
        if (s == RVK_OK) s = flush(dl);
        if (s == RVK_OK)
            s = rvk::launch_draw_pixmap(ctx, dl.px, dl.w, dl.h, ib.x, ib.y, layer, ib.w, ib.h, g.opacity,
                                        convert_blend_mode(g.blend_mode));
        rvk::scratch_release(ctx, buf);
        fail(s);
    }
};

}  // namespace

namespace rvk {

rvk_status render_walk(rvk_context *ctx, const rvk_scene *scene, Ts ts, uint32_t w, uint32_t h, uint8_t *px,
                       uint64_t *overflow_rows, Walk walk, rvk_node_ref node, const IRect *max_bbox) {
    Renderer r{ctx, *scene, max_bbox ? *max_bbox : max_filter_bbox(w, h)};
    r.build_aux();
    void *ovbuf = nullptr;
    RVK_TRY(rvk::scratch_alloc(ctx, 256, &ovbuf));
    r.dev_overflow = (int32_t *)ovbuf;
    {
        const rvk_status s = rvk::launch_fill_sub(ctx, (uint8_t *)ovbuf, 1, 1, 0, 0, 1, 1, 0u);  // counter = 0
        if (s != RVK_OK) {
            rvk::scratch_release(ctx, ovbuf);
            return s;
        }
    }
    auto run = [&](DrawList &dl) {
        switch (walk) {
            case Walk::kRoot: r.render_nodes(scene->groups[0], ts, dl); break;
            case Walk::kNode:  // render::render_node (render.rs:21-47)
                if (node.kind == RVK_NODE_GROUP) r.render_group(scene->groups[node.index], ts, dl);
                else if (node.kind == RVK_NODE_IMAGE) r.render_image(node.index, ts, dl);
                else r.render_path(scene->paths[node.index], ts, dl);
                break;
            case Walk::kClearFills: r.clip_children(scene->groups[0], true, ts, dl, 0); break;
            case Walk::kChildren: r.render_nodes(scene->groups[node.index], ts, dl); break;
        }
    };
    DrawList root;
    root.w = w;
    root.h = h;
    r.mode = Renderer::kCollect;
    run(root);
    r.geoms.resize(r.tasks.size());
    r.mode = Renderer::kExecute;
    root = DrawList{};
    root.px = px;
    root.w = w;
    root.h = h;
    run(root);
    if (r.err == RVK_OK) r.fail(r.flush(root));
    r.fail(r.finish_flushes());  // the side stream joins the context stream
    // the one wait of a render: read the overflow counter
    int32_t ov = 0;
    {
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = hipMemcpyAsync(&ov, ovbuf, 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            r.fail(RVK_ERR_HIP);
        }
        ctx->stats.host_wait_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    rvk::scratch_release(ctx, ovbuf);
    for (void *t : r.pattern_tiles) rvk::scratch_release(ctx, t);
    r.overflow_rows = (uint64_t)(ov > 0 ? ov : 0);
    if (overflow_rows) *overflow_rows = r.overflow_rows;
    if (r.err == RVK_OK && r.overflow_rows) return RVK_ERR_UNSUPPORTED;
    return r.err;
}

rvk_status render_scene(rvk_context *ctx, const rvk_scene *scene, Ts ts, uint32_t w, uint32_t h, uint8_t *px,
                        uint64_t *overflow_rows) {
    return render_walk(ctx, scene, ts, w, h, px, overflow_rows, Walk::kRoot, rvk_node_ref{0, 0});
}

// Streams a batch spreads its documents over (RVK_BATCH_LANES overrides, 1-32).
uint32_t batch_lanes() {
    if (const char *e = std::getenv("RVK_BATCH_LANES")) {
        const int v = std::atoi(e);
        if (v > 0) return (uint32_t)std::min(v, 32);
    }
    return 16;
}

// A batch of documents.  Document d is drawn on lane d % lanes (sub-contexts with
// their own streams, so the kernels of several documents run at once) by a host
// thread of that lane, which runs the documents' execute walks in order.  The
// calling thread meanwhile prepares chunks of documents: every collect walk of the
// chunk, then the geometry of all their paths in one parallel pass, handing each
// document to its lane when done.  The lanes start after the work already queued on
// the context stream and the context stream continues after all of them; each
// document has its own overflow counter and the batch ends with the one wait.
// Documents whose pixmaps overlap are drawn in order on a single lane.
rvk_status render_batch(rvk_context *ctx, uint32_t count, const rvk_scene *const *scenes, const rvk_transform *tss,
                        const uint32_t *ws, const uint32_t *hs, uint8_t *const *pxs, rvk_status *statuses) {
    uint32_t nl = std::min(batch_lanes(), count);
    {
        std::vector<std::pair<uintptr_t, uintptr_t>> spans(count);
        for (uint32_t d = 0; d < count; d++)
            spans[d] = {(uintptr_t)pxs[d], (uintptr_t)pxs[d] + (uintptr_t)ws[d] * hs[d] * 4};
        std::sort(spans.begin(), spans.end());
        for (uint32_t d = 1; d < count; d++)
            if (spans[d].first < spans[d - 1].second) nl = 1;
    }
    RVK_TRY(rvk::lanes_prepare(ctx, nl));
    void *ovbuf = nullptr;
    RVK_TRY(rvk::scratch_alloc(ctx, (size_t)count * 4, &ovbuf));
    int32_t *dev_ov = (int32_t *)ovbuf;
    std::vector<hipEvent_t> evs(nl + 1, nullptr);
    rvk_status s = rvk::launch_fill_sub(ctx, (uint8_t *)ovbuf, count, 1, 0, 0, count, 1, 0u);  // counters = 0
    for (auto &e : evs)
        if (s == RVK_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) s = RVK_ERR_HIP;
    if (s == RVK_OK && hipEventRecord(evs[nl], ctx->stream) != hipSuccess) s = RVK_ERR_HIP;
    for (uint32_t l = 0; s == RVK_OK && l < nl; l++)
        if (hipStreamWaitEvent(ctx->lanes[l]->stream, evs[nl], 0) != hipSuccess) s = RVK_ERR_HIP;
    if (s != RVK_OK) (void)hipGetLastError();

    std::vector<std::unique_ptr<Renderer>> rs(count);
    std::vector<rvk_status> st(count, RVK_OK);
    std::atomic<int32_t> fatal{s};
    // the hand-over: per lane a queue of prepared documents; `pending` counts the
    // paths of prepared documents not yet drawn (the calling thread waits while it
    // exceeds two chunks, which bounds the host memory held by prepared geometry)
    struct Lane {
        std::mutex mu;
        std::condition_variable cv;
        std::vector<uint32_t> q;
        size_t head = 0;
        bool closed = false;
    };
    std::vector<Lane> lanes(nl);
    std::mutex pmu;
    std::condition_variable pcv;
    size_t pending = 0;
    constexpr uint32_t kChunkDocs = 64;
    constexpr size_t kChunkTasks = 2048;
    auto draw_doc = [&](uint32_t d) {
        Renderer &r = *rs[d];
        const size_t nt = r.tasks.size();
        if (fatal.load() == RVK_OK) {
            r.ready_end = nt;
            r.mode = Renderer::kExecute;
            r.parallel_pack = false;  // the pool belongs to the preparing thread
            DrawList root;
            root.px = pxs[d];
            root.w = ws[d];
            root.h = hs[d];
            r.render_nodes(scenes[d]->groups[0], ts_from(tss[d]), root);
            if (r.err == RVK_OK) r.fail(r.flush(root));
            r.fail(r.finish_flushes());
            st[d] = r.err;
            if (r.err == RVK_ERR_HIP || r.err == RVK_ERR_OUT_OF_MEMORY) {  // the stream is unusable
                int32_t expect = RVK_OK;
                fatal.compare_exchange_strong(expect, r.err);
            }
        }
        rs[d].reset();
        {
            std::lock_guard<std::mutex> g(pmu);
            pending -= nt;
        }
        pcv.notify_all();
    };
    std::vector<std::thread> threads;
    for (uint32_t l = 0; l < nl; l++)
        threads.emplace_back([&, l] {
            (void)hipSetDevice(ctx->device);
            Lane &ln = lanes[l];
            for (;;) {
                uint32_t d;
                {
                    std::unique_lock<std::mutex> g(ln.mu);
                    ln.cv.wait(g, [&] { return ln.closed || ln.head < ln.q.size(); });
                    if (ln.head == ln.q.size()) return;  // closed and drained
                    d = ln.q[ln.head++];
                }
                draw_doc(d);
            }
        });
    for (uint32_t d0 = 0; d0 < count && fatal.load() == RVK_OK;) {
        {
            const auto t0 = std::chrono::steady_clock::now();
            std::unique_lock<std::mutex> g(pmu);
            pcv.wait(g, [&] { return pending <= 2 * kChunkTasks || fatal.load() != RVK_OK; });
            ctx->stats.host_wait_ms +=
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        size_t ntasks = 0;
        uint32_t d1 = d0;
        for (; d1 < count && d1 - d0 < kChunkDocs && ntasks < kChunkTasks; d1++) {
            rs[d1].reset(new Renderer{ctx->lanes[d1 % nl], *scenes[d1], max_filter_bbox(ws[d1], hs[d1])});
            Renderer &r = *rs[d1];
            r.build_aux();
            r.dev_overflow = dev_ov + d1;
            DrawList root;
            root.w = ws[d1];
            root.h = hs[d1];
            r.mode = Renderer::kCollect;
            r.render_nodes(scenes[d1]->groups[0], ts_from(tss[d1]), root);
            r.geoms.resize(r.tasks.size());
            ntasks += r.tasks.size();
        }
        std::vector<std::pair<Renderer *, uint32_t>> all;
        all.reserve(ntasks);
        for (uint32_t d = d0; d < d1; d++)
            for (size_t k = 0; k < rs[d]->tasks.size(); k++) all.emplace_back(rs[d].get(), (uint32_t)k);
        // longest first (the pool hands items out in order), so that the pass does not
        // end on one long stroke: estimated cost = verbs, times 8 with a stroke outline
        auto cost = [](const std::pair<Renderer *, uint32_t> &a) {
            const auto &t = a.first->tasks[a.second];
            const bool stroked = !t.fill_only && t.path->stroke.kind != RVK_PAINT_NONE;
            return (uint64_t)t.path->verb_count * (stroked ? 8u : 1u);
        };
        std::stable_sort(all.begin(), all.end(), [&](const auto &a, const auto &b) { return cost(a) > cost(b); });
        const auto t0 = std::chrono::steady_clock::now();
        parallel_for(all.size(), [&](size_t i) {
            Renderer &r = *all[i].first;
            const auto &t = r.tasks[all[i].second];
            compute_geom(*t.scene, *t.path, t.ts, t.pw, t.ph, t.fill_only, t.clip, t.lx, t.ly,
                         r.geoms[all[i].second]);
        });
        ctx->stats.host_geom_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        {
            std::lock_guard<std::mutex> g(pmu);
            pending += ntasks;
        }
        for (uint32_t d = d0; d < d1; d++) {
            Lane &ln = lanes[d % nl];
            {
                std::lock_guard<std::mutex> g(ln.mu);
                ln.q.push_back(d);
            }
            ln.cv.notify_one();
        }
        d0 = d1;
    }
    for (Lane &ln : lanes) {
        {
            std::lock_guard<std::mutex> g(ln.mu);
            ln.closed = true;
        }
        ln.cv.notify_one();
    }
    {
        const auto t0 = std::chrono::steady_clock::now();
        for (auto &t : threads) t.join();
        ctx->stats.host_wait_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    s = fatal.load();
    // join: the context stream waits for every lane
    for (uint32_t l = 0; l < nl; l++) {
        if (!evs[l] || hipEventRecord(evs[l], ctx->lanes[l]->stream) != hipSuccess ||
            hipStreamWaitEvent(ctx->stream, evs[l], 0) != hipSuccess) {
            (void)hipGetLastError();
            if (s == RVK_OK) s = RVK_ERR_HIP;
            (void)hipStreamSynchronize(ctx->lanes[l]->stream);
        }
    }
    std::vector<int32_t> ov(count, 0);
    {
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = hipMemcpyAsync(ov.data(), ovbuf, (size_t)count * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (s == RVK_OK) s = RVK_ERR_HIP;
        }
        ctx->stats.host_wait_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    for (auto e : evs)
        if (e) (void)hipEventDestroy(e);
    rvk::scratch_release(ctx, ovbuf);
    rvk::lanes_collect(ctx, nl);
    rvk_status first = s;
    for (uint32_t d = 0; d < count; d++) {
        if (s != RVK_OK && st[d] == RVK_OK) st[d] = s;
        if (st[d] == RVK_OK && ov[d] > 0) st[d] = RVK_ERR_UNSUPPORTED;
        if (first == RVK_OK) first = st[d];
        if (statuses) statuses[d] = st[d];
    }
    return first;
}

}  // namespace rvk

extern "C" {

static bool scene_ok(const rvk_scene *s) {
    return s && s->group_count > 0 && s->groups && (s->path_count == 0 || (s->paths && s->verbs && s->points)) &&
           (s->clip_path_count == 0 || s->clip_paths) && (s->mask_count == 0 || s->masks) &&
           (s->image_count == 0 || s->images);
}

rvk_status rvk_render(rvk_context *ctx, const rvk_scene *scene, rvk_transform transform, uint32_t width,
                      uint32_t height, uint8_t *dev_pixmap) {
    if (!ctx || !scene_ok(scene) || !dev_pixmap || !rvk::valid_dims(width, height)) return RVK_ERR_INVALID_ARGUMENT;
    ctx->stats = rvk_stats{};
    const auto t0 = std::chrono::steady_clock::now();
    const rvk_status s = rvk::render_scene(ctx, scene, ts_from(transform), width, height, dev_pixmap, nullptr);
    ctx->stats.host_build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() -
        ctx->stats.host_wait_ms;
    return s;
}

rvk_status rvk_render_batch(rvk_context *ctx, uint32_t count, const rvk_scene *const *scenes,
                            const rvk_transform *transforms, const uint32_t *widths, const uint32_t *heights,
                            uint8_t *const *dev_pixmaps, rvk_status *statuses) {
    if (!ctx || (count > 0 && (!scenes || !transforms || !widths || !heights || !dev_pixmaps)))
        return RVK_ERR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < count; i++)
        if (!scene_ok(scenes[i]) || !dev_pixmaps[i] || !rvk::valid_dims(widths[i], heights[i]))
            return RVK_ERR_INVALID_ARGUMENT;
    ctx->stats = rvk_stats{};
    if (count == 0) return RVK_OK;
    const auto t0 = std::chrono::steady_clock::now();
    const rvk_status s =
        rvk::render_batch(ctx, count, scenes, transforms, widths, heights, dev_pixmaps, statuses);
    ctx->stats.host_build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() -
        ctx->stats.host_wait_ms;
    return s;
}

// resvg::render_node (lib.rs:55-70): nothing without a layer bounding box, else the
// node drawn with the transform pre-translated by minus the box's origin.
rvk_status rvk_render_node(rvk_context *ctx, const rvk_scene *scene, rvk_node_ref node, const rvk_rect *bbox,
                           rvk_transform transform, uint32_t width, uint32_t height, uint8_t *dev_pixmap,
                           int32_t *rendered) {
    if (!ctx || !scene_ok(scene) || !dev_pixmap || !rendered || !rvk::valid_dims(width, height))
        return RVK_ERR_INVALID_ARGUMENT;
    if ((node.kind == RVK_NODE_GROUP && node.index >= scene->group_count) ||
        (node.kind == RVK_NODE_PATH && node.index >= scene->path_count) ||
        (node.kind == RVK_NODE_IMAGE && node.index >= scene->image_count) ||
        (node.kind != RVK_NODE_GROUP && node.kind != RVK_NODE_PATH && node.kind != RVK_NODE_IMAGE))
        return RVK_ERR_INVALID_ARGUMENT;
    *rendered = 0;
    ctx->stats = rvk_stats{};
    // Node::abs_layer_bounding_box() is a NonZeroRect: None unless positive and finite
    if (!bbox || !(bbox->w > 0.0f) || !(bbox->h > 0.0f) || !std::isfinite(bbox->x) || !std::isfinite(bbox->y) ||
        !std::isfinite(bbox->w) || !std::isfinite(bbox->h))
        return RVK_OK;
    const Ts ts = ts_concat(ts_from(transform), Ts{1.0f, 0.0f, 0.0f, 1.0f, -bbox->x, -bbox->y});  // pre_translate
    const auto t0 = std::chrono::steady_clock::now();
    const rvk_status s = rvk::render_walk(ctx, scene, ts, width, height, dev_pixmap, nullptr, rvk::Walk::kNode, node);
    ctx->stats.host_build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() -
        ctx->stats.host_wait_ms;
    if (s == RVK_OK) *rendered = 1;
    return s;
}

// One path under a root group, built from the caller's arrays (they stay in place):
// what path.rs's fill_path / stroke_path draw when called on their own.
struct OnePath {
    rvk_group root{};
    rvk_path path{};
    rvk_node_ref child{RVK_NODE_PATH, 0};
    rvk_scene sc{};
};

// Checks the verbs against the points they consume; fills `o` (no paint set).
static rvk_status one_path(const rvk_path_geometry *g, uint32_t anti_alias, OnePath &o) {
    if (!g || (g->verb_count && !g->verbs) || (g->point_count && !g->points)) return RVK_ERR_INVALID_ARGUMENT;
    uint64_t need = 0;
    for (uint32_t i = 0; i < g->verb_count; i++) {
        switch (g->verbs[i]) {
            case RVK_VERB_MOVE: case RVK_VERB_LINE: need += 1; break;
            case RVK_VERB_QUAD: need += 2; break;
            case RVK_VERB_CUBIC: need += 3; break;
            case RVK_VERB_CLOSE: break;
            default: return RVK_ERR_INVALID_ARGUMENT;
        }
    }
    if (need > g->point_count) return RVK_ERR_INVALID_ARGUMENT;
    o.root.transform = rvk_transform{1, 0, 0, 1, 0, 0};
    o.root.opacity = 1.0f;
    o.root.blend_mode = RVK_MIX_NORMAL;
    o.root.clip_path = o.root.mask = -1;
    o.root.layer_bbox = rvk_rect{0, 0, 1, 1};
    o.root.child_offset = 0;
    o.root.child_count = 1;
    rvk_path &p = o.path;
    p.verb_offset = 0;
    p.verb_count = g->verb_count;
    p.point_offset = 0;
    p.visible = 1;
    p.anti_alias = anti_alias ? 1 : 0;
    p.fill.kind = p.stroke.kind = RVK_PAINT_NONE;
    p.fill_rule = RVK_FILL_NONZERO;
    p.stroke_miter_limit = 4.0f;
    // tiny_skia_path::Path::bounds(): the bounds of every point
    float l = INFINITY, t = INFINITY, r = -INFINITY, b = -INFINITY;
    for (uint64_t i = 0; i < need; i++) {
        l = std::min(l, g->points[2 * i]); r = std::max(r, g->points[2 * i]);
        t = std::min(t, g->points[2 * i + 1]); b = std::max(b, g->points[2 * i + 1]);
    }
    p.bounds = need ? rvk_rect{l, t, r - l, b - t} : rvk_rect{0, 0, 0, 0};
    o.sc.groups = &o.root;
    o.sc.group_count = 1;
    o.sc.paths = &o.path;
    o.sc.path_count = 1;
    o.sc.children = &o.child;
    o.sc.verbs = g->verbs;
    o.sc.points = g->points;
    return need ? RVK_OK : RVK_ERR_UNSUPPORTED;  // no points: nothing to draw (mapped to OK by callers)
}

static bool paint_ok(const rvk_paint *pt, const rvk_stop *stops) {
    if (!pt || pt->kind > RVK_PAINT_RADIAL || pt->spread > RVK_SPREAD_REPEAT) return false;
    return pt->kind < RVK_PAINT_LINEAR || pt->stop_count == 0 || stops;
}

rvk_status rvk_fill_path(rvk_context *ctx, uint8_t *dev_px, uint32_t width, uint32_t height,
                         const rvk_path_geometry *path, uint32_t fill_rule, const rvk_paint *paint,
                         const rvk_stop *stops, rvk_transform transform, uint32_t blend_mode, uint32_t anti_alias) {
    if (!ctx || !dev_px || !rvk::valid_dims(width, height) || !paint_ok(paint, stops) || fill_rule > RVK_FILL_EVENODD)
        return RVK_ERR_INVALID_ARGUMENT;
    if (blend_mode != RVK_BLEND_SOURCE_OVER && blend_mode != RVK_BLEND_CLEAR) return RVK_ERR_UNSUPPORTED;
    OnePath o;
    const rvk_status v = one_path(path, anti_alias, o);
    if (v != RVK_OK) return v == RVK_ERR_UNSUPPORTED ? RVK_OK : v;
    o.path.fill = *paint;
    o.path.fill_rule = fill_rule;
    o.sc.stops = stops;
    ctx->stats = rvk_stats{};
    return rvk::render_walk(ctx, &o.sc, ts_from(transform), width, height, dev_px, nullptr,
                            blend_mode == RVK_BLEND_CLEAR ? rvk::Walk::kClearFills : rvk::Walk::kRoot,
                            rvk_node_ref{0, 0});
}

rvk_status rvk_stroke_path(rvk_context *ctx, uint8_t *dev_px, uint32_t width, uint32_t height,
                           const rvk_path_geometry *path, const rvk_stroke_style *stroke, const rvk_paint *paint,
                           const rvk_stop *stops, rvk_transform transform, uint32_t anti_alias) {
    if (!ctx || !dev_px || !rvk::valid_dims(width, height) || !paint_ok(paint, stops) || !stroke ||
        stroke->cap > RVK_CAP_SQUARE || stroke->join > RVK_JOIN_BEVEL || (stroke->dash_count && !stroke->dashes))
        return RVK_ERR_INVALID_ARGUMENT;
    OnePath o;
    const rvk_status v = one_path(path, anti_alias, o);
    if (v != RVK_OK) return v == RVK_ERR_UNSUPPORTED ? RVK_OK : v;
    o.path.stroke = *paint;
    o.path.stroke_width = stroke->width;
    o.path.stroke_miter_limit = stroke->miter_limit;
    o.path.stroke_cap = stroke->cap;
    o.path.stroke_join = stroke->join;
    o.path.dash_offset = 0;
    o.path.dash_count = stroke->dash_count;
    o.path.dash_phase = stroke->dash_offset;
    o.sc.stops = stops;
    o.sc.dashes = stroke->dashes;
    ctx->stats = rvk_stats{};
    return rvk::render_walk(ctx, &o.sc, ts_from(transform), width, height, dev_px, nullptr, rvk::Walk::kRoot,
                            rvk_node_ref{0, 0});
}

// Mask::fill_path on a new mask: the coverage is the alpha of an opaque white fill
// (an opaque shader composites with Source, so alpha = coverage) on a cleared pixmap.
rvk_status rvk_mask_fill_path(rvk_context *ctx, uint8_t *dev_mask, uint32_t width, uint32_t height,
                              const rvk_path_geometry *path, uint32_t fill_rule, uint32_t anti_alias,
                              rvk_transform transform) {
    if (!ctx || !dev_mask || !rvk::valid_dims(width, height) || fill_rule > RVK_FILL_EVENODD)
        return RVK_ERR_INVALID_ARGUMENT;
    OnePath o;
    const rvk_status v = one_path(path, anti_alias, o);
    if (v != RVK_OK && v != RVK_ERR_UNSUPPORTED) return v;
    const size_t n = (size_t)width * height;
    if (v == RVK_ERR_UNSUPPORTED) return rvk::launch_fill_u8(ctx, dev_mask, n, 0);
    o.path.fill.kind = RVK_PAINT_COLOR;
    o.path.fill.color[0] = o.path.fill.color[1] = o.path.fill.color[2] = 255;
    o.path.fill.opacity = 1.0f;
    o.path.fill_rule = fill_rule;
    void *buf = nullptr;
    RVK_TRY(rvk::scratch_alloc(ctx, n * 4, &buf));
    ctx->stats = rvk_stats{};
    rvk_status s = rvk::launch_fill_sub(ctx, (uint8_t *)buf, width, height, 0, 0, width, height, 0u);
    if (s == RVK_OK)
        s = rvk::render_walk(ctx, &o.sc, ts_from(transform), width, height, (uint8_t *)buf, nullptr, rvk::Walk::kRoot,
                             rvk_node_ref{0, 0});
    if (s == RVK_OK) s = rvk::launch_mask_from_pixmap(ctx, (const uint8_t *)buf, dev_mask, n, RVK_MASK_TYPE_ALPHA);
    rvk::scratch_release(ctx, buf);
    return s;
}

rvk_status rvk_render_host(rvk_context *ctx, const rvk_scene *scene, rvk_transform transform, uint32_t width,
                           uint32_t height, uint8_t *host_pixmap) {
    if (!ctx || !scene_ok(scene) || !host_pixmap || !rvk::valid_dims(width, height)) return RVK_ERR_INVALID_ARGUMENT;
    ctx->stats = rvk_stats{};
    const size_t bytes = (size_t)width * height * 4;
    void *dev = nullptr;
    RVK_TRY(rvk::scratch_alloc(ctx, bytes, &dev));
    rvk_status s = rvk::upload_sync(ctx, dev, host_pixmap, bytes);
    if (s == RVK_OK) s = rvk::render_scene(ctx, scene, ts_from(transform), width, height, (uint8_t *)dev, nullptr);
    if (s == RVK_OK) {
        hipError_t e = hipMemcpyAsync(host_pixmap, dev, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) s = RVK_ERR_HIP;
    }
    rvk::scratch_release(ctx, dev);
    return s;
}

}  // extern "C"
