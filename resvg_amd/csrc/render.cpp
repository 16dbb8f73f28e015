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
#include <cstdlib>
#include <memory>
#include <thread>

using namespace rvkh;
using namespace rvkr;

namespace rvk {
rvk_status launch_expand(rvk_context *ctx, const EdgeRec *edges, uint32_t nedges, const EdgeCov *cov,
                         PieceDev *pieces, uint32_t *npieces, int32_t *overflow, double algo_bytes);
rvk_status launch_coverage(rvk_context *ctx, const CovJob *jobs, uint32_t njobs, const uint32_t *groups128,
                           uint32_t n128, const uint32_t *groups256, uint32_t n256, const uint32_t *groups512,
                           uint32_t n512, const EdgeCov *edges, const uint2 *windows, const PieceDev *pieces,
                           const uint32_t *npieces, uint8_t *arena, int32_t *overflow, double algo_bytes);
rvk_status launch_draw_items(rvk_context *ctx, uint8_t *target, uint32_t w, uint32_t h, int32_t x0, int32_t y0,
                             int32_t x1, int32_t y1, const DrawItem *items, uint32_t nitems, const CovJob *jobs,
                             const uint8_t *arena, const ShaderParams *shaders, const HairJob *hair_jobs,
                             const uint32_t *hair_tiles, const uint32_t *hair_bins, const HairLine *hair_lines,
                             double algo_bytes);
}

namespace {

// Coverage bytes per flush (1 GiB); RVK_FLUSH_BYTES lowers it, which tests use to
// drive many flushes through small documents.
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
    std::vector<CovJob> jobs;
    std::vector<EdgeRec> edges;       // every job's edges, each job's sorted by first row
    std::vector<EdgeCov> ecov;        // the coverage kernel's view of the same edges
    std::vector<uint2> windows;       // per job and row group: [lo, hi) of the job's edges
    std::vector<uint32_t> active;     // per job and row group: edges whose rows meet the group
    uint64_t piece_slots = 0;         // piece capacity of all edges
    std::vector<ShaderParams> shaders;
    std::vector<HairJob> hair_jobs;   // hairline strokes
    std::vector<uint32_t> hair_tiles; // per job: bin offsets of its tiles (+1 end)
    std::vector<uint32_t> hair_bins;  // line indices, each tile's in drawing order
    std::vector<HairLine> hair_lines;
    uint64_t arena = 0;
    uint32_t rows = 0;
    int32_t bx0 = INT32_MAX, by0 = INT32_MAX, bx1 = INT32_MIN, by1 = INT32_MIN;
};

uint32_t convert_blend_mode(uint32_t m) {  // render.rs:254-273
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

// Edges of one fill or stroke, computed off the walk (the expensive host part).
struct EdgeSet {
    bool ok = false;
    std::vector<EdgeRec> edges;
    std::vector<uint32_t> caps;
    float mnx = 0.0f, mxx = 0.0f;  // control-point x range of the device-space segments
};

struct PathGeom {
    EdgeSet fill, stroke;
    bool hair = false;               // the stroke is an anti-aliased hairline
    float hair_opacity = 1.0f;       // folded into its paint
    std::vector<HairLine> hair_lines;
};

// Device-space geometry of one path drawn into a pw x ph pixmap: tiny-skia's
// fill_path preconditions and edges, and (unless `fill_only`) the stroke outline's
// edges or, for a stroke at most one device pixel wide, its hairline segments.
void compute_geom(const rvk_scene &sc, const rvk_path &p, Ts ts, uint32_t pw, uint32_t ph, bool fill_only,
                  PathGeom &g) {
    auto finish = [&](std::vector<Seg> &segs, EdgeSet &es) {
        if (segs.empty() || bounds_nearly_zero(segs)) return;
        float mnx = INFINITY, mxx = -INFINITY;
        for (const Seg &s : segs)
            for (int i = 0; i < s.n; i++) { mnx = std::min(mnx, s.p[i].x); mxx = std::max(mxx, s.p[i].x); }
        if (!std::isfinite(mnx) || !std::isfinite(mxx)) return;
        build_edges(segs, p.anti_alias ? 2 : 0, es.edges, es.caps);
        es.mnx = mnx;
        es.mxx = mxx;
        es.ok = !es.edges.empty();
    };
    if (!p.visible) return;
    if (p.fill.kind != RVK_PAINT_NONE && p.bounds.w != 0.0f && p.bounds.h != 0.0f) {
        std::vector<Seg> segs;
        fill_segments(sc, p, ts, segs);
        finish(segs, g.fill);
    }
    if (!fill_only && p.stroke.kind != RVK_PAINT_NONE && p.stroke_width >= 0.0f) {
        float cov = 1.0f;
        if (rvks::treat_as_hairline(p.stroke_width, p.anti_alias != 0, ts.sx, ts.ky, ts.kx, ts.sy, cov)) {
            g.hair = true;
            g.hair_opacity = rvks::hairline_paint_opacity(cov);
            hairline_lines(sc, p, ts, pw, ph, g.hair_lines);
            return;
        }
        std::vector<Seg> segs;
        stroke_segments(sc, p, ts, segs);
        finish(segs, g.stroke);
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

template <class F>
void parallel_for(size_t n, F &&fn) {
    const unsigned t = (unsigned)std::min<size_t>(host_threads(), n);
    if (t <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
    };
    std::vector<std::thread> pool;
    for (unsigned k = 1; k < t; k++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
}

struct Renderer {
    rvk_context *ctx;
    const rvk_scene &sc;
    IRect max_bbox;
    rvk_status err = RVK_OK;
    uint64_t overflow_rows = 0;
    int32_t *dev_overflow = nullptr;  // per render: crossings or rows the kernels could not hold
    const uint64_t budget = arena_budget();

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
    };
    std::vector<Task> tasks;
    std::vector<PathGeom> geoms;
    size_t next_task = 0, ready_end = 0;
    static constexpr size_t kChunk = 256;

    PathGeom take_geom() {
        const size_t k = next_task++;
        if (k >= tasks.size()) return PathGeom{};
        if (k >= ready_end) {
            const size_t end = std::min(tasks.size(), k + kChunk);
            parallel_for(end - k, [&](size_t i) {
                const Task &t = tasks[k + i];
                compute_geom(*t.scene, *t.path, t.ts, t.pw, t.ph, t.fill_only, geoms[k + i]);
            });
            ready_end = end;
        }
        return std::move(geoms[k]);
    }

    // Adds a coverage job for device-space segments, restricted to `clip` (pixels
    // of the job's pixmap).  Returns the job index or -1 when nothing is covered.
    int32_t add_job(DrawList &dl, const EdgeSet &es, uint32_t rule, bool aa, IRect clip) {
        if (!es.ok) return -1;
        const int shift = aa ? 2 : 0;
        const std::vector<EdgeRec> &er = es.edges;
        const std::vector<uint32_t> &caps = es.caps;
        // last covered row of each edge: exact for lines; for curves the row of the
        // end point plus one (the stepped points stay within a fraction of a row of
        // the monotonic curve; the expansion kernel checks the bound)
        std::vector<int32_t> ylast(er.size());
        int32_t ymin = INT32_MAX, ymax = INT32_MIN;
        for (size_t i = 0; i < er.size(); i++) {
            const EdgeRec &e = er[i];
            ylast[i] = e.type == 0 ? e.last_y : (((e.ly >> 10) + 32) >> 6) + 1;
            ymin = std::min(ymin, e.first_y);
            ymax = std::max(ymax, ylast[i]);
        }
        int64_t x0 = std::max<int64_t>((int64_t)std::floor(es.mnx) - 1, clip.x);
        int64_t x1 = std::min<int64_t>((int64_t)std::ceil(es.mxx) + 2, (int64_t)clip.x + clip.w);
        int64_t y0 = std::max<int64_t>(ymin >> shift, clip.y);
        int64_t y1 = std::min<int64_t>((ymax >> shift) + 1, (int64_t)clip.y + clip.h);
        if (x1 <= x0 || y1 <= y0) return -1;
        CovJob j{};
        j.x0 = (int32_t)x0; j.y0 = (int32_t)y0; j.w = (int32_t)(x1 - x0); j.h = (int32_t)(y1 - y0);
        j.rule = rule;
        j.aa = aa ? 1 : 0;
        // edges sorted by first row (ids keep tiny-skia's edge-list order)
        std::vector<uint32_t> order(er.size());
        for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return er[a].first_y < er[b].first_y; });
        j.edge_offset = (uint32_t)dl.edges.size();
        j.edge_count = (uint32_t)er.size();
        std::vector<int32_t> prefmax(er.size());
        int32_t pm = INT32_MIN;
        for (size_t k = 0; k < order.size(); k++) {
            const uint32_t i = order[k];
            EdgeRec e = er[i];
            e.piece_offset = (uint32_t)dl.piece_slots;
            dl.piece_slots += caps[i];
            dl.edges.push_back(e);
            dl.ecov.push_back(EdgeCov{e.first_y, ylast[i], e.winding, (int32_t)i, e.piece_offset, 0});
            pm = std::max(pm, ylast[i]);
            prefmax[k] = pm;
        }
        // per group of kGroupRows pixel rows: the edges that start at or before the
        // group's last row, minus the leading ones that all end before its first row
        j.window_offset = (uint32_t)dl.windows.size();
        const int32_t groups = (j.h + kGroupRows - 1) / kGroupRows;
        uint32_t lo = 0, hi = 0;
        for (int32_t g = 0; g < groups; g++) {
            const int32_t r0 = j.y0 + g * kGroupRows, r1 = std::min(j.y0 + j.h, r0 + kGroupRows);
            const int32_t s0 = r0 << shift, s1 = (r1 << shift) - 1;
            while (lo < j.edge_count && prefmax[lo] < s0) lo++;
            while (hi < j.edge_count && dl.ecov[j.edge_offset + hi].first_y <= s1) hi++;
            dl.windows.push_back(uint2{lo, std::max(lo, hi)});
        }
        // edges whose row range meets each group (difference array): an edge crosses a
        // sub-scanline at most once, so this bounds every sub-scanline's crossings
        std::vector<int32_t> diff((size_t)groups + 1, 0);
        for (size_t i = 0; i < er.size(); i++) {
            const int64_t fr = (int64_t)(er[i].first_y >> shift) - j.y0, lr = (int64_t)(ylast[i] >> shift) - j.y0;
            if (lr < 0 || fr >= j.h) continue;
            const int64_t ga = std::max<int64_t>(fr, 0) / kGroupRows, gb = std::min<int64_t>(lr, j.h - 1) / kGroupRows;
            diff[ga] += 1;
            diff[gb + 1] -= 1;
        }
        for (int32_t g = 0, run = 0; g < groups; g++) {
            run += diff[g];
            dl.active.push_back((uint32_t)run);
        }
        const uint64_t stride = ((uint64_t)j.w + 3) & ~3ull;
        j.cov_offset = dl.arena;
        dl.arena += (stride * (uint64_t)j.h + 255) & ~255ull;
        j.row_start = dl.rows;
        dl.rows += (uint32_t)j.h;
        dl.jobs.push_back(j);
        return (int32_t)dl.jobs.size() - 1;
    }

    // Adds a hairline job for prepared lines, binned into tiles of the part of their
    // pixel bounds inside `clip`.  Returns the job index or -1 when nothing is drawn.
    int32_t add_hair_job(DrawList &dl, const std::vector<HairLine> &lines, IRect clip) {
        if (lines.empty()) return -1;
        struct Box { int32_t x0, y0, x1, y1; };
        std::vector<Box> boxes(lines.size());
        int32_t jx0 = INT32_MAX, jy0 = INT32_MAX, jx1 = INT32_MIN, jy1 = INT32_MIN;
        for (size_t i = 0; i < lines.size(); i++) {
            const HairLine &L = lines[i];
            // minor pixels touched: f >> 16 and the one before, f at both ends
            const int64_t fa = (int64_t)L.fstart + 0x8000;
            const int64_t fb = fa + (int64_t)(L.istop - L.istart - 1) * L.slope;
            const int32_t m0 = (int32_t)(std::min(fa, fb) >> 16) - 1, m1 = (int32_t)(std::max(fa, fb) >> 16) + 1;
            Box b = (L.flags & kHairYMajor) ? Box{m0, L.istart, m1, L.istop} : Box{L.istart, m0, L.istop, m1};
            if (L.flags & kHairClipped) {
                b.x0 = std::max(b.x0, L.cl); b.y0 = std::max(b.y0, L.ct);
                b.x1 = std::min(b.x1, L.cr); b.y1 = std::min(b.y1, L.cb);
            }
            b.x0 = std::max(b.x0, clip.x); b.y0 = std::max(b.y0, clip.y);
            b.x1 = std::min(b.x1, clip.x + clip.w); b.y1 = std::min(b.y1, clip.y + clip.h);
            boxes[i] = b;
            if (b.x0 >= b.x1 || b.y0 >= b.y1) continue;
            jx0 = std::min(jx0, b.x0); jy0 = std::min(jy0, b.y0);
            jx1 = std::max(jx1, b.x1); jy1 = std::max(jy1, b.y1);
        }
        if (jx0 >= jx1 || jy0 >= jy1) return -1;
        HairJob j{};
        j.x0 = jx0; j.y0 = jy0; j.w = jx1 - jx0; j.h = jy1 - jy0;
        j.tiles_x = (uint32_t)((j.w + kHairTile - 1) / kHairTile);
        j.tiles_y = (uint32_t)((j.h + kHairTile - 1) / kHairTile);
        const size_t nt = (size_t)j.tiles_x * j.tiles_y;
        std::vector<uint32_t> count(nt + 1, 0);
        auto tiles_of = [&](const Box &b, auto &&fn) {
            const int32_t tx0 = (b.x0 - j.x0) / kHairTile, tx1 = (b.x1 - 1 - j.x0) / kHairTile;
            const int32_t ty0 = (b.y0 - j.y0) / kHairTile, ty1 = (b.y1 - 1 - j.y0) / kHairTile;
            for (int32_t ty = ty0; ty <= ty1; ty++)
                for (int32_t tx = tx0; tx <= tx1; tx++) fn((size_t)ty * j.tiles_x + tx);
        };
        for (const Box &b : boxes)
            if (b.x0 < b.x1 && b.y0 < b.y1) tiles_of(b, [&](size_t t) { count[t]++; });
        j.tile_offset = (uint32_t)dl.hair_tiles.size();
        const uint32_t base = (uint32_t)dl.hair_bins.size();
        uint32_t run = base;
        for (size_t t = 0; t < nt; t++) {
            dl.hair_tiles.push_back(run);
            const uint32_t c = count[t];
            count[t] = run;  // becomes the fill cursor
            run += c;
        }
        dl.hair_tiles.push_back(run);
        dl.hair_bins.resize(run);
        const uint32_t line_base = (uint32_t)dl.hair_lines.size();
        for (size_t i = 0; i < lines.size(); i++) {
            const Box &b = boxes[i];
            if (b.x0 < b.x1 && b.y0 < b.y1)
                tiles_of(b, [&](size_t t) { dl.hair_bins[count[t]++] = line_base + (uint32_t)i; });
        }
        dl.hair_lines.insert(dl.hair_lines.end(), lines.begin(), lines.end());
        dl.hair_jobs.push_back(j);
        return (int32_t)dl.hair_jobs.size() - 1;
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
            tasks.push_back(Task{&S, &p, ts, pw, ph, fill_only});
            return;
        }
        if (mode == kSkip) {
            next_task++;
            return;
        }
        const PathGeom g = take_geom();
        if (!p.visible) return;
        ShaderParams sp;
        // path::fill_path (path.rs:26-75)
        if (g.fill.ok && make_shader(S, p.fill, ts, sp)) {
            it.fill_job = add_job(dl, g.fill, p.fill_rule, p.anti_alias != 0, clip);
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
            if (make_shader(S, p.stroke, ts, sp, g.hair_opacity)) {
                it.stroke_job = add_hair_job(dl, g.hair_lines, clip);
                if (it.stroke_job >= 0) {
                    it.stroke_hair = 1;
                    it.stroke_shader = (int32_t)dl.shaders.size();
                    dl.shaders.push_back(sp);
                }
            }
        } else if (g.stroke.ok && make_shader(S, p.stroke, ts, sp)) {
            it.stroke_job = add_job(dl, g.stroke, RVK_FILL_NONZERO, p.anti_alias != 0, clip);
            if (it.stroke_job >= 0) {
                it.stroke_shader = (int32_t)dl.shaders.size();
                dl.shaders.push_back(sp);
            }
        }
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
                const HairJob &hj = dl.hair_jobs[j];
                jx = hj.x0; jy = hj.y0; jw = hj.w; jh = hj.h;
            } else {
                const CovJob &jb = dl.jobs[j];
                jx = jb.x0; jy = jb.y0; jw = jb.w; jh = jb.h;
            }
            x0 = std::min(x0, jx + it.lx); y0 = std::min(y0, jy + it.ly);
            x1 = std::max(x1, jx + jw + it.lx); y1 = std::max(y1, jy + jh + it.ly);
        }
        it.x0 = std::max(x0, 0); it.y0 = std::max(y0, 0);
        it.x1 = std::min(x1, (int32_t)dl.w); it.y1 = std::min(y1, (int32_t)dl.h);
        if (it.x1 <= it.x0 || it.y1 <= it.y0) return;
        dl.bx0 = std::min(dl.bx0, it.x0); dl.by0 = std::min(dl.by0, it.y0);
        dl.bx1 = std::max(dl.bx1, it.x1); dl.by1 = std::max(dl.by1, it.y1);
        dl.items.push_back(it);
        ctx->stats.paths++;
    }

    rvk_status flush(DrawList &dl) {
        if (dl.items.empty()) {
            dl = DrawList{dl.px, dl.w, dl.h};
            return RVK_OK;
        }
        // one device block: jobs | edges | ecov | windows | shaders | items || npieces |
        // pieces | arena (the part after || is written by the device).  The metadata is
        // packed straight into a pinned staging buffer and copied without a host wait;
        // device buffers are reused only by later work on the same stream.
        // row groups by the variant of the coverage kernel that handles them: a
        // group's edge window bounds the crossings of any of its sub-scanlines
        std::vector<uint32_t> groups_tier[3];
        for (const CovJob &j : dl.jobs) {
            const int32_t ng = (j.h + kGroupRows - 1) / kGroupRows;
            for (int32_t g = 0; g < ng; g++) {
                const uint32_t n = dl.active[j.window_offset + g];
                const int tier = n <= kCoverageKeys[0] ? 0 : n <= kCoverageKeys[1] ? 1 : 2;
                groups_tier[tier].push_back(j.row_start + (uint32_t)(g * kGroupRows));
            }
        }
        auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
        const size_t o_jobs = 0;
        const size_t o_edges = al(o_jobs + dl.jobs.size() * sizeof(CovJob));
        const size_t o_ecov = al(o_edges + dl.edges.size() * sizeof(EdgeRec));
        const size_t o_win = al(o_ecov + dl.ecov.size() * sizeof(EdgeCov));
        const size_t o_sh = al(o_win + dl.windows.size() * sizeof(uint2));
        const size_t o_it = al(o_sh + dl.shaders.size() * sizeof(ShaderParams));
        const size_t o_hj = al(o_it + dl.items.size() * sizeof(DrawItem));
        const size_t o_ht = al(o_hj + dl.hair_jobs.size() * sizeof(HairJob));
        const size_t o_hb = al(o_ht + dl.hair_tiles.size() * 4);
        const size_t o_hl = al(o_hb + dl.hair_bins.size() * 4);
        size_t o_grp[3];
        o_grp[0] = al(o_hl + dl.hair_lines.size() * sizeof(HairLine));
        o_grp[1] = al(o_grp[0] + groups_tier[0].size() * 4);
        o_grp[2] = al(o_grp[1] + groups_tier[1].size() * 4);
        const size_t meta = al(o_grp[2] + groups_tier[2].size() * 4);
        const size_t o_np = meta;
        const size_t o_pc = al(o_np + dl.edges.size() * 4);
        const size_t o_arena = al(o_pc + (size_t)dl.piece_slots * sizeof(PieceDev));
        void *hostv = nullptr;
        uint32_t slot = 0;
        RVK_TRY(rvk::staging_acquire(ctx, meta, &hostv, &slot));
        uint8_t *host = (uint8_t *)hostv;
        auto put = [&](size_t off, const void *src, size_t n) { if (n) std::memcpy(host + off, src, n); };
        put(o_jobs, dl.jobs.data(), dl.jobs.size() * sizeof(CovJob));
        put(o_edges, dl.edges.data(), dl.edges.size() * sizeof(EdgeRec));
        put(o_ecov, dl.ecov.data(), dl.ecov.size() * sizeof(EdgeCov));
        put(o_win, dl.windows.data(), dl.windows.size() * sizeof(uint2));
        put(o_sh, dl.shaders.data(), dl.shaders.size() * sizeof(ShaderParams));
        put(o_it, dl.items.data(), dl.items.size() * sizeof(DrawItem));
        put(o_hj, dl.hair_jobs.data(), dl.hair_jobs.size() * sizeof(HairJob));
        put(o_ht, dl.hair_tiles.data(), dl.hair_tiles.size() * 4);
        put(o_hb, dl.hair_bins.data(), dl.hair_bins.size() * 4);
        put(o_hl, dl.hair_lines.data(), dl.hair_lines.size() * sizeof(HairLine));
        for (int t = 0; t < 3; t++) put(o_grp[t], groups_tier[t].data(), groups_tier[t].size() * 4);
        void *dev = nullptr;
        RVK_TRY(rvk::scratch_alloc(ctx, o_arena + dl.arena, &dev));
        uint8_t *d = (uint8_t *)dev;
        // algorithmic bytes: expansion reads each edge once and writes its pieces;
        // coverage writes one byte per job pixel and reads edges, windows and pieces
        // once; the compositor reads and writes each target pixel of its rect once
        // and reads each coverage byte once
        double cov_px = 0.0;
        for (const CovJob &j : dl.jobs) cov_px += (double)j.w * (double)j.h;
        const double piece_bytes = (double)dl.piece_slots * sizeof(PieceDev);
        const double exp_bytes = (double)dl.edges.size() * (sizeof(EdgeRec) + 4) + piece_bytes;
        const double cov_bytes = cov_px + (double)dl.ecov.size() * sizeof(EdgeCov) +
                                 (double)dl.windows.size() * sizeof(uint2) + piece_bytes;
        const double draw_bytes = 8.0 * (double)(dl.bx1 - dl.bx0) * (double)(dl.by1 - dl.by0) + cov_px +
                                  (double)(dl.hair_lines.size() * sizeof(HairLine) + dl.hair_bins.size() * 4);
        ctx->stats.edge_pieces += dl.piece_slots;
        ctx->stats.band_entries += dl.windows.size();
        ctx->stats.upload_bytes += meta;
        rvk_status st = rvk::upload_staged(ctx, slot, d, meta);
        if (st == RVK_OK)
            st = rvk::launch_expand(ctx, (const EdgeRec *)(d + o_edges), (uint32_t)dl.edges.size(),
                                    (const EdgeCov *)(d + o_ecov), (PieceDev *)(d + o_pc), (uint32_t *)(d + o_np),
                                    dev_overflow, exp_bytes);
        if (st == RVK_OK)
            st = rvk::launch_coverage(ctx, (const CovJob *)(d + o_jobs), (uint32_t)dl.jobs.size(),
                                      (const uint32_t *)(d + o_grp[0]), (uint32_t)groups_tier[0].size(),
                                      (const uint32_t *)(d + o_grp[1]), (uint32_t)groups_tier[1].size(),
                                      (const uint32_t *)(d + o_grp[2]), (uint32_t)groups_tier[2].size(),
                                      (const EdgeCov *)(d + o_ecov), (const uint2 *)(d + o_win),
                                      (const PieceDev *)(d + o_pc), (const uint32_t *)(d + o_np), d + o_arena,
                                      dev_overflow, cov_bytes);
        if (st == RVK_OK)
            st = rvk::launch_draw_items(ctx, dl.px, dl.w, dl.h, dl.bx0, dl.by0, dl.bx1, dl.by1,
                                        (const DrawItem *)(d + o_it), (uint32_t)dl.items.size(),
                                        (const CovJob *)(d + o_jobs), d + o_arena, (const ShaderParams *)(d + o_sh),
                                        (const HairJob *)(d + o_hj), (const uint32_t *)(d + o_ht),
                                        (const uint32_t *)(d + o_hb), (const HairLine *)(d + o_hl), draw_bytes);
        rvk::scratch_release(ctx, dev);
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
    static constexpr int kMaxNesting = 64;  // deeper clip / mask chains are malformed input

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
        aux.scene.verbs = aux.verbs.data();
        aux.scene.points = aux.points.data();
        aux.scene.paths = aux.paths.data();
        aux.scene.path_count = (uint32_t)aux.paths.size();
    }

    void render_nodes(const rvk_group &g, Ts ts, DrawList &dl) {
        for (uint32_t i = 0; i < g.child_count && err == RVK_OK; i++) {
            const rvk_node_ref &n = sc.children[g.child_offset + i];
            if (n.kind == RVK_NODE_GROUP) render_group(sc.groups[n.index], ts, dl);
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
        for (uint32_t i = 0; s == RVK_OK && err == RVK_OK && i < g.filter_count; i++)
            s = rvk::filter_apply(ctx, &sc.filters[g.filter_offset + i], &sc.filter_data, lts, layer, ib.w, ib.h,
                                  nullptr);
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

rvk_status render_scene(rvk_context *ctx, const rvk_scene *scene, Ts ts, uint32_t w, uint32_t h, uint8_t *px,
                        uint64_t *overflow_rows) {
    Renderer r{ctx, *scene, max_filter_bbox(w, h)};
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
    DrawList root;
    root.w = w;
    root.h = h;
    r.mode = Renderer::kCollect;
    r.render_nodes(scene->groups[0], ts, root);
    r.geoms.resize(r.tasks.size());
    r.mode = Renderer::kExecute;
    root = DrawList{};
    root.px = px;
    root.w = w;
    root.h = h;
    r.render_nodes(scene->groups[0], ts, root);
    if (r.err == RVK_OK) r.fail(r.flush(root));
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
    r.overflow_rows = (uint64_t)(ov > 0 ? ov : 0);
    if (overflow_rows) *overflow_rows = r.overflow_rows;
    if (r.err == RVK_OK && r.overflow_rows) return RVK_ERR_UNSUPPORTED;
    return r.err;
}

}  // namespace rvk

extern "C" {

static bool scene_ok(const rvk_scene *s) {
    return s && s->group_count > 0 && s->groups && (s->path_count == 0 || (s->paths && s->verbs && s->points)) &&
           (s->clip_path_count == 0 || s->clip_paths) && (s->mask_count == 0 || s->masks);
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
