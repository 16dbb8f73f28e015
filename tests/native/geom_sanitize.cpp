// Host geometry under AddressSanitizer and UndefinedBehaviorSanitizer: the product's
// path iteration, edge clipping, edge building, coverage-job and hairline-job
// preparation, stroker, dasher and shader setup (raster_host.cpp, edge_clip.cpp,
// stroker.cpp, dash.cpp, hairline.cpp) run over seeded random scenes, including
// paths far outside the pixmap, huge and degenerate coordinates, NaN, dashes and
// hairlines.  Every prepared job is checked against the invariants the device
// kernels index by (windows inside the job's edges, increasing piece offsets within
// the job's pieces, one coverage tier per band, hairline bins inside the lines).
//
// Built by tests/test_host_sanitizers.py with g++ -fsanitize=address,undefined
// -fno-sanitize-recover=all; prints "ok <jobs> <hair jobs>" and exits 0, or names the
// first broken invariant and exits 1.
#include "../../resvg_amd/csrc/raster.h"
#include "../../resvg_amd/csrc/stroker.h"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace rvkr;
using rvkh::IRect;
using rvkh::Ts;

namespace {

int fail(const char *what, long a, long b) {
    std::printf("invariant broken: %s (%ld, %ld)\n", what, a, b);
    std::exit(1);
}

void check_job(const JobPrep &P) {
    if (!P.ok) return;
    const CovJob &j = P.j;
    if (P.edges.size() != j.edge_count || P.ecov.size() != j.edge_count) fail("edge arrays", (long)P.edges.size(), j.edge_count);
    if (P.windows.size() != j.tiles_y || P.tier.size() != j.tiles_y) fail("bands", (long)P.windows.size(), j.tiles_y);
    if (j.tiles_x > kMaxTilesX) fail("tiles_x", j.tiles_x, kMaxTilesX);
    if (j.w <= 0 || j.h <= 0) fail("job rect", j.w, j.h);
    if (j.x0 < j.gx0 || j.x0 + j.w > j.gx0 + (int64_t)j.tiles_x * kTile) fail("grid x", j.x0, j.gx0);
    if (j.y0 < j.gy0 || j.y0 + j.h > j.gy0 + (int64_t)j.tiles_y * kTile) fail("grid y", j.y0, j.gy0);
    uint32_t tiers = 0;
    for (int t = 0; t < 3; t++) tiers += P.tier_count[t];
    if (tiers != j.tiles_y) fail("tier counts", tiers, j.tiles_y);
    for (size_t g = 0; g < P.windows.size(); g++) {
        if (P.windows[g].x > P.windows[g].y || P.windows[g].y > j.edge_count) fail("window", (long)g, P.windows[g].y);
        if (P.tier[g] > 2) fail("tier", (long)g, P.tier[g]);
    }
    uint64_t prev = 0;
    for (size_t k = 0; k < P.edges.size(); k++) {
        const uint64_t o = P.edges[k].piece_offset;
        if (o < prev || o > P.pieces) fail("piece offset", (long)k, (long)o);
        if (P.ecov[k].piece_offset != P.edges[k].piece_offset) fail("piece mirror", (long)k, (long)o);
        if ((uint32_t)P.ecov[k].id >= j.edge_count) fail("edge id", (long)k, P.ecov[k].id);
        if (k && P.ecov[k].first_y < P.ecov[k - 1].first_y) fail("edge order", (long)k, P.ecov[k].first_y);
        prev = o;
    }
}

void check_hair(const HairPrep &P) {
    if (!P.ok) return;
    const uint64_t tiles = (uint64_t)P.j.tiles_x * P.j.tiles_y;
    if (P.tiles.size() != tiles + 1) fail("hair tiles", (long)P.tiles.size(), (long)tiles + 1);
    for (size_t t = 0; t + 1 < P.tiles.size(); t++)
        if (P.tiles[t] > P.tiles[t + 1]) fail("hair tile order", (long)t, P.tiles[t]);
    if (P.tiles.back() != P.bins.size()) fail("hair bin end", P.tiles.back(), (long)P.bins.size());
    for (size_t b = 0; b < P.bins.size(); b++)
        if (P.bins[b] >= P.lines.size()) fail("hair bin", (long)b, P.bins[b]);
}

}  // namespace

int main(int argc, char **argv) {
    const int nscenes = argc > 1 ? std::atoi(argv[1]) : 60;
    long jobs = 0, hair = 0;
    for (int seed = 0; seed < nscenes; seed++) {
        std::mt19937 rng(1234u + (uint32_t)seed);
        auto uni = [&](float a, float b) { return std::uniform_real_distribution<float>(a, b)(rng); };
        auto pick = [&](int n) { return (int)(rng() % (uint32_t)n); };
        const uint32_t W = 32 + (uint32_t)pick(400), H = 32 + (uint32_t)pick(300);
        // coordinates: mostly around the pixmap, sometimes far outside or huge
        const float span = seed % 7 == 0 ? 1e7f : seed % 5 == 0 ? 4.0f * (float)W : 1.5f * (float)W;
        std::vector<uint8_t> verbs;
        std::vector<float> pts;
        std::vector<rvk_path> paths;
        std::vector<float> dashes;
        std::vector<rvk_stop> stops;
        for (int k = 0; k < 4; k++) stops.push_back(rvk_stop{(float)k / 3.0f, {(uint8_t)rng(), 9, 200, 0}, 1.0f});
        const int npaths = 6 + pick(10);
        for (int i = 0; i < npaths; i++) {
            rvk_path p{};
            p.verb_offset = (uint32_t)verbs.size();
            p.point_offset = (uint32_t)(pts.size() / 2);
            const int nseg = 1 + pick(9);
            auto pt = [&]() {
                float x = uni(-0.25f * span, span), y = uni(-0.25f * span, span);
                if (pick(40) == 0) x = NAN;
                if (pick(50) == 0) y = 1e30f;
                pts.push_back(x);
                pts.push_back(y);
            };
            verbs.push_back(RVK_VERB_MOVE);
            pt();
            for (int s = 0; s < nseg; s++) {
                const int v = 1 + pick(3);
                verbs.push_back((uint8_t)v);
                for (int q = 0; q < v; q++) pt();
            }
            if (pick(2)) verbs.push_back(RVK_VERB_CLOSE);
            p.verb_count = (uint32_t)verbs.size() - p.verb_offset;
            p.visible = 1;
            p.anti_alias = pick(5) != 0;
            p.fill.kind = pick(4) == 0 ? RVK_PAINT_LINEAR : RVK_PAINT_COLOR;
            p.fill.opacity = 1.0f;
            p.fill.x2 = 50.0f;
            p.fill.transform = rvk_transform{1, 0, 0, 1, 0, 0};
            p.fill.stop_count = 4;
            p.fill_rule = (uint32_t)pick(2);
            p.stroke.kind = pick(2) ? RVK_PAINT_COLOR : RVK_PAINT_NONE;
            p.stroke.opacity = 1.0f;
            p.stroke_width = pick(4) == 0 ? uni(0.1f, 1.0f) : uni(1.0f, 20.0f);
            p.stroke_miter_limit = uni(1.0f, 10.0f);
            p.stroke_cap = (uint32_t)pick(3);
            p.stroke_join = (uint32_t)pick(4);
            if (pick(4) == 0) {
                p.dash_offset = (uint32_t)dashes.size();
                p.dash_count = 2;
                dashes.push_back(uni(0.5f, 20.0f));
                dashes.push_back(uni(0.0f, 20.0f));
                p.dash_phase = uni(-10.0f, 10.0f);
            }
            p.bounds = rvk_rect{0.0f, 0.0f, 1.0f, 1.0f};
            paths.push_back(p);
        }
        rvk_scene sc{};
        sc.paths = paths.data();
        sc.path_count = (uint32_t)paths.size();
        sc.verbs = verbs.data();
        sc.points = pts.data();
        sc.stops = stops.data();
        sc.dashes = dashes.data();
        const float a = uni(0.3f, 2.0f), k = pick(3) == 0 ? uni(-0.5f, 0.5f) : 0.0f;
        const Ts ts{a, k, -k, a, uni(-50.0f, 50.0f), uni(-50.0f, 50.0f)};
        const IRect clip{0, 0, (int32_t)W, (int32_t)H};
        for (const rvk_path &p : paths) {
            auto finish = [&](std::vector<Seg> &segs, uint32_t rule) {
                if (!segs_drawable(segs) || bounds_nearly_zero(segs)) return;
                EdgeSet es;
                float mnx = INFINITY, mxx = -INFINITY;
                for (const Seg &s : segs)
                    for (int q = 0; q < s.n; q++) { mnx = std::fmin(mnx, s.p[q].x); mxx = std::fmax(mxx, s.p[q].x); }
                if (segs_inside_pixmap(segs, p.anti_alias != 0, W, H)) {
                    build_edges(segs, p.anti_alias ? 2 : 0, es.edges, es.caps);
                } else {
                    std::vector<Seg> clipped;
                    clip_segments(segs, W, H, clipped);
                    for (const Seg &s : clipped)  // end points inside (control points may lie out)
                        for (int q = 0; q < s.n; q += s.n - 1)
                            if (!(s.p[q].x >= 0.0f && s.p[q].x <= (float)W && s.p[q].y >= 0.0f && s.p[q].y <= (float)H))
                                fail("clipped point outside", (long)s.p[q].x, (long)s.p[q].y);
                    build_edges(clipped, p.anti_alias ? 2 : 0, es.edges, es.caps, true);
                }
                es.mnx = mnx;
                es.mxx = mxx;
                es.ok = !es.edges.empty();
                JobPrep P;
                prep_job(es, rule, p.anti_alias != 0, clip, pick(3) * 8 - 8, pick(3) * 4, P);
                check_job(P);
                jobs += P.ok;
            };
            std::vector<Seg> segs;
            fill_segments(sc, p, ts, segs);
            finish(segs, p.fill_rule);
            ShaderParams sp;
            (void)make_shader(sc, p.fill, ts, sp);
            if (p.stroke.kind != RVK_PAINT_NONE) {
                float cov = 1.0f;
                if (rvks::treat_as_hairline(p.stroke_width, p.anti_alias != 0, ts.sx, ts.ky, ts.kx, ts.sy, cov)) {
                    std::vector<HairLine> lines;
                    hairline_lines(sc, p, ts, W, H, lines);
                    HairPrep hp;
                    prep_hair(std::move(lines), clip, hp);
                    check_hair(hp);
                    hair += hp.ok;
                } else {
                    std::vector<Seg> ssegs;
                    stroke_segments(sc, p, ts, ssegs);
                    finish(ssegs, RVK_FILL_NONZERO);
                }
            }
            // a pattern paint over the path: tile size and shader setup
            rvk_paint pat{};
            pat.kind = RVK_PAINT_PATTERN;
            pat.opacity = uni(0.0f, 1.0f);
            pat.transform = rvk_transform{uni(-2, 2), uni(-1, 1), uni(-1, 1), uni(-2, 2), uni(-9, 9), uni(-9, 9)};
            pat.rect = rvk_rect{uni(-5, 5), uni(-5, 5), uni(0.0f, 40.0f), uni(0.0f, 40.0f)};
            float sx, sy;
            uint32_t tw, th;
            if (pattern_tile_size(pat, ts, sx, sy, tw, th)) {
                static uint32_t tile_stub;
                (void)make_pattern_shader(pat, ts, sx, sy, &tile_stub, tw, th, sp);
            }
        }
    }
    std::printf("ok %ld %ld\n", jobs, hair);
    return 0;
}
