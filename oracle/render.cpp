// Oracle restatement of resvg's render walk (crates/resvg/src/lib.rs:34-43,
// render.rs:10-143, path.rs:6-24) over the flattened scene.  Test infrastructure only.
#include "oracle.h"
#include "common.h"
#include "geom.h"
#include <algorithm>
#include <cmath>
#include <vector>

using namespace orc;

namespace orc {
bool fill_path_into(const rvk_scene &sc, const rvk_path &path, Ts ts, uint8_t *px, uint32_t w, uint32_t h);
bool stroke_path_into(const rvk_scene &sc, const rvk_path &path, Ts ts, uint8_t *px, uint32_t w, uint32_t h);
bool clear_path_into(const rvk_scene &sc, const rvk_path &path, Ts ts, uint8_t *px, uint32_t w, uint32_t h);
bool draw_image_into(const rvk_image &img, Ts ts, uint8_t *px, uint32_t w, uint32_t h);
void rect_mask(const rvk_rect &r, Ts ts, uint32_t w, uint32_t h, std::vector<uint8_t> &mask);
}

namespace {

struct Ctx { IRect max_bbox; };

// convert_blend_mode (render.rs:145-164)
uint32_t convert_blend_mode(uint32_t m) {
    static const uint32_t map[RVK_MIX_COUNT] = {
        RVK_BLEND_SOURCE_OVER, RVK_BLEND_MULTIPLY, RVK_BLEND_SCREEN, RVK_BLEND_OVERLAY, RVK_BLEND_DARKEN,
        RVK_BLEND_LIGHTEN, RVK_BLEND_COLOR_DODGE, RVK_BLEND_COLOR_BURN, RVK_BLEND_HARD_LIGHT, RVK_BLEND_SOFT_LIGHT,
        RVK_BLEND_DIFFERENCE, RVK_BLEND_EXCLUSION, RVK_BLEND_HUE, RVK_BLEND_SATURATION, RVK_BLEND_COLOR,
        RVK_BLEND_LUMINOSITY};
    return m < RVK_MIX_COUNT ? map[m] : RVK_BLEND_SOURCE_OVER;
}

void render_nodes(const rvk_scene &sc, const rvk_group &g, const Ctx &ctx, Ts ts, uint8_t *px, uint32_t w,
                  uint32_t h);

// path::render (path.rs:6-24); a pattern paint renders its tile with this walk's
// Context (render_pattern_pixmap passes `ctx` on)
void render_path(const rvk_scene &sc, const rvk_path &p, const Ctx &ctx, Ts ts, uint8_t *px, uint32_t w,
                 uint32_t h) {
    if (!p.visible) return;
    const PatternHook saved = pattern_hook;
    const bool has_pattern = p.fill.kind == RVK_PAINT_PATTERN || p.stroke.kind == RVK_PAINT_PATTERN;
    if (has_pattern) {
        pattern_hook = [&sc, &ctx](uint32_t root, const float *t, uint8_t *img, uint32_t iw, uint32_t ih) {
            if (root >= sc.group_count) return false;
            render_nodes(sc, sc.groups[root], ctx, Ts{t[0], t[1], t[2], t[3], t[4], t[5]}, img, iw, ih);
            return true;
        };
    }
    auto fill = [&] {
        prof::Scope timed(prof::kFill);
        fill_path_into(sc, p, ts, px, w, h);
    };
    auto stroke = [&] {
        prof::Scope timed(prof::kStroke);
        stroke_path_into(sc, p, ts, px, w, h);
    };
    if (!p.paint_order_stroke_first) {
        fill();
        stroke();
    } else {
        stroke();
        fill();
    }
    if (has_pattern) pattern_hook = saved;
}

// ---------------------------------------------------------------------------
// clip paths (clip.rs) and masks (mask.rs)
// ---------------------------------------------------------------------------
constexpr int kMaxNesting = 64;  // clip/mask chains deeper than this are malformed input

inline uint32_t div255(uint32_t v) { return (v + 255) >> 8; }

// Pixmap::apply_mask: lowp LoadMaskU8 + DestinationIn on every channel
void apply_mask(uint8_t *px, const std::vector<uint8_t> &m) {
    for (size_t i = 0; i < m.size(); i++)
        for (int k = 0; k < 4; k++) px[4 * i + k] = (uint8_t)div255(px[4 * i + k] * m[i]);
}

// Mask::from_pixmap (tiny-skia mask.rs): the alpha channel, or the luminance of the
// demultiplied colour (Rec. 709 coefficients 0.2126, 0.7152, 0.0722) times alpha,
// scaled to 255 and rounded up; pinned by the masking/mask fixtures (only this form
// reproduces masking/mask/simple-case exactly)
void mask_from_pixmap(const std::vector<uint8_t> &px, uint32_t kind, std::vector<uint8_t> &m) {
    m.resize(px.size() / 4);
    for (size_t i = 0; i < m.size(); i++) {
        const uint8_t *p = &px[4 * i];
        if (kind == RVK_MASK_ALPHA) {
            m[i] = p[3];
            continue;
        }
        float r = (float)p[0] / 255.0f, g = (float)p[1] / 255.0f, b = (float)p[2] / 255.0f;
        const float a = (float)p[3] / 255.0f;
        if (p[3] != 0) {
            r /= a;
            g /= a;
            b /= a;
        }
        const float luma = r * 0.2126f + g * 0.7152f + b * 0.0722f;
        float v = (luma * a) * 255.0f;
        v = std::min(std::max(v, 0.0f), 255.0f);
        m[i] = (uint8_t)std::ceil(v);
    }
}

void clip_apply(const rvk_scene &sc, int32_t idx, Ts ts, uint8_t *px, uint32_t w, uint32_t h, int depth);

// clip::draw_children (clip.rs:30-70): fills of the paths, groups with their
// transforms (a group with its own clip-path goes through clip_group)
void clip_draw_children(const rvk_scene &sc, const rvk_group &parent, bool clear, Ts ts, uint8_t *px, uint32_t w,
                        uint32_t h, int depth) {
    for (uint32_t i = 0; i < parent.child_count; i++) {
        const rvk_node_ref &n = sc.children[parent.child_offset + i];
        if (n.kind == RVK_NODE_IMAGE) continue;  // not clip path content
        if (n.kind == RVK_NODE_PATH) {
            const rvk_path &p = sc.paths[n.index];
            if (!p.visible) continue;
            if (clear) clear_path_into(sc, p, ts, px, w, h);
            else fill_path_into(sc, p, ts, px, w, h);
        } else {
            const rvk_group &g = sc.groups[n.index];
            const Ts gts = ts_concat(ts, ts_from(g.transform));
            if (g.clip_path >= 0) {
                // clip_group (clip.rs:72-98): the children on their own canvas, clipped,
                // then Xor-ed into the clip pixmap
                std::vector<uint8_t> tmp((size_t)w * h * 4, 0);
                clip_draw_children(sc, g, false, gts, tmp.data(), w, h, depth + 1);
                clip_apply(sc, g.clip_path, gts, tmp.data(), w, h, depth + 1);
                orc_draw_pixmap(px, w, h, 0, 0, tmp.data(), w, h, 1.0f, RVK_BLEND_XOR);
            } else {
                clip_draw_children(sc, g, clear, gts, px, w, h, depth + 1);
            }
        }
    }
}

// clip::apply (clip.rs:6-28)
void clip_apply(const rvk_scene &sc, int32_t idx, Ts ts, uint8_t *px, uint32_t w, uint32_t h, int depth) {
    if (idx < 0 || (uint32_t)idx >= sc.clip_path_count || depth > kMaxNesting) return;
    const rvk_clip_path &cp = sc.clip_paths[idx];
    std::vector<uint8_t> clip((size_t)w * h * 4, 0);
    for (size_t i = 0; i < (size_t)w * h; i++) clip[4 * i + 3] = 255;  // Color::BLACK
    clip_draw_children(sc, sc.groups[cp.root], true, ts_concat(ts, ts_from(cp.transform)), clip.data(), w, h, depth);
    if (cp.clip_path >= 0) clip_apply(sc, cp.clip_path, ts, px, w, h, depth + 1);
    std::vector<uint8_t> m;
    mask_from_pixmap(clip, RVK_MASK_ALPHA, m);
    for (auto &v : m) v = (uint8_t)(255 - v);  // Mask::invert
    apply_mask(px, m);
}

void mask_apply(const rvk_scene &sc, const Ctx &ctx, int32_t idx, Ts ts, uint8_t *px, uint32_t w, uint32_t h,
                int depth);

bool int_rect_from_xywh(int64_t x, int64_t y, int64_t w, int64_t h, IRect &out) {
    if (w <= 0 || h <= 0 || x < INT32_MIN || y < INT32_MIN || x + w > INT32_MAX || y + h > INT32_MAX ||
        w > INT32_MAX || h > INT32_MAX)
        return false;
    out = IRect{(int32_t)x, (int32_t)y, (int32_t)w, (int32_t)h};
    return true;
}

// render_group (render.rs:49-143)
void render_group(const rvk_scene &sc, const rvk_group &g, const Ctx &ctx, Ts ts, uint8_t *px, uint32_t w,
                  uint32_t h) {
    ts = ts_concat(ts, ts_from(g.transform));
    bool isolate = g.isolate || g.opacity != 1.0f || g.clip_path >= 0 || g.mask >= 0 || g.filter_count > 0 ||
                   g.blend_mode != RVK_MIX_NORMAL;
    if (!isolate) {
        render_nodes(sc, g, ctx, ts, px, w, h);
        return;
    }
    FRect bb = rect_transform(g.layer_bbox, ts);
    if (!bb.ok) return;
    IRect ib;
    float bw = bb.r - bb.l, bh = bb.b - bb.t;
    if (g.filter_count == 0) {
        int64_t x = (int64_t)as_i32(std::floor(bb.l)) - 2, y = (int64_t)as_i32(std::floor(bb.t)) - 2;
        if (x < INT32_MIN || y < INT32_MIN) return;
        int64_t ww = (int64_t)as_u32(std::ceil(bw)) + 4, hh = (int64_t)as_u32(std::ceil(bh)) + 4;
        if (ww > UINT32_MAX || hh > UINT32_MAX) return;
        if (!int_rect_from_xywh(x, y, ww, hh, ib)) return;
        if (!fit_to_rect(ib, ctx.max_bbox, ib)) return;
    } else {
        IRect t;
        if (!int_rect_from_xywh(as_i32(std::floor(bb.l)), as_i32(std::floor(bb.t)),
                                as_u32(std::max(std::ceil(bw), 1.0f)), as_u32(std::max(std::ceil(bh), 1.0f)), t))
            return;
        if (!fit_to_rect(t, ctx.max_bbox, ib)) return;
    }
    float dx = bb.l, dy = bb.t;
    dx -= bb.l - (float)ib.x;
    dy -= bb.t - (float)ib.y;
    Ts lts = ts_concat(Ts{1, 0, 0, 1, -dx, -dy}, ts);
    std::vector<uint8_t> layer((size_t)ib.w * ib.h * 4, 0);
    render_nodes(sc, g, ctx, lts, layer.data(), ib.w, ib.h);
    if (g.filter_count > 0) {
        // feImage content: render_nodes of a scene group with max_bbox the primitive's
        // pixmap (filter/mod.rs:878-882)
        const FeImageHook saved = fe_image_hook;
        fe_image_hook = [&sc](uint32_t root, const float *t, uint8_t *img, uint32_t iw, uint32_t ih) {
            if (root >= sc.group_count) return false;
            render_nodes(sc, sc.groups[root], Ctx{IRect{0, 0, (int32_t)iw, (int32_t)ih}},
                         Ts{t[0], t[1], t[2], t[3], t[4], t[5]}, img, iw, ih);
            return true;
        };
        for (uint32_t i = 0; i < g.filter_count; i++)
            orc_filter_apply(&sc.filters[g.filter_offset + i], &sc.filter_data,
                             rvk_transform{lts.sx, lts.ky, lts.kx, lts.sy, lts.tx, lts.ty}, layer.data(), ib.w, ib.h);
        fe_image_hook = saved;
    }
    if (g.clip_path >= 0) clip_apply(sc, g.clip_path, lts, layer.data(), ib.w, ib.h, 0);
    if (g.mask >= 0) mask_apply(sc, ctx, g.mask, lts, layer.data(), ib.w, ib.h, 0);
    prof::Scope timed(prof::kLayer);
    orc_draw_pixmap(px, w, h, ib.x, ib.y, layer.data(), ib.w, ib.h, g.opacity, convert_blend_mode(g.blend_mode));
}

// mask::apply (mask.rs:6-46)
void mask_apply(const rvk_scene &sc, const Ctx &ctx, int32_t idx, Ts ts, uint8_t *px, uint32_t w, uint32_t h,
                int depth) {
    if (idx < 0 || (uint32_t)idx >= sc.mask_count || depth > kMaxNesting) return;
    const rvk_mask &mk = sc.masks[idx];
    const rvk_group &root = sc.groups[mk.root];
    if (root.child_count == 0) {
        std::fill(px, px + (size_t)w * h * 4, (uint8_t)0);
        return;
    }
    std::vector<uint8_t> mp((size_t)w * h * 4, 0);
    std::vector<uint8_t> region;
    rect_mask(mk.rect, ts, w, h, region);
    render_nodes(sc, root, ctx, ts, mp.data(), w, h);
    apply_mask(mp.data(), region);
    if (mk.mask >= 0) mask_apply(sc, ctx, mk.mask, ts, px, w, h, depth + 1);
    std::vector<uint8_t> m;
    mask_from_pixmap(mp, mk.kind, m);
    apply_mask(px, m);
}

void render_nodes(const rvk_scene &sc, const rvk_group &g, const Ctx &ctx, Ts ts, uint8_t *px, uint32_t w,
                  uint32_t h) {
    for (uint32_t i = 0; i < g.child_count; i++) {
        const rvk_node_ref &n = sc.children[g.child_offset + i];
        if (n.kind == RVK_NODE_GROUP) render_group(sc, sc.groups[n.index], ctx, ts, px, w, h);
        else if (n.kind == RVK_NODE_IMAGE) draw_image_into(sc.images[n.index], ts, px, w, h);  // image::render
        else render_path(sc, sc.paths[n.index], ctx, ts, px, w, h);
    }
}

// max_filter_bbox (lib.rs:86-97)
IRect max_filter_bbox(uint32_t w, uint32_t h) {
    int64_t x = -2 * (int64_t)std::min<uint32_t>(w, INT32_MAX), y = -2 * (int64_t)std::min<uint32_t>(h, INT32_MAX);
    int64_t ww = 5 * (int64_t)w, hh = 5 * (int64_t)h;
    IRect r;
    if (x >= INT32_MIN && y >= INT32_MIN && int_rect_from_xywh(x, y, std::min<int64_t>(ww, UINT32_MAX),
                                                               std::min<int64_t>(hh, UINT32_MAX), r))
        return r;
    return IRect{INT32_MIN / 2, INT32_MIN / 2, INT32_MAX / 2 - INT32_MIN / 2, INT32_MAX / 2 - INT32_MIN / 2};
}

}  // namespace

int32_t orc_render(const rvk_scene *scene, rvk_transform ts, uint32_t w, uint32_t h, uint8_t *px) {
    if (!scene || scene->group_count == 0) return RVK_ERR_INVALID_ARGUMENT;
    Ctx ctx{max_filter_bbox(w, h)};
    render_nodes(*scene, scene->groups[0], ctx, ts_from(ts), px, w, h);
    return RVK_OK;
}

// resvg::render_node (lib.rs:55-70): None (0) without a layer bounding box, else the
// node (render.rs:21-47) drawn with the transform pre-translated by minus the box's
// origin (1)
int32_t orc_render_node(const rvk_scene *scene, uint32_t kind, uint32_t index, const rvk_rect *bbox,
                        rvk_transform ts, uint32_t w, uint32_t h, uint8_t *px) {
    if (!bbox || !(bbox->w > 0.0f) || !(bbox->h > 0.0f) || !std::isfinite(bbox->x) || !std::isfinite(bbox->y) ||
        !std::isfinite(bbox->w) || !std::isfinite(bbox->h))
        return 0;
    Ctx ctx{max_filter_bbox(w, h)};
    const Ts t = ts_concat(ts_from(ts), Ts{1.0f, 0.0f, 0.0f, 1.0f, -bbox->x, -bbox->y});
    if (kind == RVK_NODE_GROUP) render_group(*scene, scene->groups[index], ctx, t, px, w, h);
    else if (kind == RVK_NODE_IMAGE) draw_image_into(scene->images[index], t, px, w, h);
    else render_path(*scene, scene->paths[index], ctx, t, px, w, h);
    return 1;
}

// tiny_skia::Mask::from_pixmap and Pixmap::apply_mask on their own
void orc_mask_from_pixmap(const uint8_t *px, uint32_t w, uint32_t h, uint32_t luminance, uint8_t *mask) {
    std::vector<uint8_t> p(px, px + (size_t)w * h * 4), m;
    mask_from_pixmap(p, luminance ? RVK_MASK_LUMINANCE : RVK_MASK_ALPHA, m);
    std::copy(m.begin(), m.end(), mask);
}
void orc_apply_mask(uint8_t *px, uint32_t w, uint32_t h, const uint8_t *mask) {
    std::vector<uint8_t> m(mask, mask + (size_t)w * h);
    apply_mask(px, m);
}

// ---------------------------------------------------------------------------
// per-operation time split (CPU baseline only)
// ---------------------------------------------------------------------------
namespace orc {
thread_local FeImageHook fe_image_hook;
thread_local PatternHook pattern_hook;
namespace prof {
bool enabled = false;
double ms[kSlots] = {};
}  // namespace prof
}  // namespace orc

extern "C" void orc_prof_enable(int on) {
    orc::prof::enabled = on != 0;
    std::fill(orc::prof::ms, orc::prof::ms + orc::prof::kSlots, 0.0);
}

extern "C" uint32_t orc_prof_read(double *out, uint32_t n) {
    const uint32_t k = std::min<uint32_t>(n, orc::prof::kSlots);
    for (uint32_t i = 0; i < k; i++) out[i] = orc::prof::ms[i];
    return k;
}
