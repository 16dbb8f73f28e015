// Filter-graph execution on the device: the control flow of resvg's filter::apply
// (crates/resvg/src/filter/mod.rs:335-564, 581-1147) over device images, plus the
// per-primitive C entry points.  Region math runs on the host in f32 exactly as the
// reference; every pixel operation is a gfx950 kernel.  Images are shared,
// immutable-once-published buffers (the reference's Rc<Pixmap>); a primitive that
// holds the only reference to its input works in place (Image::take).
#include "rvk_internal.h"
#include "host_geom.h"
#include <algorithm>
#include <memory>

using namespace rvkh;

namespace {

struct DImg {
    rvk_context *ctx = nullptr;
    uint8_t *px = nullptr;
    uint32_t w = 0, h = 0;
    IRect region{};
    uint32_t cs = RVK_CI_SRGB;
    bool owned = true;
    ~DImg() {
        if (owned && px) rvk::scratch_release(ctx, px);
    }
    size_t bytes() const { return (size_t)w * h * 4; }
};
using Img = std::shared_ptr<DImg>;

struct FilterError { rvk_status code; };

#define FTRY(expr)                                   \
    do {                                             \
        rvk_status s_ = (expr);                      \
        if (s_ != RVK_OK) throw FilterError{s_};     \
    } while (0)

// Pixmap::try_create (zeroed) + Image::from_image
Img new_img(rvk_context *ctx, uint32_t w, uint32_t h, uint32_t cs, bool zero = true) {
    if (w == 0 || h == 0) throw FilterError{RVK_ERR_INVALID_REGION};
    auto im = std::make_shared<DImg>();
    im->ctx = ctx;
    im->w = w; im->h = h;
    im->cs = cs;
    im->region = IRect{0, 0, (int32_t)w, (int32_t)h};
    void *p = nullptr;
    FTRY(rvk::scratch_alloc(ctx, im->bytes(), &p));
    im->px = (uint8_t *)p;
    if (zero) FTRY(rvk::launch_fill_sub(ctx, im->px, w, h, 0, 0, w, h, 0u));
    return im;
}

Img clone_img(const Img &a) {
    Img c = new_img(a->ctx, a->w, a->h, a->cs, false);
    c->region = a->region;
    FTRY(rvk::launch_copy(a->ctx, c->px, a->px, a->bytes()));
    return c;
}

// Image::take: the pixmap itself when unshared, otherwise a copy.
Img take(Img a) {
    if (a.use_count() == 1 && a->owned) return a;
    return clone_img(a);
}

// Image::into_color_space (mod.rs:286-308)
Img into_cs(Img im, uint32_t cs) {
    if (im->cs == cs) return im;
    IRect region = im->region;
    Img c = take(std::move(im));
    FTRY(rvk::launch_into_cs(c->ctx, c->px, c->w, c->h, cs == RVK_CI_LINEAR_RGB));
    c->cs = cs;
    c->region = region;
    return c;
}

Img from_image(Img im, uint32_t cs) {
    im->region = IRect{0, 0, (int32_t)im->w, (int32_t)im->h};
    im->cs = cs;
    return im;
}

void draw(const Img &dst, const Img &src, int32_t x, int32_t y, uint32_t mode) {
    FTRY(rvk::launch_draw_pixmap(dst->ctx, dst->px, dst->w, dst->h, x, y, src->px, src->w, src->h, 1.0f, mode));
}

struct Exec {
    rvk_context *ctx;
    const rvk_filter_data *data;
    Ts ts;
    IRect region;
    Img source;  // the layer, borrowed
    std::vector<std::pair<int32_t, Img>> results;
    const rvk::FeImageHook *fe_image = nullptr;  // draws feImage content (render walks only)

    // get_input (mod.rs:523-564)
    Img input(int32_t in) {
        if (in >= 0) {
            for (auto it = results.rbegin(); it != results.rend(); ++it)
                if (it->first == in) return it->second;
            in = RVK_INPUT_SOURCE_GRAPHIC;
        }
        if (in == RVK_INPUT_SOURCE_ALPHA) {
            Img c = clone_img(source);
            c->region = region;
            c->cs = RVK_CI_SRGB;
            FTRY(rvk::launch_zero_rgb(ctx, c->px, (size_t)c->w * c->h));
            return c;
        }
        // SourceGraphic: the reference clones the layer for every use; here each use
        // gets an unowned view of it, which take() copies the first time a primitive
        // writes to it (the layer itself is not written until the graph has run)
        auto c = std::make_shared<DImg>();
        c->ctx = ctx;
        c->px = source->px;
        c->w = source->w;
        c->h = source->h;
        c->owned = false;
        c->region = region;
        c->cs = RVK_CI_SRGB;
        return c;
    }

    void blur(const Img &im, double sx, double sy, bool box) {
        if (box) FTRY(rvk::launch_box_blur(ctx, im->px, im->w, im->h, sx, sy));
        else FTRY(rvk::launch_iir_blur(ctx, im->px, im->w, im->h, sx, sy));
    }

    Img primitive(const rvk_primitive &p, IRect sub) {
        const uint32_t cs = p.color_interpolation;
        switch (p.kind) {
            case RVK_FE_BLEND: {  // mod.rs:695-729
                Img i1 = into_cs(input(p.in1), cs);
                Img i2 = into_cs(input(p.in2), cs);
                Img out = new_img(ctx, region.w, region.h, cs);
                draw(out, i2, 0, 0, RVK_BLEND_SOURCE_OVER);
                if (p.u.blend.mode >= RVK_BLEND_COUNT) throw FilterError{RVK_ERR_INVALID_ARGUMENT};
                draw(out, i1, 0, 0, p.u.blend.mode);
                return out;
            }
            case RVK_FE_DROP_SHADOW: {  // mod.rs:581-643
                const auto &d = p.u.drop_shadow;
                Img in = input(p.in1);
                float dx, dy;
                scale_coords(d.dx, d.dy, ts, dx, dy);
                Img out = new_img(ctx, in->w, in->h, cs);
                Img inp = into_cs(in, cs);
                in.reset();
                Img shadow = clone_img(inp);
                double sx, sy;
                bool box;
                if (resolve_std_dev(d.std_dev_x, d.std_dev_y, ts, sx, sy, box)) blur(shadow, sx, sy, box);
                FTRY(launch_shadow_flood(ctx, shadow->px, (size_t)shadow->w * shadow->h, d.color, d.opacity));
                FTRY(rvk::launch_into_cs(ctx, shadow->px, shadow->w, shadow->h, cs == RVK_CI_LINEAR_RGB));
                draw(out, shadow, as_i32(dx), as_i32(dy), RVK_BLEND_SOURCE_OVER);
                draw(out, inp, 0, 0, RVK_BLEND_SOURCE_OVER);
                return out;
            }
            case RVK_FE_FLOOD: {  // mod.rs:820-835
                Img out = new_img(ctx, region.w, region.h, RVK_CI_SRGB, false);
                uint8_t rgba[4] = {p.u.flood.color[0], p.u.flood.color[1], p.u.flood.color[2],
                                   opacity_to_u8(p.u.flood.opacity)};
                uint8_t pm[4];
                premultiply_color(rgba, pm);
                FTRY(rvk::launch_fill(ctx, out->px, out->w, out->h, pm));
                return out;
            }
            case RVK_FE_GAUSSIAN_BLUR: {  // mod.rs:645-666
                Img in = input(p.in1);
                double sx, sy;
                bool box;
                if (!resolve_std_dev(p.u.blur.std_dev_x, p.u.blur.std_dev_y, ts, sx, sy, box)) return in;
                Img out = take(into_cs(std::move(in), cs));
                blur(out, sx, sy, box);
                return from_image(out, cs);
            }
            case RVK_FE_OFFSET: {  // mod.rs:668-693
                Img in = input(p.in1);
                float dx, dy;
                scale_coords(p.u.offset.dx, p.u.offset.dy, ts, dx, dy);
                if (approx_zero_ulps(dx, 4) && approx_zero_ulps(dy, 4)) return in;
                Img out = new_img(ctx, in->w, in->h, in->cs);
                draw(out, in, as_i32(dx), as_i32(dy), RVK_BLEND_SOURCE_OVER);
                return out;
            }
            case RVK_FE_COMPOSITE: {  // mod.rs:731-793
                Img i1 = into_cs(input(p.in1), cs);
                Img i2 = into_cs(input(p.in2), cs);
                Img out = new_img(ctx, region.w, region.h, cs);
                const auto &k = p.u.composite;
                if (k.op == RVK_COMPOSITE_ARITHMETIC) {
                    size_t n = std::min({i1->bytes(), i2->bytes(), out->bytes()}) / 4;
                    FTRY(rvk::launch_arithmetic(ctx, i1->px, i2->px, out->px, n, k.k1, k.k2, k.k3, k.k4));
                    return out;
                }
                if (k.op > RVK_COMPOSITE_ARITHMETIC) throw FilterError{RVK_ERR_INVALID_ARGUMENT};
                draw(out, i2, 0, 0, RVK_BLEND_SOURCE_OVER);
                uint32_t mode = k.op == RVK_COMPOSITE_IN ? RVK_BLEND_SOURCE_IN
                              : k.op == RVK_COMPOSITE_OUT ? RVK_BLEND_SOURCE_OUT
                              : k.op == RVK_COMPOSITE_ATOP ? RVK_BLEND_SOURCE_ATOP
                              : k.op == RVK_COMPOSITE_XOR ? RVK_BLEND_XOR : RVK_BLEND_SOURCE_OVER;
                draw(out, i1, 0, 0, mode);
                return out;
            }
            case RVK_FE_MERGE: {  // mod.rs:795-818
                Img out = new_img(ctx, region.w, region.h, cs);
                for (uint32_t i = 0; i < p.u.merge.input_count; i++) {
                    if (!data->inputs) throw FilterError{RVK_ERR_INVALID_ARGUMENT};
                    Img in = into_cs(input(data->inputs[p.u.merge.input_offset + i]), cs);
                    draw(out, in, 0, 0, RVK_BLEND_SOURCE_OVER);
                }
                return out;
            }
            case RVK_FE_COMPONENT_TRANSFER: {  // mod.rs:891-903 (demultiply, transfer, multiply fused)
                Img out = take(into_cs(input(p.in1), cs));
                FTRY(rvk::launch_component_transfer(ctx, out->px, out->w, out->h, &p, data->values, 1));
                return from_image(out, cs);
            }
            case RVK_FE_COLOR_MATRIX: {  // mod.rs:905-917 (demultiply, matrix, multiply fused)
                Img out = take(into_cs(input(p.in1), cs));
                FTRY(rvk::launch_color_matrix(ctx, out->px, out->w, out->h, &p, 1));
                return from_image(out, cs);
            }
            case RVK_FE_CONVOLVE_MATRIX: {  // mod.rs:919-933
                Img out = take(into_cs(input(p.in1), cs));
                if (p.u.convolve.preserve_alpha) FTRY(rvk::launch_demultiply(ctx, out->px, out->w, out->h));
                FTRY(rvk::launch_convolve(ctx, out->px, out->w, out->h, &p, data->values));
                return from_image(out, cs);
            }
            case RVK_FE_MORPHOLOGY: {  // mod.rs:935-956
                Img in = into_cs(input(p.in1), cs);
                Img out = new_img(ctx, in->w, in->h, cs, false);  // written whole; the kernel reads `in` only
                out->region = in->region;
                float rx, ry;
                scale_coords(p.u.morphology.radius_x, p.u.morphology.radius_y, ts, rx, ry);
                if (!(rx > 0.0f && ry > 0.0f))
                    FTRY(rvk::launch_fill_sub(ctx, out->px, out->w, out->h, 0, 0, out->w, out->h, 0u));
                else
                    FTRY(rvk::launch_morphology_to(ctx, in->px, out->px, out->w, out->h, p.u.morphology.op, rx, ry));
                return from_image(out, cs);
            }
            case RVK_FE_DISPLACEMENT_MAP: {  // mod.rs:958-986
                Img p1 = into_cs(input(p.in1), cs);
                Img p2 = into_cs(input(p.in2), cs);
                Img out = new_img(ctx, region.w, region.h, cs);
                float sx, sy;
                scale_coords(p.u.displacement.scale, p.u.displacement.scale, ts, sx, sy);
                if (p1->w == out->w && p1->h == out->h && p2->w == out->w && p2->h == out->h)
                    FTRY(rvk::launch_displacement(ctx, p1->px, p2->px, out->px, out->w, out->h, sx, sy,
                                                  p.u.displacement.scale, p.u.displacement.x_channel,
                                                  p.u.displacement.y_channel));
                return out;
            }
            case RVK_FE_TURBULENCE: {  // mod.rs:988-1018 (multiply_alpha fused)
                float sx, sy;
                ts_get_scale(ts, sx, sy);
                bool degenerate = approx_zero_ulps(sx, 4) || approx_zero_ulps(sy, 4);
                Img out = new_img(ctx, region.w, region.h, cs, degenerate);
                if (degenerate) return out;
                const auto &t = p.u.turbulence;
                FTRY(rvk::launch_turbulence(ctx, out->px, out->w, out->h, (double)region.x - (double)ts.tx,
                                            (double)region.y - (double)ts.ty, sx, sy, t.base_frequency_x,
                                            t.base_frequency_y, t.num_octaves, t.seed, t.stitch_tiles,
                                            t.fractal_noise, 1));
                return out;
            }
            case RVK_FE_DIFFUSE_LIGHTING:
            case RVK_FE_SPECULAR_LIGHTING: {  // mod.rs:1020-1060
                Img in = input(p.in1);
                Img out = new_img(ctx, region.w, region.h, cs);
                rvk_light_source l = transform_light(p.u.lighting.light, region, ts);
                if (in->w == out->w && in->h == out->h)
                    FTRY(rvk::launch_lighting(ctx, in->px, out->px, out->w, out->h,
                                              p.kind == RVK_FE_SPECULAR_LIGHTING, p.u.lighting.surface_scale,
                                              p.u.lighting.constant, p.u.lighting.specular_exponent,
                                              p.u.lighting.color, &l));
                return out;
            }
            case RVK_FE_TILE: {  // apply_tile (mod.rs:837-859): the input as is, tagged sRGB
                Img in = input(p.in1);
                // the input's region relative to the filter region, cut to the input
                // pixmap (copy_region = clone_rect, which intersects)
                const int64_t sx = (int64_t)in->region.x - region.x, sy = (int64_t)in->region.y - region.y;
                const int64_t x0 = std::max<int64_t>(sx, 0), y0 = std::max<int64_t>(sy, 0);
                const int64_t x1 = std::min<int64_t>(sx + in->region.w, in->w);
                const int64_t y1 = std::min<int64_t>(sy + in->region.h, in->h);
                if (x1 <= x0 || y1 <= y0) throw FilterError{RVK_ERR_INVALID_REGION};
                // Pattern(Repeat) translated by (sx, sy): pixel x reads tile column
                // (x - sx) mod tw, i.e. (x + phase) mod tw
                auto phase = [](int64_t s, int64_t t) { int64_t r = (-s) % t; return (uint32_t)(r < 0 ? r + t : r); };
                Img out = new_img(ctx, region.w, region.h, RVK_CI_SRGB, false);
                FTRY(rvk::launch_tile(ctx, out->px, out->w, out->h, in->px, in->w, in->h, (uint32_t)x0, (uint32_t)y0,
                                      (uint32_t)(x1 - x0), (uint32_t)(y1 - y0), phase(sx, x1 - x0),
                                      phase(sy, y1 - y0)));
                return out;
            }
            case RVK_FE_IMAGE: {  // apply_image (mod.rs:861-889): the content drawn into a
                // region-sized pixmap with the layer scale, moved by the subregion's origin
                if (!fe_image || !*fe_image) throw FilterError{RVK_ERR_UNSUPPORTED};
                Img out = new_img(ctx, region.w, region.h, RVK_CI_SRGB);
                float sx, sy;
                ts_get_scale(ts, sx, sy);
                FTRY((*fe_image)(p.u.image.root, Ts{sx, 0.0f, 0.0f, sy, (float)sub.x, (float)sub.y}, out->px, out->w,
                                 out->h));
                return from_image(out, RVK_CI_SRGB);
            }
            default:
                throw FilterError{RVK_ERR_INVALID_ARGUMENT};
        }
    }

    // Four Clear fill_rects outside the subregion (mod.rs:471-501).
    void clear_outside(const Img &im, IRect sub) {
        const int64_t w = im->w, h = im->h;
        if (sub.y > 0) FTRY(rvk::launch_fill_sub(ctx, im->px, im->w, im->h, 0, 0, w, sub.y, 0u));
        if (sub.x > 0) FTRY(rvk::launch_fill_sub(ctx, im->px, im->w, im->h, 0, 0, sub.x, h, 0u));
        FTRY(rvk::launch_fill_sub(ctx, im->px, im->w, im->h, ir_right(sub), 0, (int64_t)ir_right(sub) + w, h, 0u));
        FTRY(rvk::launch_fill_sub(ctx, im->px, im->w, im->h, 0, ir_bottom(sub), w, (int64_t)ir_bottom(sub) + h, 0u));
    }

    // apply_inner (mod.rs:357-521)
    Img run(const rvk_filter &f) {
        FRect fr = rect_transform(f.rect, ts);
        if (!fr.ok) throw FilterError{RVK_ERR_INVALID_REGION};
        if (!fit_to_rect(frect_to_int(fr), IRect{0, 0, (int32_t)source->w, (int32_t)source->h}, region))
            throw FilterError{RVK_ERR_INVALID_REGION};
        for (uint32_t i = 0; i < f.primitive_count; i++) {
            const rvk_primitive &p = data->primitives[f.primitive_offset + i];
            FRect pr = rect_transform(p.rect, ts);
            if (!pr.ok) throw FilterError{RVK_ERR_INVALID_REGION};
            IRect sub = frect_to_int(pr);
            if (p.kind == RVK_FE_OFFSET && p.in1 >= 0) {
                for (auto it = results.rbegin(); it != results.rend(); ++it)
                    if (it->first == p.in1) { sub = it->second->region; break; }
            }
            Img res = primitive(p, sub);
            if (!ir_eq(region, sub)) {
                IRect sub2 = p.kind == RVK_FE_OFFSET ? IRect{0, 0, region.w, region.h}
                                                     : IRect{sub.x - region.x, sub.y - region.y, sub.w, sub.h};
                uint32_t rcs = res->cs;
                Img c = take(std::move(res));
                clear_outside(c, sub2);
                c->region = sub;
                c->cs = rcs;
                res = c;
            }
            results.emplace_back(p.result, res);
        }
        if (results.empty()) throw FilterError{RVK_OK};  // Error::NoResults: clear, no warning
        Img last = results.back().second;
        results.clear();
        return last;
    }
};

}  // namespace

namespace rvk {

// filter::apply + apply_to_canvas (mod.rs:335-355, 1100-1114) on a device layer.
rvk_status filter_apply(rvk_context *ctx, const rvk_filter *filter, const rvk_filter_data *data, Ts ts, uint8_t *px,
                        uint32_t w, uint32_t h, rvk_status *out_error, const FeImageHook *fe_image) {
    rvk_status ferr = RVK_OK;
    try {
        auto src = std::make_shared<DImg>();
        src->ctx = ctx;
        src->px = px;
        src->w = w; src->h = h;
        src->owned = false;
        Exec ex{ctx, data, ts, IRect{}, src, {}, fe_image};
        Img res = ex.run(*filter);
        ex.source.reset();
        res = into_cs(std::move(res), RVK_CI_SRGB);
        if (res->px == px) {
            // the result is the layer itself (SourceGraphic passed through unchanged)
        } else if (res->w == w && res->h == h) {
            // SourceOver onto a cleared pixmap of the same size is a copy
            RVK_TRY(launch_copy(ctx, px, res->px, res->bytes()));
        } else {
            RVK_TRY(launch_fill_sub(ctx, px, w, h, 0, 0, w, h, 0u));
            RVK_TRY(launch_draw_pixmap(ctx, px, w, h, 0, 0, res->px, res->w, res->h, 1.0f, RVK_BLEND_SOURCE_OVER));
        }
        ctx->stats.filters++;
        if (out_error) *out_error = RVK_OK;
        return RVK_OK;
    } catch (FilterError e) {
        ferr = e.code;
    }
    if (ferr == RVK_ERR_HIP || ferr == RVK_ERR_OUT_OF_MEMORY || ferr == RVK_ERR_INVALID_ARGUMENT ||
        ferr == RVK_ERR_UNSUPPORTED) {
        if (out_error) *out_error = ferr;
        if (ferr == RVK_ERR_HIP || ferr == RVK_ERR_INVALID_ARGUMENT) return ferr;
    }
    // clear on error (mod.rs:343-346)
    RVK_TRY(launch_fill_sub(ctx, px, w, h, 0, 0, w, h, 0u));
    if (out_error) *out_error = ferr;
    return ferr == RVK_ERR_OUT_OF_MEMORY ? ferr : RVK_OK;
}

}  // namespace rvk

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------
extern "C" {

#define CHECK_PX(ctx, px, w, h)                                            \
    do {                                                                   \
        if (!(ctx) || !(px) || !rvk::valid_dims((w), (h))) return RVK_ERR_INVALID_ARGUMENT; \
    } while (0)

rvk_status rvk_into_linear_rgb(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_into_cs(ctx, px, w, h, 1);
}
rvk_status rvk_into_srgb(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_into_cs(ctx, px, w, h, 0);
}
rvk_status rvk_demultiply_alpha(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_demultiply(ctx, px, w, h);
}
rvk_status rvk_multiply_alpha(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_multiply(ctx, px, w, h);
}
rvk_status rvk_box_blur(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_box_blur(ctx, px, w, h, sx, sy);
}
rvk_status rvk_box_blur_noalloc(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_box_blur_noalloc(ctx, px, w, h, sx, sy);
}
rvk_status rvk_iir_blur(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_iir_blur(ctx, px, w, h, sx, sy);
}
rvk_status rvk_morphology(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, uint32_t op, float rx, float ry) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_morphology(ctx, px, w, h, op, rx, ry);
}
rvk_status rvk_convolve_matrix(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p,
                               const float *values) {
    CHECK_PX(ctx, px, w, h);
    if (!p) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_convolve(ctx, px, w, h, p, values);
}
rvk_status rvk_color_matrix(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p) {
    CHECK_PX(ctx, px, w, h);
    if (!p) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_color_matrix(ctx, px, w, h, p, 0);
}
rvk_status rvk_component_transfer(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *p,
                                  const float *values) {
    CHECK_PX(ctx, px, w, h);
    if (!p) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_component_transfer(ctx, px, w, h, p, values, 0);
}
rvk_status rvk_composite_arithmetic(rvk_context *ctx, const uint8_t *s1, const uint8_t *s2, uint8_t *dst, uint32_t w,
                                    uint32_t h, float k1, float k2, float k3, float k4) {
    CHECK_PX(ctx, dst, w, h);
    if (!s1 || !s2) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_arithmetic(ctx, s1, s2, dst, (size_t)w * h, k1, k2, k3, k4);
}
rvk_status rvk_turbulence(rvk_context *ctx, uint8_t *dst, uint32_t w, uint32_t h, double ox, double oy, double sx,
                          double sy, double bfx, double bfy, uint32_t octaves, int32_t seed, uint32_t stitch,
                          uint32_t fractal) {
    CHECK_PX(ctx, dst, w, h);
    return rvk::launch_turbulence(ctx, dst, w, h, ox, oy, sx, sy, bfx, bfy, octaves, seed, stitch, fractal, 0);
}
rvk_status rvk_lighting(rvk_context *ctx, const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                        uint32_t specular, float surface_scale, float constant, float spec_exp,
                        const uint8_t color[4], const rvk_light_source *light) {
    CHECK_PX(ctx, dst, w, h);
    if (!src || !color || !light) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_lighting(ctx, src, dst, w, h, specular, surface_scale, constant, spec_exp, color, light);
}
rvk_status rvk_displacement_map(rvk_context *ctx, const uint8_t *src, const uint8_t *map, uint8_t *dst, uint32_t w,
                                uint32_t h, float sx, float sy, float scale, uint32_t xch, uint32_t ych) {
    CHECK_PX(ctx, dst, w, h);
    if (!src || !map) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_displacement(ctx, src, map, dst, w, h, sx, sy, scale, xch, ych);
}
rvk_status rvk_draw_pixmap(rvk_context *ctx, uint8_t *dst, uint32_t dw, uint32_t dh, int32_t x, int32_t y,
                           const uint8_t *src, uint32_t sw, uint32_t sh, float opacity, uint32_t mode) {
    CHECK_PX(ctx, dst, dw, dh);
    if (!src || !rvk::valid_dims(sw, sh)) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_draw_pixmap(ctx, dst, dw, dh, x, y, src, sw, sh, opacity, mode);
}
rvk_status rvk_fill(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const uint8_t rgba[4]) {
    CHECK_PX(ctx, px, w, h);
    if (!rgba) return RVK_ERR_INVALID_ARGUMENT;
    uint8_t pm[4];
    premultiply_color(rgba, pm);
    return rvk::launch_fill(ctx, px, w, h, pm);
}
// tiny_skia::Mask operations (clip.rs:25-27, mask.rs:32-45)
static bool mask_aligned(const void *mask, const void *px) {
    return ((uintptr_t)mask % 4) == 0 && ((uintptr_t)px % 16) == 0;
}
rvk_status rvk_mask_from_pixmap(rvk_context *ctx, const uint8_t *px, uint32_t w, uint32_t h, uint32_t type,
                                uint8_t *mask) {
    CHECK_PX(ctx, px, w, h);
    if (!mask || type > RVK_MASK_TYPE_LUMINANCE || !mask_aligned(mask, px)) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_mask_from_pixmap(ctx, px, mask, (size_t)w * h, type);
}
rvk_status rvk_mask_invert(rvk_context *ctx, uint8_t *mask, uint32_t w, uint32_t h) {
    CHECK_PX(ctx, mask, w, h);
    if (!mask_aligned(mask, nullptr)) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_mask_invert(ctx, mask, (size_t)w * h);
}
rvk_status rvk_apply_mask(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, const uint8_t *mask) {
    CHECK_PX(ctx, px, w, h);
    if (!mask || !mask_aligned(mask, px)) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::launch_apply_mask_u8(ctx, px, mask, (size_t)w * h);
}
// Pixmap::fill_rect with BlendMode::Clear on a pixel-aligned rect: zero the rect's pixels
rvk_status rvk_clear_rect(rvk_context *ctx, uint8_t *px, uint32_t w, uint32_t h, int32_t x, int32_t y, uint32_t rw,
                          uint32_t rh) {
    CHECK_PX(ctx, px, w, h);
    return rvk::launch_fill_sub(ctx, px, w, h, x, y, (int64_t)x + rw, (int64_t)y + rh, 0u);
}
rvk_status rvk_filter_apply(rvk_context *ctx, const rvk_filter *filter, const rvk_filter_data *data,
                            rvk_transform ts, uint8_t *px, uint32_t w, uint32_t h, rvk_status *out_error) {
    CHECK_PX(ctx, px, w, h);
    if (!filter || !data || (filter->primitive_count && !data->primitives)) return RVK_ERR_INVALID_ARGUMENT;
    return rvk::filter_apply(ctx, filter, data, ts_from(ts), px, w, h, out_error);
}

}  // extern "C"
