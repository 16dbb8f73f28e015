// Oracle restatement of resvg's filter-graph evaluation (filter::apply,
// crates/resvg/src/filter/mod.rs:335-564, 581-1147).  Test infrastructure only.
#include "oracle.h"
#include "common.h"
#include "geom.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

using namespace orc;

namespace {

struct Img {
    std::vector<uint8_t> px;
    uint32_t w = 0, h = 0;
    IRect region{};
    uint32_t cs = RVK_CI_SRGB;
};
using ImgP = std::shared_ptr<const Img>;

struct Result { int32_t name; ImgP img; };

struct FilterError { int32_t code; };

ImgP make_img(uint32_t w, uint32_t h, uint32_t cs) {  // Pixmap::try_create + Image::from_image
    if (w == 0 || h == 0) throw FilterError{RVK_ERR_INVALID_REGION};
    auto im = std::make_shared<Img>();
    im->px.assign((size_t)w * h * 4, 0);
    im->w = w; im->h = h;
    im->region = IRect{0, 0, (int32_t)w, (int32_t)h};
    im->cs = cs;
    return im;
}

std::shared_ptr<Img> clone_img(const ImgP &a) { return std::make_shared<Img>(*a); }

// Image::into_color_space (mod.rs:286-308)
ImgP into_cs(const ImgP &im, uint32_t cs) {
    if (im->cs == cs) return im;
    auto c = clone_img(im);
    if (cs == RVK_CI_SRGB) orc_into_srgb(c->px.data(), c->w, c->h);
    else orc_into_linear_rgb(c->px.data(), c->w, c->h);
    c->cs = cs;
    return c;
}

std::shared_ptr<Img> from_pixels(std::shared_ptr<Img> im, uint32_t cs) {  // Image::from_image
    im->region = IRect{0, 0, (int32_t)im->w, (int32_t)im->h};
    im->cs = cs;
    return im;
}

// scale_coordinates / resolve_std_dev (mod.rs:1119-1147)
void scale_coords(float x, float y, Ts ts, float &ox, float &oy) {
    float sx, sy;
    ts_get_scale(ts, sx, sy);
    ox = x * sx;
    oy = y * sy;
}

bool resolve_std_dev(float sdx, float sdy, Ts ts, double &ox, double &oy, bool &box) {
    float x, y;
    scale_coords(sdx, sdy, ts, x, y);
    if (approx_eq_ulps(x, 0.0f, 4) && approx_eq_ulps(y, 0.0f, 4)) return false;
    if (x < 0.05f) x = 0.0f;
    if (y < 0.05f) y = 0.0f;
    box = x >= 2.0f || y >= 2.0f;
    ox = x; oy = y;
    return true;
}

void blur_in_place(Img &im, double sx, double sy, bool box) {
    if (box) orc_box_blur(im.px.data(), im.w, im.h, sx, sy);
    else orc_iir_blur(im.px.data(), im.w, im.h, sx, sy);
}

struct Ctx {
    const rvk_filter_data *data;
    Ts ts;
    IRect region;
    const Img *source;
    std::vector<Result> results;
};

// get_input (mod.rs:523-564)
ImgP get_input(Ctx &c, int32_t input) {
    if (input >= 0) {
        for (auto it = c.results.rbegin(); it != c.results.rend(); ++it)
            if (it->name == input) return it->img;
        input = RVK_INPUT_SOURCE_GRAPHIC;
    }
    auto im = std::make_shared<Img>(*c.source);
    im->region = c.region;
    im->cs = RVK_CI_SRGB;
    if (input == RVK_INPUT_SOURCE_ALPHA)
        for (size_t i = 0; i < (size_t)im->w * im->h; i++) im->px[4 * i] = im->px[4 * i + 1] = im->px[4 * i + 2] = 0;
    return im;
}

void draw(Img &dst, const Img &src, int32_t x, int32_t y, uint32_t mode) {
    orc_draw_pixmap(dst.px.data(), dst.w, dst.h, x, y, src.px.data(), src.w, src.h, 1.0f, mode);
}

// transform_light_source (mod.rs:1063-1098)
rvk_light_source transform_light(rvk_light_source l, IRect region, Ts ts) {
    const float SQRT_2 = 1.41421353816986083984375f;
    if (l.kind == RVK_LIGHT_POINT) {
        float x = l.x, y = l.y;
        ts_map(ts, x, y);
        l.x = x - (float)region.x;
        l.y = y - (float)region.y;
        l.z = l.z * std::sqrt(ts.sx * ts.sx + ts.sy * ts.sy) / SQRT_2;
    } else if (l.kind == RVK_LIGHT_SPOT) {
        float sz = std::sqrt(ts.sx * ts.sx + ts.sy * ts.sy) / SQRT_2;
        float x = l.x, y = l.y;
        ts_map(ts, x, y);
        l.x = x - (float)region.x;
        l.y = y - (float)region.y;
        l.z *= sz;
        x = l.points_at_x; y = l.points_at_y;
        ts_map(ts, x, y);
        l.points_at_x = x - (float)region.x;
        l.points_at_y = y - (float)region.y;
        l.points_at_z *= sz;
    }
    return l;
}

ImgP apply_primitive(Ctx &c, const rvk_primitive &p, IRect subregion) {
    prof::Scope timed((int)p.kind);
    const uint32_t cs = p.color_interpolation;
    const IRect region = c.region;
    switch (p.kind) {
        case RVK_FE_BLEND: {  // mod.rs:695-729
            auto i1 = into_cs(get_input(c, p.in1), cs);
            auto i2 = into_cs(get_input(c, p.in2), cs);
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, cs));
            draw(*out, *i2, 0, 0, RVK_BLEND_SOURCE_OVER);
            draw(*out, *i1, 0, 0, p.u.blend.mode);
            return out;
        }
        case RVK_FE_DROP_SHADOW: {  // mod.rs:581-643
            const auto &d = p.u.drop_shadow;
            auto in = get_input(c, p.in1);
            float dx, dy;
            scale_coords(d.dx, d.dy, c.ts, dx, dy);
            auto out = std::const_pointer_cast<Img>(make_img(in->w, in->h, cs));
            auto input = into_cs(in, cs);
            auto shadow = clone_img(input);
            double sx, sy; bool box;
            if (resolve_std_dev(d.std_dev_x, d.std_dev_y, c.ts, sx, sy, box)) blur_in_place(*shadow, sx, sy, box);
            uint8_t op8 = as_u8(std::min(std::max(d.opacity, 0.0f), 1.0f) * 255.0f + 0.5f);
            float cr = (float)d.color[0] / 255.0f, cg = (float)d.color[1] / 255.0f,
                  cb = (float)d.color[2] / 255.0f, ca = (float)op8 / 255.0f;
            for (size_t i = 0; i < (size_t)shadow->w * shadow->h; i++) {
                uint8_t *q = &shadow->px[4 * i];
                float o = (float)q[3] / 255.0f;
                float a = f32_bound(0.0f, ca * f32_bound(0.0f, o, 1.0f), 1.0f);
                float r = cr, g = cg, b = cb;
                if (a != 1.0f) { r = r * a; g = g * a; b = b * a; }
                q[0] = as_u8(r * 255.0f + 0.5f); q[1] = as_u8(g * 255.0f + 0.5f);
                q[2] = as_u8(b * 255.0f + 0.5f); q[3] = as_u8(a * 255.0f + 0.5f);
            }
            if (cs == RVK_CI_SRGB) orc_into_srgb(shadow->px.data(), shadow->w, shadow->h);
            else orc_into_linear_rgb(shadow->px.data(), shadow->w, shadow->h);
            draw(*out, *shadow, as_i32(dx), as_i32(dy), RVK_BLEND_SOURCE_OVER);
            draw(*out, *input, 0, 0, RVK_BLEND_SOURCE_OVER);
            return out;
        }
        case RVK_FE_FLOOD: {  // mod.rs:820-835
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, RVK_CI_SRGB));
            uint8_t rgba[4] = {p.u.flood.color[0], p.u.flood.color[1], p.u.flood.color[2],
                               as_u8(std::min(std::max(p.u.flood.opacity, 0.0f), 1.0f) * 255.0f + 0.5f)};
            orc_fill(out->px.data(), out->w, out->h, rgba);
            return out;
        }
        case RVK_FE_GAUSSIAN_BLUR: {  // mod.rs:645-666
            auto in = get_input(c, p.in1);
            double sx, sy; bool box;
            if (!resolve_std_dev(p.u.blur.std_dev_x, p.u.blur.std_dev_y, c.ts, sx, sy, box)) return in;
            auto out = clone_img(into_cs(in, cs));
            blur_in_place(*out, sx, sy, box);
            return from_pixels(out, cs);
        }
        case RVK_FE_OFFSET: {  // mod.rs:668-693
            auto in = get_input(c, p.in1);
            float dx, dy;
            scale_coords(p.u.offset.dx, p.u.offset.dy, c.ts, dx, dy);
            if (approx_zero_ulps(dx, 4) && approx_zero_ulps(dy, 4)) return in;
            auto out = std::const_pointer_cast<Img>(make_img(in->w, in->h, in->cs));
            draw(*out, *in, as_i32(dx), as_i32(dy), RVK_BLEND_SOURCE_OVER);
            return out;
        }
        case RVK_FE_COMPOSITE: {  // mod.rs:731-793
            auto i1 = into_cs(get_input(c, p.in1), cs);
            auto i2 = into_cs(get_input(c, p.in2), cs);
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, cs));
            const auto &k = p.u.composite;
            if (k.op == RVK_COMPOSITE_ARITHMETIC) {
                size_t n = std::min({(size_t)i1->w * i1->h, (size_t)i2->w * i2->h, (size_t)out->w * out->h});
                // composite::arithmetic zips the three buffers pixel by pixel
                std::vector<uint8_t> t(out->px);
                orc_composite_arithmetic(i1->px.data(), i2->px.data(), t.data(), (uint32_t)n, 1, k.k1, k.k2,
                                         k.k3, k.k4);
                out->px = t;
                return out;
            }
            draw(*out, *i2, 0, 0, RVK_BLEND_SOURCE_OVER);
            uint32_t mode = k.op == RVK_COMPOSITE_IN ? RVK_BLEND_SOURCE_IN
                          : k.op == RVK_COMPOSITE_OUT ? RVK_BLEND_SOURCE_OUT
                          : k.op == RVK_COMPOSITE_ATOP ? RVK_BLEND_SOURCE_ATOP
                          : k.op == RVK_COMPOSITE_XOR ? RVK_BLEND_XOR : RVK_BLEND_SOURCE_OVER;
            draw(*out, *i1, 0, 0, mode);
            return out;
        }
        case RVK_FE_MERGE: {  // mod.rs:795-818
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, cs));
            for (uint32_t i = 0; i < p.u.merge.input_count; i++) {
                auto in = into_cs(get_input(c, c.data->inputs[p.u.merge.input_offset + i]), cs);
                draw(*out, *in, 0, 0, RVK_BLEND_SOURCE_OVER);
            }
            return out;
        }
        case RVK_FE_COMPONENT_TRANSFER: {  // mod.rs:891-903
            auto out = clone_img(into_cs(get_input(c, p.in1), cs));
            orc_demultiply_alpha(out->px.data(), out->w, out->h);
            orc_component_transfer(out->px.data(), out->w, out->h, &p, c.data->values);
            orc_multiply_alpha(out->px.data(), out->w, out->h);
            return from_pixels(out, cs);
        }
        case RVK_FE_COLOR_MATRIX: {  // mod.rs:905-917
            auto out = clone_img(into_cs(get_input(c, p.in1), cs));
            orc_demultiply_alpha(out->px.data(), out->w, out->h);
            orc_color_matrix(out->px.data(), out->w, out->h, &p);
            orc_multiply_alpha(out->px.data(), out->w, out->h);
            return from_pixels(out, cs);
        }
        case RVK_FE_CONVOLVE_MATRIX: {  // mod.rs:919-933
            auto out = clone_img(into_cs(get_input(c, p.in1), cs));
            if (p.u.convolve.preserve_alpha) orc_demultiply_alpha(out->px.data(), out->w, out->h);
            orc_convolve_matrix(out->px.data(), out->w, out->h, &p, c.data->values);
            return from_pixels(out, cs);
        }
        case RVK_FE_MORPHOLOGY: {  // mod.rs:935-956
            auto out = clone_img(into_cs(get_input(c, p.in1), cs));
            float rx, ry;
            scale_coords(p.u.morphology.radius_x, p.u.morphology.radius_y, c.ts, rx, ry);
            if (!(rx > 0.0f && ry > 0.0f)) std::fill(out->px.begin(), out->px.end(), 0);
            else orc_morphology(out->px.data(), out->w, out->h, p.u.morphology.op, rx, ry);
            return from_pixels(out, cs);
        }
        case RVK_FE_DISPLACEMENT_MAP: {  // mod.rs:958-986
            auto p1 = into_cs(get_input(c, p.in1), cs);
            auto p2 = into_cs(get_input(c, p.in2), cs);
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, cs));
            float sx, sy;
            scale_coords(p.u.displacement.scale, p.u.displacement.scale, c.ts, sx, sy);
            if (p1->w == out->w && p1->h == out->h && p2->w == out->w && p2->h == out->h)
                orc_displacement_map(p1->px.data(), p2->px.data(), out->px.data(), out->w, out->h, sx, sy,
                                     p.u.displacement.scale, p.u.displacement.x_channel, p.u.displacement.y_channel);
            return out;
        }
        case RVK_FE_TURBULENCE: {  // mod.rs:988-1018
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, cs));
            float sx, sy;
            ts_get_scale(c.ts, sx, sy);
            if (approx_zero_ulps(sx, 4) || approx_zero_ulps(sy, 4)) return out;
            const auto &t = p.u.turbulence;
            orc_turbulence(out->px.data(), out->w, out->h, (double)region.x - (double)c.ts.tx,
                           (double)region.y - (double)c.ts.ty, sx, sy, t.base_frequency_x, t.base_frequency_y,
                           t.num_octaves, t.seed, t.stitch_tiles, t.fractal_noise);
            orc_multiply_alpha(out->px.data(), out->w, out->h);
            return out;
        }
        case RVK_FE_DIFFUSE_LIGHTING:
        case RVK_FE_SPECULAR_LIGHTING: {  // mod.rs:1020-1060
            auto in = get_input(c, p.in1);
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, cs));
            rvk_light_source l = transform_light(p.u.lighting.light, region, c.ts);
            if (in->w == out->w && in->h == out->h)
                orc_lighting(in->px.data(), out->px.data(), out->w, out->h, p.kind == RVK_FE_SPECULAR_LIGHTING,
                             p.u.lighting.surface_scale, p.u.lighting.constant, p.u.lighting.specular_exponent,
                             p.u.lighting.color, &l);
            return out;
        }
        case RVK_FE_IMAGE: {  // apply_image (mod.rs:861-889)
            // render_nodes of the content into a region-sized pixmap with transform
            // [sx 0 0 sy subregion.x subregion.y] and max_bbox the pixmap
            if (!orc::fe_image_hook) throw FilterError{RVK_ERR_UNSUPPORTED};
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, RVK_CI_SRGB));
            float sx, sy;
            ts_get_scale(c.ts, sx, sy);
            const float t6[6] = {sx, 0.0f, 0.0f, sy, (float)subregion.x, (float)subregion.y};
            if (!orc::fe_image_hook(p.u.image.root, t6, out->px.data(), out->w, out->h))
                throw FilterError{RVK_ERR_INVALID_ARGUMENT};
            return from_pixels(out, RVK_CI_SRGB);
        }
        case RVK_FE_TILE: {  // apply_tile (mod.rs:837-859)
            auto in = get_input(c, p.in1);
            // subregion = input.region translated by -region.xy; copy_region is
            // tiny-skia's clone_rect, which cuts the rect to the pixmap
            const int64_t sx = (int64_t)in->region.x - region.x, sy = (int64_t)in->region.y - region.y;
            const int64_t l = std::max<int64_t>(sx, 0), t = std::max<int64_t>(sy, 0);
            const int64_t r = std::min<int64_t>(sx + in->region.w, in->w);
            const int64_t b = std::min<int64_t>(sy + in->region.h, in->h);
            if (r <= l || b <= t) throw FilterError{RVK_ERR_INVALID_REGION};
            const int64_t tw = r - l, th = b - t;
            std::vector<uint8_t> tile((size_t)tw * th * 4);
            for (int64_t y = 0; y < th; y++)
                std::memcpy(&tile[(size_t)(y * tw) * 4], &in->px[((size_t)(t + y) * in->w + l) * 4], (size_t)tw * 4);
            // fill_rect of the whole region with Pattern(tile, Repeat, translate(sx, sy)):
            // the pixel centre in tile space, wrapped by the repeat tiling
            // (v - floor(v / n) n), then sampled nearest (an integer translation
            // reduces every filter quality to nearest)
            auto out = std::const_pointer_cast<Img>(make_img(region.w, region.h, RVK_CI_SRGB));
            auto wrap = [](float v, float n) { return v - std::floor(v / n) * n; };
            for (uint32_t y = 0; y < out->h; y++) {
                const float fy = wrap((float)y + 0.5f - (float)sy, (float)th);
                const int64_t iy = std::min<int64_t>(std::max<int64_t>((int64_t)fy, 0), th - 1);
                for (uint32_t x = 0; x < out->w; x++) {
                    const float fx = wrap((float)x + 0.5f - (float)sx, (float)tw);
                    const int64_t ix = std::min<int64_t>(std::max<int64_t>((int64_t)fx, 0), tw - 1);
                    std::memcpy(&out->px[((size_t)y * out->w + x) * 4], &tile[(size_t)(iy * tw + ix) * 4], 4);
                }
            }
            return out;
        }
        default:
            throw FilterError{RVK_ERR_UNSUPPORTED};
    }
}

// Clears the pixels of `im` outside `sub` with four Clear fill_rects (mod.rs:471-501).
void clear_outside(Img &im, IRect sub) {
    const int64_t w = im.w, h = im.h;
    auto clear_rect = [&](int64_t x0, int64_t y0, int64_t x1, int64_t y1) {
        x0 = std::max<int64_t>(x0, 0); y0 = std::max<int64_t>(y0, 0);
        x1 = std::min<int64_t>(x1, w); y1 = std::min<int64_t>(y1, h);
        for (int64_t y = y0; y < y1; y++)
            for (int64_t x = x0; x < x1; x++) std::memset(&im.px[4 * (size_t)(y * w + x)], 0, 4);
    };
    if (sub.y > 0) clear_rect(0, 0, w, sub.y);
    if (sub.x > 0) clear_rect(0, 0, sub.x, h);
    clear_rect(ir_right(sub), 0, (int64_t)ir_right(sub) + w, h);
    clear_rect(0, ir_bottom(sub), w, (int64_t)ir_bottom(sub) + h);
}

ImgP apply_inner(const rvk_filter &f, const rvk_filter_data &data, Ts ts, const Img &source) {  // mod.rs:357-521
    FRect fr = rect_transform(f.rect, ts);
    if (!fr.ok) throw FilterError{RVK_ERR_INVALID_REGION};
    IRect region;
    if (!fit_to_rect(frect_to_int(fr), IRect{0, 0, (int32_t)source.w, (int32_t)source.h}, region))
        throw FilterError{RVK_ERR_INVALID_REGION};
    Ctx c{&data, ts, region, &source, {}};
    for (uint32_t i = 0; i < f.primitive_count; i++) {
        const rvk_primitive &p = data.primitives[f.primitive_offset + i];
        FRect pr = rect_transform(p.rect, ts);
        if (!pr.ok) throw FilterError{RVK_ERR_INVALID_REGION};
        IRect sub = frect_to_int(pr);
        if (p.kind == RVK_FE_OFFSET && p.in1 >= 0) {
            for (auto it = c.results.rbegin(); it != c.results.rend(); ++it)
                if (it->name == p.in1) { sub = it->img->region; break; }
        }
        ImgP res = apply_primitive(c, p, sub);
        if (!ir_eq(region, sub)) {
            IRect sub2 = p.kind == RVK_FE_OFFSET ? IRect{0, 0, region.w, region.h}
                                                 : IRect{sub.x - region.x, sub.y - region.y, sub.w, sub.h};
            auto cl = clone_img(res);
            clear_outside(*cl, sub2);
            cl->region = sub;
            res = cl;
        }
        c.results.push_back(Result{p.result, res});
    }
    if (c.results.empty()) throw FilterError{-1};
    return c.results.back().img;
}

}  // namespace

// filter::apply (mod.rs:335-355) + apply_to_canvas (mod.rs:1100-1114)
int32_t orc_filter_apply(const rvk_filter *filter, const rvk_filter_data *data, rvk_transform ts, uint8_t *px,
                         uint32_t w, uint32_t h) {
    orc::prof::Scope timed(orc::prof::kFilter);
    Img source;
    source.px.assign(px, px + (size_t)w * h * 4);
    source.w = w; source.h = h;
    int32_t code = RVK_OK;
    try {
        ImgP res = apply_inner(*filter, *data, ts_from(ts), source);
        res = into_cs(res, RVK_CI_SRGB);
        std::memset(px, 0, (size_t)w * h * 4);
        orc_draw_pixmap(px, w, h, 0, 0, res->px.data(), res->w, res->h, 1.0f, RVK_BLEND_SOURCE_OVER);
        return RVK_OK;
    } catch (FilterError e) {
        code = e.code;
    }
    std::memset(px, 0, (size_t)w * h * 4);
    return code < 0 ? RVK_OK : code;
}
