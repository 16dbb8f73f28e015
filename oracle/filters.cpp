// Oracle restatement of resvg's filter primitives (crates/resvg/src/filter/*.rs)
// and of the tiny-skia f32 ("highp") pixel pipeline used by draw_pixmap/fill.
// Test infrastructure only: never linked into the product library.
#include "oracle.h"
#include "common.h"
#include <vector>

using namespace orc;

// ---------------------------------------------------------------------------
// Colour space (filter/mod.rs:115-238)
// ---------------------------------------------------------------------------
// Literal golden data: SRGB_TO_LINEAR_RGB_TABLE (filter/mod.rs:162-179).
static const uint8_t kSrgbToLinear[256] = {
    0,   0,   0,   0,   0,   0,  0,    1,   1,   1,   1,   1,   1,   1,   1,   1,
    1,   1,   2,   2,   2,   2,  2,    2,   2,   2,   3,   3,   3,   3,   3,   3,
    4,   4,   4,   4,   4,   5,  5,    5,   5,   6,   6,   6,   6,   7,   7,   7,
    8,   8,   8,   8,   9,   9,  9,   10,  10,  10,  11,  11,  12,  12,  12,  13,
    13,  13,  14,  14,  15,  15,  16,  16,  17,  17,  17,  18,  18,  19,  19,  20,
    20,  21,  22,  22,  23,  23,  24,  24,  25,  25,  26,  27,  27,  28,  29,  29,
    30,  30,  31,  32,  32,  33,  34,  35,  35,  36,  37,  37,  38,  39,  40,  41,
    41,  42,  43,  44,  45,  45,  46,  47,  48,  49,  50,  51,  51,  52,  53,  54,
    55,  56,  57,  58,  59,  60,  61,  62,  63,  64,  65,  66,  67,  68,  69,  70,
    71,  72,  73,  74,  76,  77,  78,  79,  80,  81,  82,  84,  85,  86,  87,  88,
    90,  91,  92,  93,  95,  96,  97,  99, 100, 101, 103, 104, 105, 107, 108, 109,
    111, 112, 114, 115, 116, 118, 119, 121, 122, 124, 125, 127, 128, 130, 131, 133,
    134, 136, 138, 139, 141, 142, 144, 146, 147, 149, 151, 152, 154, 156, 157, 159,
    161, 163, 164, 166, 168, 170, 171, 173, 175, 177, 179, 181, 183, 184, 186, 188,
    190, 192, 194, 196, 198, 200, 202, 204, 206, 208, 210, 212, 214, 216, 218, 220,
    222, 224, 226, 229, 231, 233, 235, 237, 239, 242, 244, 246, 248, 250, 253, 255,
};
// Literal golden data: LINEAR_RGB_TO_SRGB_TABLE (filter/mod.rs:195-212).
static const uint8_t kLinearToSrgb[256] = {
    0,  13,  22,  28,  34,  38,  42,  46,  50,  53,  56,  59,  61,  64,  66,  69,
    71,  73,  75,  77,  79,  81,  83,  85,  86,  88,  90,  92,  93,  95,  96,  98,
    99, 101, 102, 104, 105, 106, 108, 109, 110, 112, 113, 114, 115, 117, 118, 119,
    120, 121, 122, 124, 125, 126, 127, 128, 129, 130, 131, 132, 133, 134, 135, 136,
    137, 138, 139, 140, 141, 142, 143, 144, 145, 146, 147, 148, 148, 149, 150, 151,
    152, 153, 154, 155, 155, 156, 157, 158, 159, 159, 160, 161, 162, 163, 163, 164,
    165, 166, 167, 167, 168, 169, 170, 170, 171, 172, 173, 173, 174, 175, 175, 176,
    177, 178, 178, 179, 180, 180, 181, 182, 182, 183, 184, 185, 185, 186, 187, 187,
    188, 189, 189, 190, 190, 191, 192, 192, 193, 194, 194, 195, 196, 196, 197, 197,
    198, 199, 199, 200, 200, 201, 202, 202, 203, 203, 204, 205, 205, 206, 206, 207,
    208, 208, 209, 209, 210, 210, 211, 212, 212, 213, 213, 214, 214, 215, 215, 216,
    216, 217, 218, 218, 219, 219, 220, 220, 221, 221, 222, 222, 223, 223, 224, 224,
    225, 226, 226, 227, 227, 228, 228, 229, 229, 230, 230, 231, 231, 232, 232, 233,
    233, 234, 234, 235, 235, 236, 236, 237, 237, 238, 238, 238, 239, 239, 240, 240,
    241, 241, 242, 242, 243, 243, 244, 244, 245, 245, 246, 246, 246, 247, 247, 248,
    248, 249, 249, 250, 250, 251, 251, 251, 252, 252, 253, 253, 254, 254, 255, 255,
};

const uint8_t *orc_srgb_to_linear_table(void) { return kSrgbToLinear; }
const uint8_t *orc_linear_to_srgb_table(void) { return kLinearToSrgb; }

// multiply_alpha (filter/mod.rs:129-136)
void orc_multiply_alpha(uint8_t *px, uint32_t w, uint32_t h) {
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        uint8_t *p = px + 4 * i;
        float a = (float)p[3] / 255.0f;
        for (int c = 2; c >= 0; c--) p[c] = as_u8((float)p[c] * a + 0.5f);
    }
}

// demultiply_alpha (filter/mod.rs:139-146)
void orc_demultiply_alpha(uint8_t *px, uint32_t w, uint32_t h) {
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        uint8_t *p = px + 4 * i;
        float a = (float)p[3] / 255.0f;
        for (int c = 2; c >= 0; c--) p[c] = as_u8((float)p[c] / a + 0.5f);
    }
}

static void lut_rgb(uint8_t *px, uint32_t w, uint32_t h, const uint8_t *lut) {
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) px[4 * i + c] = lut[px[4 * i + c]];
}

// PixmapExt::into_linear_rgb (filter/mod.rs:121-125)
void orc_into_linear_rgb(uint8_t *px, uint32_t w, uint32_t h) {
    orc_demultiply_alpha(px, w, h);
    lut_rgb(px, w, h, kSrgbToLinear);
    orc_multiply_alpha(px, w, h);
}

// PixmapExt::into_srgb (filter/mod.rs:115-119)
void orc_into_srgb(uint8_t *px, uint32_t w, uint32_t h) {
    orc_demultiply_alpha(px, w, h);
    lut_rgb(px, w, h, kLinearToSrgb);
    orc_multiply_alpha(px, w, h);
}

// ---------------------------------------------------------------------------
// Box blur (filter/box_blur.rs)
// ---------------------------------------------------------------------------
// create_box_gauss (box_blur.rs:37-71)
void orc_box_gauss_sizes(float sigma, int32_t out[5]) {
    if (sigma > 0.0f) {
        const float n = 5.0f;
        float w_ideal = std::sqrt(12.0f * sigma * sigma / n) + 1.0f;
        int32_t wl = as_i32(std::floor(w_ideal));
        if (wl % 2 == 0) wl -= 1;
        int32_t wu = wl + 2;
        float wlf = (float)wl;
        float m_ideal = (12.0f * sigma * sigma - n * wlf * wlf - 4.0f * n * wlf - 3.0f * n) /
                        (-4.0f * wlf - 4.0f);
        size_t m = as_usize(std::round(m_ideal));  // f32::round: half away from zero
        for (size_t i = 0; i < 5; i++) out[i] = i < m ? wl : wu;
    } else {
        for (int i = 0; i < 5; i++) out[i] = 1;
    }
}

// box.rs `round`: magic-number round to nearest even (box_blur.rs:327-331)
static inline float magic_round(float x) {
    volatile float t = x + 12582912.0f;
    return t - 12582912.0f;
}

// One running-sum pass along lines (box_blur_vert / box_blur_horz,
// box_blur.rs:85-320).  `stride` = pixel step along a line, `lstride` = step
// between lines, `n` = line length, `lines` = number of lines.  Pixels outside
// the line are transparent (fv = lv = RGBA8::default()).
static void box_pass(const uint8_t *src, uint8_t *dst, size_t n, size_t lines, size_t stride,
                     size_t lstride, size_t r) {
    const float iarr = 1.0f / (float)(r + r + 1);
    for (size_t line = 0; line < lines; line++) {
        const uint8_t *s = src + 4 * line * lstride;
        uint8_t *d = dst + 4 * line * lstride;
        long long acc[4] = {0, 0, 0, 0};
        auto px = [&](size_t i) { return s + 4 * i * stride; };
        auto emit = [&](size_t ti) {
            uint8_t *o = d + 4 * ti * stride;
            for (int c = 0; c < 4; c++) o[c] = as_u8(magic_round((float)acc[c] * iarr));
        };
        for (size_t j = 0; j < std::min(r, n); j++)
            for (int c = 0; c < 4; c++) acc[c] += px(j)[c];
        size_t ti = 0, li = 0, ri = r;
        for (size_t k = 0; k < std::min(n, r + 1); k++) {
            if (ri < n) for (int c = 0; c < 4; c++) acc[c] += px(ri)[c];
            ri++;
            emit(ti++);
        }
        if (n <= r) continue;
        for (size_t k = r + 1; k < n - r; k++) {
            for (int c = 0; c < 4; c++) acc[c] += (long long)px(ri)[c] - (long long)px(li)[c];
            ri++; li++;
            emit(ti++);
        }
        for (size_t k = 0; k < std::min(n - r - 1, r); k++) {
            for (int c = 0; c < 4; c++) acc[c] -= px(li)[c];
            li++;
            emit(ti++);
        }
    }
}

// box_blur::apply (box_blur.rs:23-34): five rounds of [vertical, horizontal].
void orc_box_blur(uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    int32_t bh[5], bv[5];
    orc_box_gauss_sizes((float)sx, bh);
    orc_box_gauss_sizes((float)sy, bv);
    size_t bytes = (size_t)w * h * 4;
    std::vector<uint8_t> back(px, px + bytes);
    for (int i = 0; i < 5; i++) {
        size_t rh = (size_t)((bh[i] - 1) / 2), rv = (size_t)((bv[i] - 1) / 2);
        if (rv == 0) std::memcpy(back.data(), px, bytes);
        else box_pass(px, back.data(), h, w, w, 1, rv);
        if (rh == 0) std::memcpy(px, back.data(), bytes);
        else box_pass(back.data(), px, w, h, 1, w, rh);
    }
}

// ---------------------------------------------------------------------------
// IIR blur (filter/iir_blur.rs)
// ---------------------------------------------------------------------------
static void iir_coefficients(double sigma, double *lambda, double *dnu) {  // iir_blur.rs:142-146
    double l = (sigma * sigma) / (2.0 * 4.0);
    *lambda = l;
    *dnu = (1.0 + 2.0 * l - std::sqrt(1.0 + 4.0 * l)) / (2.0 * l);
}

static void iir_2d(double *buf, size_t w, size_t h, double sx, double sy) {  // iir_blur.rs:79-140
    double lx = 1.0, dx = 1.0, ly = 1.0, dy = 1.0;
    if (sx > 0.0) {
        iir_coefficients(sx, &lx, &dx);
        for (size_t y = 0; y < h; y++) {
            double *row = buf + w * y;
            for (int s = 0; s < 4; s++) {
                for (size_t x = 1; x < w; x++) row[x] += dx * row[x - 1];
                for (size_t x = w - 1; x > 0; x--) row[x - 1] += dx * row[x];
            }
        }
    }
    if (sy > 0.0) {
        iir_coefficients(sy, &ly, &dy);
        size_t len = w * h;
        for (size_t x = 0; x < w; x++) {
            for (int s = 0; s < 4; s++) {
                for (size_t y = w; y < len; y += w) buf[x + y] += dy * buf[x + y - w];
                for (size_t y = len - w; y > 0; y -= w) buf[x + y - w] += dy * buf[x + y];
            }
        }
    }
    double base = std::sqrt(dx * dy) / std::sqrt(lx * ly);
    double p = base * base;  // powi(base, 8) by repeated squaring
    p = p * p;
    p = p * p;
    for (size_t i = 0; i < w * h; i++) buf[i] *= p;
}

// iir_blur::apply (iir_blur.rs:47-77)
void orc_iir_blur(uint8_t *px, uint32_t w, uint32_t h, double sx, double sy) {
    size_t n = (size_t)w * h;
    std::vector<double> buf(n);
    for (int ch = 0; ch < 4; ch++) {
        for (size_t i = 0; i < n; i++) buf[i] = (double)px[4 * i + ch] / 255.0;
        iir_2d(buf.data(), w, h, sx, sy);
        for (size_t i = 0; i < n; i++) px[4 * i + ch] = as_u8d(buf[i] * 255.0);
    }
}

// ---------------------------------------------------------------------------
// Morphology (filter/morphology.rs:15-73)
// ---------------------------------------------------------------------------
void orc_morphology(uint8_t *px, uint32_t w, uint32_t h, uint32_t op, float rx, float ry) {
    uint32_t columns = std::min(as_u32(std::ceil(rx)) * 2u, w);
    uint32_t rows = std::min(as_u32(std::ceil(ry)) * 2u, h);
    uint32_t tx0 = as_u32(std::floor((float)columns / 2.0f));
    uint32_t ty0 = as_u32(std::floor((float)rows / 2.0f));
    int32_t wmax = (int32_t)w - 1, hmax = (int32_t)h - 1;
    std::vector<uint8_t> out((size_t)w * h * 4);
    const bool erode = op == RVK_MORPH_ERODE;
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            uint8_t v[4];
            for (int c = 0; c < 4; c++) v[c] = erode ? 255 : 0;
            for (uint32_t oy = 0; oy < rows; oy++) {
                for (uint32_t ox = 0; ox < columns; ox++) {
                    int32_t tx = (int32_t)x - (int32_t)tx0 + (int32_t)ox;
                    int32_t ty = (int32_t)y - (int32_t)ty0 + (int32_t)oy;
                    if (tx < 0 || tx > wmax || ty < 0 || ty > hmax) continue;
                    const uint8_t *p = px + 4 * ((size_t)ty * w + tx);
                    for (int c = 0; c < 4; c++) v[c] = erode ? std::min(v[c], p[c]) : std::max(v[c], p[c]);
                }
            }
            std::memcpy(&out[4 * ((size_t)y * w + x)], v, 4);
        }
    }
    std::memcpy(px, out.data(), out.size());
}

// ---------------------------------------------------------------------------
// ConvolveMatrix (filter/convolve_matrix.rs:15-111)
// ---------------------------------------------------------------------------
void orc_convolve_matrix(uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *prim, const float *values) {
    const auto &cm = prim->u.convolve;
    const float *k = values + cm.data_offset;
    int32_t wmax = (int32_t)w - 1, hmax = (int32_t)h - 1;
    std::vector<uint8_t> out((size_t)w * h * 4);
    const bool pa = cm.preserve_alpha != 0;
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            float nr = 0.0f, ng = 0.0f, nb = 0.0f, na = 0.0f;
            for (uint32_t oy = 0; oy < cm.rows; oy++) {
                for (uint32_t ox = 0; ox < cm.columns; ox++) {
                    int32_t tx = (int32_t)x - (int32_t)cm.target_x + (int32_t)ox;
                    int32_t ty = (int32_t)y - (int32_t)cm.target_y + (int32_t)oy;
                    if (cm.edge_mode == RVK_EDGE_NONE) {
                        if (tx < 0 || tx > wmax || ty < 0 || ty > hmax) continue;
                    } else if (cm.edge_mode == RVK_EDGE_DUPLICATE) {
                        tx = std::max(0, std::min(wmax, tx));
                        ty = std::max(0, std::min(hmax, ty));
                    } else {
                        while (tx < 0) tx += (int32_t)w;
                        tx %= (int32_t)w;
                        while (ty < 0) ty += (int32_t)h;
                        ty %= (int32_t)h;
                    }
                    float kv = k[(cm.rows - oy - 1) * cm.columns + (cm.columns - ox - 1)];
                    const uint8_t *p = px + 4 * ((size_t)ty * w + tx);
                    nr += (float)p[0] / 255.0f * kv;
                    ng += (float)p[1] / 255.0f * kv;
                    nb += (float)p[2] / 255.0f * kv;
                    if (!pa) na += (float)p[3] / 255.0f * kv;
                }
            }
            if (pa) na = (float)px[4 * ((size_t)y * w + x) + 3] / 255.0f;
            else na = na / cm.divisor + cm.bias;
            float bna = f32_bound(0.0f, na, 1.0f);
            auto calc = [&](float v) {
                v = v / cm.divisor + cm.bias * na;
                v = pa ? f32_bound(0.0f, v, 1.0f) * bna : f32_bound(0.0f, v, bna);
                return as_u8(v * 255.0f + 0.5f);
            };
            uint8_t *o = &out[4 * ((size_t)y * w + x)];
            o[0] = calc(nr); o[1] = calc(ng); o[2] = calc(nb);
            o[3] = as_u8(bna * 255.0f + 0.5f);
        }
    }
    std::memcpy(px, out.data(), out.size());
}

// ---------------------------------------------------------------------------
// ColorMatrix (filter/color_matrix.rs:11-110), unpremultiplied input
// ---------------------------------------------------------------------------
static inline uint8_t from_normalized(float c) { return as_u8(f32_bound(0.0f, c, 1.0f) * 255.0f); }

void orc_color_matrix(uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *prim) {
    const auto &cm = prim->u.color_matrix;
    size_t n = (size_t)w * h;
    float m[20];
    std::memcpy(m, cm.values, sizeof m);
    if (cm.kind == RVK_COLOR_MATRIX_SATURATE) {
        float v = std::max(cm.values[0], 0.0f);
        float s[9] = {0.213f + 0.787f * v, 0.715f - 0.715f * v, 0.072f - 0.072f * v,
                      0.213f - 0.213f * v, 0.715f + 0.285f * v, 0.072f - 0.072f * v,
                      0.213f - 0.213f * v, 0.715f - 0.715f * v, 0.072f + 0.928f * v};
        std::memcpy(m, s, sizeof s);
    } else if (cm.kind == RVK_COLOR_MATRIX_HUE_ROTATE) {
        float ang = to_radians(cm.values[0]);
        float a1 = std::cos(ang), a2 = std::sin(ang);
        float s[9] = {0.213f + 0.787f * a1 - 0.213f * a2, 0.715f - 0.715f * a1 - 0.715f * a2,
                      0.072f - 0.072f * a1 + 0.928f * a2, 0.213f - 0.213f * a1 + 0.143f * a2,
                      0.715f + 0.285f * a1 + 0.140f * a2, 0.072f - 0.072f * a1 - 0.283f * a2,
                      0.213f - 0.213f * a1 - 0.787f * a2, 0.715f - 0.715f * a1 + 0.715f * a2,
                      0.072f + 0.928f * a1 + 0.072f * a2};
        std::memcpy(m, s, sizeof s);
    }
    for (size_t i = 0; i < n; i++) {
        uint8_t *p = px + 4 * i;
        float r = (float)p[0] / 255.0f, g = (float)p[1] / 255.0f, b = (float)p[2] / 255.0f,
              a = (float)p[3] / 255.0f;
        switch (cm.kind) {
            case RVK_COLOR_MATRIX_MATRIX: {
                float nr = r * m[0] + g * m[1] + b * m[2] + a * m[3] + m[4];
                float ng = r * m[5] + g * m[6] + b * m[7] + a * m[8] + m[9];
                float nb = r * m[10] + g * m[11] + b * m[12] + a * m[13] + m[14];
                float na = r * m[15] + g * m[16] + b * m[17] + a * m[18] + m[19];
                p[0] = from_normalized(nr); p[1] = from_normalized(ng);
                p[2] = from_normalized(nb); p[3] = from_normalized(na);
                break;
            }
            case RVK_COLOR_MATRIX_SATURATE:
            case RVK_COLOR_MATRIX_HUE_ROTATE: {
                float nr = r * m[0] + g * m[1] + b * m[2];
                float ng = r * m[3] + g * m[4] + b * m[5];
                float nb = r * m[6] + g * m[7] + b * m[8];
                p[0] = from_normalized(nr); p[1] = from_normalized(ng); p[2] = from_normalized(nb);
                break;
            }
            default: {
                float na = r * 0.2125f + g * 0.7154f + b * 0.0721f;
                p[0] = p[1] = p[2] = 0;
                p[3] = from_normalized(na);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// ComponentTransfer (filter/component_transfer.rs:10-72), unpremultiplied input
// ---------------------------------------------------------------------------
static bool transfer_is_dummy(const rvk_transfer_fn &f) {
    switch (f.kind) {
        case RVK_TRANSFER_IDENTITY: return true;
        case RVK_TRANSFER_TABLE:
        case RVK_TRANSFER_DISCRETE: return f.table_len == 0;
        default: return false;
    }
}

static uint8_t transfer(const rvk_transfer_fn &f, const float *values, uint8_t cu) {
    float c = (float)cu / 255.0f;
    const float *v = values + f.table_offset;
    switch (f.kind) {
        case RVK_TRANSFER_TABLE: {
            size_t n = f.table_len - 1;
            size_t k = as_usize(std::floor(c * (float)n));
            k = std::min(k, n);
            if (k == n) c = v[k];
            else {
                float vk = v[k], vk1 = v[k + 1], kf = (float)k, nf = (float)n;
                c = vk + (c - kf / nf) * nf * (vk1 - vk);
            }
            break;
        }
        case RVK_TRANSFER_DISCRETE: {
            size_t n = f.table_len;
            size_t k = as_usize(std::floor(c * (float)n));
            c = v[std::min(k, n - 1)];
            break;
        }
        case RVK_TRANSFER_LINEAR: c = f.a * c + f.b; break;
        case RVK_TRANSFER_GAMMA: c = f.a * std::pow(c, f.b) + f.c; break;
        default: break;
    }
    return as_u8(f32_bound(0.0f, c, 1.0f) * 255.0f);
}

void orc_component_transfer(uint8_t *px, uint32_t w, uint32_t h, const rvk_primitive *prim, const float *values) {
    const auto &ct = prim->u.component_transfer;
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        uint8_t *p = px + 4 * i;
        // channel order R, B, G, A as in component_transfer.rs:12-26 (independent)
        const int order[4] = {0, 2, 1, 3};
        for (int j = 0; j < 4; j++) {
            int ch = order[j];
            if (!transfer_is_dummy(ct.func[ch])) p[ch] = transfer(ct.func[ch], values, p[ch]);
        }
    }
}

// ---------------------------------------------------------------------------
// Arithmetic composite (filter/composite.rs:14-50)
// ---------------------------------------------------------------------------
void orc_composite_arithmetic(const uint8_t *s1, const uint8_t *s2, uint8_t *dst, uint32_t w, uint32_t h,
                              float k1, float k2, float k3, float k4) {
    auto calc = [&](uint8_t c1, uint8_t c2, float mx) {
        float i1 = (float)c1 / 255.0f, i2 = (float)c2 / 255.0f;
        float r = k1 * i1 * i2 + k2 * i1 + k3 * i2 + k4;
        return f32_bound(0.0f, r, mx);
    };
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        const uint8_t *a = s1 + 4 * i, *b = s2 + 4 * i;
        float al = calc(a[3], b[3], 1.0f);
        if (approx_zero_ulps(al, 4)) continue;
        uint8_t *o = dst + 4 * i;
        o[0] = as_u8(calc(a[0], b[0], al) * 255.0f);
        o[1] = as_u8(calc(a[1], b[1], al) * 255.0f);
        o[2] = as_u8(calc(a[2], b[2], al) * 255.0f);
        o[3] = as_u8(al * 255.0f);
    }
}

// ---------------------------------------------------------------------------
// Turbulence (filter/turbulence.rs)
// ---------------------------------------------------------------------------
static const int32_t RAND_M = 2147483647, RAND_A = 16807, RAND_Q = 127773, RAND_R = 2836;
static const int32_t BSIZE = 0x100, BM = 0xff, PERLIN_N = 0x1000;
static const size_t BLEN = 0x100 + 0x100 + 2;

int32_t orc_turbulence_random(int32_t seed) {  // turbulence.rs:287-294
    int32_t r = RAND_A * (seed % RAND_Q) - RAND_R * (seed / RAND_Q);
    if (r <= 0) r += RAND_M;
    return r;
}

struct TurbTables { int32_t lattice[BLEN]; double grad[4][BLEN][2]; };

static void turb_init(int32_t seed, TurbTables &t) {  // turbulence.rs:93-140
    if (seed <= 0) seed = -(seed % (RAND_M - 1)) + 1;
    if (seed > RAND_M - 1) seed = RAND_M - 1;
    for (int k = 0; k < 4; k++) {
        for (int i = 0; i < BSIZE; i++) {
            t.lattice[i] = i;
            for (int j = 0; j < 2; j++) {
                seed = orc_turbulence_random(seed);
                t.grad[k][i][j] = (double)((seed % (BSIZE + BSIZE)) - BSIZE) / (double)BSIZE;
            }
            double s = std::sqrt(t.grad[k][i][0] * t.grad[k][i][0] + t.grad[k][i][1] * t.grad[k][i][1]);
            t.grad[k][i][0] /= s;
            t.grad[k][i][1] /= s;
        }
    }
    for (int i = BSIZE - 1; i >= 1; i--) {
        int32_t k = t.lattice[i];
        seed = orc_turbulence_random(seed);
        int32_t j = seed % BSIZE;
        t.lattice[i] = t.lattice[j];
        t.lattice[j] = k;
    }
    for (int i = 0; i < BSIZE + 2; i++) {
        t.lattice[BSIZE + i] = t.lattice[i];
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < 2; j++) t.grad[k][BSIZE + i][j] = t.grad[k][i][j];
    }
}

void orc_turbulence_init(int32_t seed, int32_t *lattice_out, double *gradient_out) {
    TurbTables t;
    turb_init(seed, t);
    std::memcpy(lattice_out, t.lattice, sizeof t.lattice);
    std::memcpy(gradient_out, t.grad, sizeof t.grad);
}

struct Stitch { int32_t width, height, wrap_x, wrap_y; };

static inline double s_curve(double t) { return t * t * (3.0 - 2.0 * t); }
static inline double lerp(double t, double a, double b) { return a + t * (b - a); }

static double noise2(int ch, double x, double y, const TurbTables &t, const Stitch *st) {  // turbulence.rs:224-285
    double tt = x + (double)PERLIN_N;
    int32_t bx0 = as_i32d(tt);
    int32_t bx1 = bx0 + 1;
    double rx0 = tt - (double)as_i64d(tt);
    double rx1 = rx0 - 1.0;
    tt = y + (double)PERLIN_N;
    int32_t by0 = as_i32d(tt);
    int32_t by1 = by0 + 1;
    double ry0 = tt - (double)as_i64d(tt);
    double ry1 = ry0 - 1.0;
    if (st) {
        if (bx0 >= st->wrap_x) bx0 -= st->width;
        if (bx1 >= st->wrap_x) bx1 -= st->width;
        if (by0 >= st->wrap_y) by0 -= st->height;
        if (by1 >= st->wrap_y) by1 -= st->height;
    }
    bx0 &= BM; bx1 &= BM; by0 &= BM; by1 &= BM;
    int32_t i = t.lattice[bx0], j = t.lattice[bx1];
    int32_t b00 = t.lattice[i + by0], b10 = t.lattice[j + by0];
    int32_t b01 = t.lattice[i + by1], b11 = t.lattice[j + by1];
    double sx = s_curve(rx0), sy = s_curve(ry0);
    const double *q = t.grad[ch][b00];
    double u = rx0 * q[0] + ry0 * q[1];
    q = t.grad[ch][b10];
    double v = rx1 * q[0] + ry0 * q[1];
    double a = lerp(sx, u, v);
    q = t.grad[ch][b01];
    u = rx0 * q[0] + ry1 * q[1];
    q = t.grad[ch][b11];
    v = rx1 * q[0] + ry1 * q[1];
    double b = lerp(sx, u, v);
    return lerp(sy, a, b);
}

static double turbulence_at(int ch, double x, double y, double tile_x, double tile_y, double tw, double th,
                            double bfx, double bfy, uint32_t octaves, bool fractal, bool stitch,
                            const TurbTables &t) {  // turbulence.rs:142-222
    Stitch st{}, *stp = nullptr;
    if (stitch) {
        if (!approx_eq_ulps_d(bfx, 0.0, 4)) {
            double lo = std::floor(tw * bfx) / tw, hi = std::ceil(tw * bfx) / tw;
            bfx = (bfx / lo < hi / bfx) ? lo : hi;
        }
        if (!approx_eq_ulps_d(bfy, 0.0, 4)) {
            double lo = std::floor(th * bfy) / th, hi = std::ceil(th * bfy) / th;
            bfy = (bfy / lo < hi / bfy) ? lo : hi;
        }
        st.width = as_i32d(tw * bfx + 0.5);
        st.height = as_i32d(th * bfy + 0.5);
        st.wrap_x = as_i32d(tile_x * bfx + (double)PERLIN_N + (double)st.width);
        st.wrap_y = as_i32d(tile_y * bfy + (double)PERLIN_N + (double)st.height);
        stp = &st;
    }
    double sum = 0.0, ratio = 1.0;
    x *= bfx;
    y *= bfy;
    for (uint32_t o = 0; o < octaves; o++) {
        double nv = noise2(ch, x, y, t, stp);
        sum += (fractal ? nv : std::fabs(nv)) / ratio;
        x *= 2.0;
        y *= 2.0;
        ratio *= 2.0;
        if (stp) {
            st.width *= 2;
            st.wrap_x = 2 * st.wrap_x - PERLIN_N;
            st.height *= 2;
            st.wrap_y = 2 * st.wrap_y - PERLIN_N;
        }
    }
    return sum;
}

// turbulence::apply (turbulence.rs:33-91): unpremultiplied output.
void orc_turbulence(uint8_t *dst, uint32_t w, uint32_t h, double ox, double oy, double sx, double sy,
                    double bfx, double bfy, uint32_t octaves, int32_t seed, uint32_t stitch, uint32_t fractal) {
    TurbTables t;
    turb_init(seed, t);
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            uint8_t *p = dst + 4 * ((size_t)y * w + x);
            for (int ch = 0; ch < 4; ch++) {
                double tx = ((double)x + ox) / sx, ty = ((double)y + oy) / sy;
                double nv = turbulence_at(ch, tx, ty, (double)x, (double)y, (double)w, (double)h, bfx, bfy,
                                          octaves, fractal != 0, stitch != 0, t);
                nv = fractal ? (nv * 255.0 + 255.0) / 2.0 : nv * 255.0;
                p[ch] = as_u8(f32_bound(0.0f, (float)nv, 255.0f) + 0.5f);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Lighting (filter/lighting.rs)
// ---------------------------------------------------------------------------
struct V3 { float x, y, z; };
static inline float v3_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float v3_len(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline V3 v3_norm_or_self(V3 a) {  // Vector3::normalized().unwrap_or(v) (lighting.rs:68-79)
    float l = v3_len(a);
    if (!approx_zero_ulps(l, 4)) return V3{a.x / l, a.y / l, a.z / l};
    return a;
}

struct Normal { float fx, fy, nx, ny; };

static inline int16_t alpha_at(const uint8_t *src, uint32_t w, uint32_t x, uint32_t y) {
    return (int16_t)src[4 * ((size_t)w * y + x) + 3];
}

// The nine Sobel-normal cases (lighting.rs:340-476).
static Normal sobel_normal(const uint8_t *s, uint32_t w, uint32_t h, uint32_t x, uint32_t y) {
    const float F12 = 1.0f / 2.0f, F13 = 1.0f / 3.0f, F14 = 1.0f / 4.0f, F23 = 2.0f / 3.0f;
    auto A = [&](uint32_t xx, uint32_t yy) { return (int32_t)alpha_at(s, w, xx, yy); };
    auto mk = [](float fx, float fy, int32_t nx, int32_t ny) {
        return Normal{fx, fy, (float)(int16_t)(-nx), (float)(int16_t)(-ny)};
    };
    uint32_t r = w - 1, b = h - 1;
    if (y == 0) {
        if (x == 0) {
            int32_t c = A(0, 0), rr = A(1, 0), bo = A(0, 1), br = A(1, 1);
            return mk(F23, F23, -2 * c + 2 * rr - bo + br, -2 * c - rr + 2 * bo + br);
        }
        if (x == r) {
            int32_t l = A(r - 1, 0), c = A(r, 0), bl = A(r - 1, 1), bo = A(r, 1);
            return mk(F23, F23, -2 * l + 2 * c - bl + bo, -l - 2 * c + bl + 2 * bo);
        }
        int32_t l = A(x - 1, 0), c = A(x, 0), rr = A(x + 1, 0), bl = A(x - 1, 1), bo = A(x, 1), br = A(x + 1, 1);
        return mk(F13, F12, -2 * l + 2 * rr - bl + br, -l - 2 * c - rr + bl + 2 * bo + br);
    }
    if (y == b) {
        if (x == 0) {
            int32_t t = A(0, b - 1), tr = A(1, b - 1), c = A(0, b), rr = A(1, b);
            return mk(F23, F23, -t + tr - 2 * c + 2 * rr, -2 * t - tr + 2 * c + rr);
        }
        if (x == r) {
            int32_t tl = A(r - 1, b - 1), t = A(r, b - 1), l = A(r - 1, b), c = A(r, b);
            return mk(F23, F23, -tl + t - 2 * l + 2 * c, -tl - 2 * t + l + 2 * c);
        }
        int32_t tl = A(x - 1, b - 1), t = A(x, b - 1), tr = A(x + 1, b - 1), l = A(x - 1, b), c = A(x, b),
                rr = A(x + 1, b);
        return mk(F13, F12, -tl + tr - 2 * l + 2 * rr, -tl - 2 * t - tr + l + 2 * c + rr);
    }
    if (x == 0) {
        int32_t t = A(0, y - 1), tr = A(1, y - 1), c = A(0, y), rr = A(1, y), bo = A(0, y + 1), br = A(1, y + 1);
        return mk(F12, F13, -t + tr - 2 * c + 2 * rr - bo + br, -2 * t - tr + 2 * bo + br);
    }
    if (x == r) {
        int32_t tl = A(r - 1, y - 1), t = A(r, y - 1), l = A(r - 1, y), c = A(r, y), bl = A(r - 1, y + 1),
                bo = A(r, y + 1);
        return mk(F12, F13, -tl + t - 2 * l + 2 * c - bl + bo, -tl - 2 * t + bl + 2 * bo);
    }
    int32_t tl = A(x - 1, y - 1), t = A(x, y - 1), tr = A(x + 1, y - 1), l = A(x - 1, y), rr = A(x + 1, y),
            bl = A(x - 1, y + 1), bo = A(x, y + 1), br = A(x + 1, y + 1);
    return mk(F14, F14, -tl + tr - 2 * l + 2 * rr - bl + br, -tl - 2 * t - tr + bl + 2 * bo + br);
}

// lighting::apply / diffuse_lighting / specular_lighting / light_color (lighting.rs:132-338)
void orc_lighting(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, uint32_t specular,
                  float surface_scale, float constant, float spec_exp, const uint8_t color[4],
                  const rvk_light_source *light) {
    if (w < 3 || h < 3) return;
    V3 lv{1.0f, 1.0f, 1.0f};
    if (light->kind == RVK_LIGHT_DISTANT) {
        float az = to_radians(light->azimuth), el = to_radians(light->elevation);
        lv = V3{std::cos(az) * std::cos(el), std::sin(az) * std::cos(el), std::sin(el)};
    }
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            Normal nrm = sobel_normal(src, w, h, x, y);
            if (light->kind != RVK_LIGHT_DISTANT) {
                float nz = (float)alpha_at(src, w, x, y) / 255.0f * surface_scale;
                V3 v{light->x - (float)x, light->y - (float)y, light->z - nz};
                lv = v3_norm_or_self(v);
            }
            // light_color (lighting.rs:309-338)
            uint8_t lc[3] = {color[0], color[1], color[2]};
            if (light->kind == RVK_LIGHT_SPOT) {
                V3 dir{light->points_at_x - light->x, light->points_at_y - light->y, light->points_at_z - light->z};
                dir = v3_norm_or_self(dir);
                float mlds = -v3_dot(lv, dir);
                bool black = mlds <= 0.0f;
                if (!black && light->has_limiting_cone_angle &&
                    mlds < std::cos(to_radians(light->limiting_cone_angle)))
                    black = true;
                if (black) {
                    lc[0] = lc[1] = lc[2] = 0;
                } else {
                    float f = std::pow(mlds, light->specular_exponent);
                    for (int c = 0; c < 3; c++) lc[c] = as_u8(f32_bound(0.0f, (float)color[c] * f, 255.0f) + 0.5f);
                }
            }
            bool nzero = approx_zero_ulps(nrm.nx, 4) && approx_zero_ulps(nrm.ny, 4);
            float factor;
            if (!specular) {
                float k;
                if (nzero) k = lv.z;
                else {
                    float s = surface_scale / 255.0f;
                    float nx = nrm.nx * s, ny = nrm.ny * s;
                    nx *= nrm.fx; ny *= nrm.fy;
                    V3 n{nx, ny, 1.0f};
                    k = v3_dot(n, lv) / v3_len(n);
                }
                factor = constant * k;
            } else {
                V3 hv{lv.x + 0.0f, lv.y + 0.0f, lv.z + 1.0f};
                float hl = v3_len(hv);
                if (approx_zero_ulps(hl, 4)) factor = 0.0f;
                else {
                    float ndh;
                    if (nzero) ndh = hv.z / hl;
                    else {
                        float s = surface_scale / 255.0f;
                        float nx = nrm.nx * s, ny = nrm.ny * s;
                        nx *= nrm.fx; ny *= nrm.fy;
                        V3 n{nx, ny, 1.0f};
                        ndh = v3_dot(n, hv) / v3_len(n) / hl;
                    }
                    float k = approx_eq_ulps(spec_exp, 1.0f, 4) ? ndh : std::pow(ndh, spec_exp);
                    factor = constant * k;
                }
            }
            uint8_t *o = dst + 4 * ((size_t)y * w + x);
            for (int c = 0; c < 3; c++) o[c] = as_u8(f32_bound(0.0f, (float)lc[c] * factor, 255.0f) + 0.5f);
            o[3] = specular ? std::max(std::max(o[0], o[1]), o[2]) : 255;
        }
    }
}

// ---------------------------------------------------------------------------
// DisplacementMap (filter/displacement_map.rs:15-62)
// ---------------------------------------------------------------------------
void orc_displacement_map(const uint8_t *src, const uint8_t *map, uint8_t *dst, uint32_t w, uint32_t h,
                          float sx, float sy, float scale, uint32_t xch, uint32_t ych) {
    auto chan = [](const uint8_t *p, uint32_t ch) {
        return ch == RVK_CHANNEL_R ? p[0] : ch == RVK_CHANNEL_G ? p[1] : ch == RVK_CHANNEL_B ? p[2] : p[3];
    };
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            const uint8_t *m = map + 4 * ((size_t)y * w + x);
            float dx = (float)chan(m, xch) / 255.0f - 0.5f;
            float dy = (float)chan(m, ych) / 255.0f - 0.5f;
            int32_t ox = as_i32(std::round((float)x + dx * sx * scale));
            int32_t oy = as_i32(std::round((float)y + dy * sy * scale));
            if (ox >= 0 && ox < (int32_t)w && oy >= 0 && oy < (int32_t)h)
                std::memcpy(dst + 4 * ((size_t)y * w + x), src + 4 * ((size_t)oy * w + ox), 4);
        }
    }
}

// ---------------------------------------------------------------------------
// tiny-skia f32 raster pipeline for draw_pixmap (restated from tiny-skia 0.12's
// highp stages: gather -> scale_1_float -> load_dst -> blend -> store).
// ---------------------------------------------------------------------------
struct F4 { float r, g, b, a; };

static inline float from_u8f(uint8_t v) { return (float)v * (1.0f / 255.0f); }
static inline uint8_t unnorm(float v) {
    v = std::max(v, 0.0f);  // NaN -> 0
    if (!(v >= 0.0f)) v = 0.0f;
    v = std::min(v, 1.0f);
    return (uint8_t)std::nearbyint(v * 255.0f);
}
static inline float inv(float v) { return 1.0f - v; }
static inline float two(float v) { return v + v; }
static inline float lum(float r, float g, float b) { return r * 0.30f + (g * 0.59f + b * 0.11f); }
static inline float sat(float r, float g, float b) {
    return std::max(r, std::max(g, b)) - std::min(r, std::min(g, b));
}
static void set_sat(float &r, float &g, float &b, float s) {
    float mn = std::min(r, std::min(g, b)), mx = std::max(r, std::max(g, b)), st = mx - mn;
    auto sc = [&](float c) { return st == 0.0f ? 0.0f : (c - mn) * s / st; };
    r = sc(r); g = sc(g); b = sc(b);
}
static void set_lum(float &r, float &g, float &b, float l) {
    float d = l - lum(r, g, b);
    r += d; g += d; b += d;
}
// clip_color as tiny-skia evaluates it: only the upper clamp toward the luminance
// takes effect (no lower `min < 0` correction); pinned by the reference fixture
// painting/mix-blend-mode/hue, which differs by up to 4 LSB when the lower clamp runs.
static void clip_color(float &r, float &g, float &b, float a) {
    float mx = std::max(r, std::max(g, b)), l = lum(r, g, b);
    auto cl = [&](float c) {
        if (mx > a && mx - l != 0.0f) c = l + (c - l) * (a - l) / (mx - l);
        return std::max(c, 0.0f);
    };
    r = cl(r); g = cl(g); b = cl(b);
}

static F4 blend(uint32_t mode, F4 s, F4 d) {
    F4 o;
    auto all4 = [&](auto f) { o.r = f(s.r, d.r); o.g = f(s.g, d.g); o.b = f(s.b, d.b); o.a = f(s.a, d.a); };
    auto sep = [&](auto f) {
        o.r = f(s.r, d.r); o.g = f(s.g, d.g); o.b = f(s.b, d.b);
        o.a = d.a * inv(s.a) + s.a;
    };
    const float sa = s.a, da = d.a;
    switch (mode) {
        case RVK_BLEND_CLEAR: o = F4{0, 0, 0, 0}; break;
        case RVK_BLEND_SOURCE: o = s; break;
        case RVK_BLEND_DESTINATION: o = d; break;
        case RVK_BLEND_SOURCE_OVER: all4([&](float x, float y) { return y * inv(sa) + x; }); break;
        case RVK_BLEND_DESTINATION_OVER: all4([&](float x, float y) { return x * inv(da) + y; }); break;
        case RVK_BLEND_SOURCE_IN: all4([&](float x, float) { return x * da; }); break;
        case RVK_BLEND_DESTINATION_IN: all4([&](float, float y) { return y * sa; }); break;
        case RVK_BLEND_SOURCE_OUT: all4([&](float x, float) { return x * inv(da); }); break;
        case RVK_BLEND_DESTINATION_OUT: all4([&](float, float y) { return y * inv(sa); }); break;
        case RVK_BLEND_SOURCE_ATOP: all4([&](float x, float y) { return x * da + y * inv(sa); }); break;
        case RVK_BLEND_DESTINATION_ATOP: all4([&](float x, float y) { return y * sa + x * inv(da); }); break;
        case RVK_BLEND_XOR: all4([&](float x, float y) { return x * inv(da) + y * inv(sa); }); break;
        case RVK_BLEND_PLUS: all4([&](float x, float y) { return std::min(x + y, 1.0f); }); break;
        case RVK_BLEND_MODULATE: all4([&](float x, float y) { return x * y; }); break;
        case RVK_BLEND_SCREEN: all4([&](float x, float y) { return x + y - x * y; }); break;
        case RVK_BLEND_MULTIPLY:
            all4([&](float x, float y) { return x * inv(da) + y * inv(sa) + x * y; });
            break;
        case RVK_BLEND_DARKEN: sep([&](float x, float y) { return x + y - std::max(x * da, y * sa); }); break;
        case RVK_BLEND_LIGHTEN: sep([&](float x, float y) { return x + y - std::min(x * da, y * sa); }); break;
        case RVK_BLEND_DIFFERENCE:
            sep([&](float x, float y) { return x + y - two(std::min(x * da, y * sa)); });
            break;
        case RVK_BLEND_EXCLUSION: sep([&](float x, float y) { return x + y - two(x * y); }); break;
        case RVK_BLEND_COLOR_BURN:
            sep([&](float x, float y) {
                if (y == da) return y + x * inv(da);
                if (x == 0.0f) return y * inv(sa);
                return sa * (da - std::min(da, (da - y) * sa * (1.0f / x))) + x * inv(da) + y * inv(sa);
            });
            break;
        case RVK_BLEND_COLOR_DODGE:
            sep([&](float x, float y) {
                if (y == 0.0f) return x * inv(da);
                if (x == sa) return x + y * inv(sa);
                return sa * std::min(da, (y * sa) * (1.0f / (sa - x))) + x * inv(da) + y * inv(sa);
            });
            break;
        case RVK_BLEND_HARD_LIGHT:
            sep([&](float x, float y) {
                return x * inv(da) + y * inv(sa) +
                       (two(x) <= sa ? two(x * y) : sa * da - two((da - y) * (sa - x)));
            });
            break;
        case RVK_BLEND_OVERLAY:
            sep([&](float x, float y) {
                return x * inv(da) + y * inv(sa) +
                       (two(y) <= da ? two(x * y) : sa * da - two((da - y) * (sa - x)));
            });
            break;
        case RVK_BLEND_SOFT_LIGHT:
            sep([&](float x, float y) {
                float m = da > 0.0f ? y / da : 0.0f, s2 = two(x), m4 = two(two(m));
                float dark_src = y * (sa + (s2 - sa) * (1.0f - m));
                float dark_dst = (m4 * m4 + m4) * (m - 1.0f) + 7.0f * m;
                float lite_dst = std::sqrt(m) - m;
                float lite_src = y * sa + da * (s2 - sa) * (two(two(y)) <= da ? dark_dst : lite_dst);
                return x * inv(da) + y * inv(sa) + (s2 <= sa ? dark_src : lite_src);
            });
            break;
        case RVK_BLEND_HUE:
        case RVK_BLEND_SATURATION:
        case RVK_BLEND_COLOR:
        case RVK_BLEND_LUMINOSITY: {
            float R, G, B;
            if (mode == RVK_BLEND_HUE) {
                R = s.r * sa; G = s.g * sa; B = s.b * sa;
                set_sat(R, G, B, sat(d.r, d.g, d.b) * sa);
                set_lum(R, G, B, lum(d.r, d.g, d.b) * sa);
            } else if (mode == RVK_BLEND_SATURATION) {
                R = d.r * sa; G = d.g * sa; B = d.b * sa;
                set_sat(R, G, B, sat(s.r, s.g, s.b) * da);
                set_lum(R, G, B, lum(d.r, d.g, d.b) * sa);
            } else if (mode == RVK_BLEND_COLOR) {
                R = s.r * da; G = s.g * da; B = s.b * da;
                set_lum(R, G, B, lum(d.r, d.g, d.b) * sa);
            } else {
                R = d.r * sa; G = d.g * sa; B = d.b * sa;
                set_lum(R, G, B, lum(s.r, s.g, s.b) * da);
            }
            clip_color(R, G, B, sa * da);
            o.r = s.r * inv(da) + d.r * inv(sa) + R;
            o.g = s.g * inv(da) + d.g * inv(sa) + G;
            o.b = s.b * inv(da) + d.b * inv(sa) + B;
            o.a = sa + da - sa * da;
            break;
        }
        default: o = s;
    }
    return o;
}

void orc_draw_pixmap(uint8_t *dst, uint32_t dw, uint32_t dh, int32_t x, int32_t y, const uint8_t *src,
                     uint32_t sw, uint32_t sh, float opacity, uint32_t blend_mode) {
    int64_t x0 = std::max<int64_t>(x, 0), y0 = std::max<int64_t>(y, 0);
    int64_t x1 = std::min<int64_t>((int64_t)x + sw, dw), y1 = std::min<int64_t>((int64_t)y + sh, dh);
    for (int64_t py = y0; py < y1; py++) {
        for (int64_t px = x0; px < x1; px++) {
            const uint8_t *sp = src + 4 * ((size_t)(py - y) * sw + (size_t)(px - x));
            uint8_t *dp = dst + 4 * ((size_t)py * dw + (size_t)px);
            F4 s{from_u8f(sp[0]), from_u8f(sp[1]), from_u8f(sp[2]), from_u8f(sp[3])};
            if (opacity != 1.0f) { s.r *= opacity; s.g *= opacity; s.b *= opacity; s.a *= opacity; }
            F4 d{from_u8f(dp[0]), from_u8f(dp[1]), from_u8f(dp[2]), from_u8f(dp[3])};
            F4 o = blend(blend_mode, s, d);
            dp[0] = unnorm(o.r); dp[1] = unnorm(o.g); dp[2] = unnorm(o.b); dp[3] = unnorm(o.a);
        }
    }
}

// Pixmap::fill(Color::from_rgba8(..)): premultiply in f32, (v*255+0.5) as u8.
void orc_fill(uint8_t *px, uint32_t w, uint32_t h, const uint8_t rgba[4]) {
    float r = (float)rgba[0] / 255.0f, g = (float)rgba[1] / 255.0f, b = (float)rgba[2] / 255.0f,
          a = (float)rgba[3] / 255.0f;
    if (a != 1.0f) { r *= a; g *= a; b *= a; }
    uint8_t c[4] = {as_u8(r * 255.0f + 0.5f), as_u8(g * 255.0f + 0.5f), as_u8(b * 255.0f + 0.5f),
                    as_u8(a * 255.0f + 0.5f)};
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) std::memcpy(px + 4 * i, c, 4);
}
