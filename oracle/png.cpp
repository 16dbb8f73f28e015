// Oracle (test infrastructure only): the scanlines of tiny-skia's Pixmap::encode_png
// (crates/resvg/src/main.rs:104-117 writes the CLI's output with it).
//
// Demultiply: PremultipliedColorU8::demultiply, restated by the reference's test
// harness at crates/resvg/tests/integration/main.rs:204-211: alpha 255 unchanged,
// otherwise c / (a / 255) + 0.5 in f64, cast to u8 the way Rust's `as u8` casts
// (saturating, NaN to 0).  Filters: ISO/IEC 15948 section 9 (types None, Sub, Up,
// Average, Paeth over 4-byte pixels; bytes left of the row and above the first row
// are 0), one per row, the type whose bytes read as signed have the least sum of
// magnitudes (ties to the lower type).  Output: per row the type byte, then the row.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "oracle.h"

namespace {

uint8_t demul(uint32_t c, uint32_t a) {
    if (a == 255) return (uint8_t)c;
    const double v = (double)c / ((double)a / 255.0) + 0.5;
    if (std::isnan(v)) return 0;
    if (v >= 255.0) return 255;
    if (v <= 0.0) return 0;
    return (uint8_t)v;
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

}  // namespace

extern "C" void orc_png_filter_rows(const uint8_t *px, uint32_t w, uint32_t h, uint8_t *out) {
    const size_t stride = 4 * (size_t)w;
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    std::vector<uint8_t> f[5];
    for (auto &v : f) v.resize(stride);
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t *row = px + (size_t)y * stride;
        for (size_t i = 0; i < stride; i += 4) {
            const uint32_t a = row[i + 3];
            for (int k = 0; k < 3; k++) cur[i + k] = demul(row[i + k], a);
            cur[i + 3] = (uint8_t)a;
        }
        uint64_t best_sum = ~0ull;
        int best = 0;
        for (int t = 0; t < 5; t++) {
            uint64_t sum = 0;
            for (size_t i = 0; i < stride; i++) {
                const int x = cur[i], a = i >= 4 ? cur[i - 4] : 0, b = prev[i], c = i >= 4 ? prev[i - 4] : 0;
                int pred = 0;
                if (t == 1) pred = a;
                else if (t == 2) pred = b;
                else if (t == 3) pred = (a + b) / 2;
                else if (t == 4) pred = paeth(a, b, c);
                const uint8_t v = (uint8_t)(x - pred);
                f[t][i] = v;
                sum += (uint64_t)std::abs((int)(int8_t)v);
            }
            if (sum < best_sum) {
                best_sum = sum;
                best = t;
            }
        }
        uint8_t *o = out + (size_t)y * (stride + 1);
        o[0] = (uint8_t)best;
        for (size_t i = 0; i < stride; i++) o[1 + i] = f[best][i];
        prev.swap(cur);
    }
}
