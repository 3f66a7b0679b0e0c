/*
 * ORACLE — test infrastructure, not product code.
 *
 * The reference's CPU texel passes, crates/lightbaker/src/filter.rs, restated in C over one
 * tile's arrays (w x h, row-major), FP32 in Rust's evaluation order (-ffp-contract=off):
 *   orc_smooth     filter.rs:20-53   Gaussian-weighted mean over valid same-surface texels
 *   orc_fill       filter.rs:58-94   ring-by-ring mean into texels without a value
 *   orc_sun_field  filter.rs:106-141 the sun's signed distance field from the hard-ray bits
 * Surface::same (filter.rs:13-16) is the chart equality and n . n' >= 0.9.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NONE_U 0xffffffffu

static int same(const uint32_t *chart, const float *normal, size_t i, size_t j) {
    const float *a = normal + i * 3, *b = normal + j * 3;
    return chart[i] == chart[j] && a[0] * b[0] + a[1] * b[1] + a[2] * b[2] >= 0.9f;
}

/* filter.rs:20-53; v and out hold C channels a texel */
void orc_smooth(int w, int h, const uint32_t *chart, const float *normal, const float *v, int C,
                const uint8_t *valid, int r, float *out) {
    size_t n = (size_t)w * h;
    memcpy(out, v, n * C * sizeof(float));
    if (r == 0) return;
    float rr = (float)r * 0.75f;
    float sigma2 = 2.0f * (rr * rr);
    float sum[16];
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            if (!valid[i]) continue;
            for (int c = 0; c < C; c++) sum[c] = 0.0f;
            float wsum = 0.0f;
            int y0 = y - r < 0 ? 0 : y - r, y1 = y + r + 1 < h ? y + r + 1 : h;
            int x0 = x - r < 0 ? 0 : x - r, x1 = x + r + 1 < w ? x + r + 1 : w;
            for (int yy = y0; yy < y1; yy++)
                for (int xx = x0; xx < x1; xx++) {
                    size_t j = (size_t)yy * w + xx;
                    if (!valid[j] || !same(chart, normal, i, j)) continue;
                    float dx = (float)xx - (float)x, dy = (float)yy - (float)y;
                    float k = expf(-(dx * dx + dy * dy) / sigma2);
                    for (int c = 0; c < C; c++) sum[c] += v[j * C + c] * k;
                    wsum += k;
                }
            for (int c = 0; c < C; c++) out[i * C + c] = sum[c] / wsum;
        }
}

/* filter.rs:58-94: `rings` rings, in place on v (C channels) and valid */
void orc_fill(int w, int h, const uint32_t *chart, float *v, int C, uint8_t *valid, int rings) {
    size_t n = (size_t)w * h;
    uint8_t *was = malloc(n ? n : 1);
    float *src = malloc((n ? n : 1) * C * sizeof(float));
    float own[16], any[16];
    for (int ring = 0; ring < rings; ring++) {
        memcpy(was, valid, n);
        memcpy(src, v, n * C * sizeof(float));
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t i = (size_t)y * w + x;
                if (was[i]) continue;
                int no = 0, na = 0;
                for (int c = 0; c < C; c++) own[c] = any[c] = 0.0f;
                int y0 = y > 0 ? y - 1 : 0, y1 = y + 2 < h ? y + 2 : h;
                int x0 = x > 0 ? x - 1 : 0, x1 = x + 2 < w ? x + 2 : w;
                for (int yy = y0; yy < y1; yy++)
                    for (int xx = x0; xx < x1; xx++) {
                        size_t j = (size_t)yy * w + xx;
                        if (!was[j]) continue;
                        if (chart[i] != NONE_U && chart[j] == chart[i]) {
                            for (int c = 0; c < C; c++) own[c] += src[j * C + c];
                            no++;
                        } else {
                            for (int c = 0; c < C; c++) any[c] += src[j * C + c];
                            na++;
                        }
                    }
                if (no > 0) {
                    for (int c = 0; c < C; c++) v[i * C + c] = own[c] / (float)no;
                    valid[i] = 1;
                } else if (na > 0 && chart[i] == NONE_U) {
                    for (int c = 0; c < C; c++) v[i * C + c] = any[c] / (float)na;
                    valid[i] = 1;
                }
            }
    }
    free(was);
    free(src);
}

/* filter.rs:106-141 */
void orc_sun_field(int w, int h, const uint32_t *chart, const float *normal, const uint32_t *hard,
                   const uint8_t *covered, int g, float radius, float *out) {
    int reach = (int)ceilf(radius) + 1;
    int gg = g * g;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            out[i] = -1.0f;
            if (!covered[i]) continue;
            int lit = __builtin_popcount(hard[i]) * 2 >= gg;
            float best = radius;
            int y0 = y - reach < 0 ? 0 : y - reach, y1 = y + reach + 1 < h ? y + reach + 1 : h;
            int x0 = x - reach < 0 ? 0 : x - reach, x1 = x + reach + 1 < w ? x + reach + 1 : w;
            for (int yy = y0; yy < y1; yy++)
                for (int xx = x0; xx < x1; xx++) {
                    size_t j = (size_t)yy * w + xx;
                    if (!covered[j] || !same(chart, normal, i, j)) continue;
                    for (int c = 0; c < gg; c++) {
                        int at = (hard[j] >> c) & 1u;
                        if (at == lit) continue;
                        float px = (float)xx + ((float)(c % g) + 0.5f) / (float)g - ((float)x + 0.5f);
                        float py = (float)yy + ((float)(c / g) + 0.5f) / (float)g - ((float)y + 0.5f);
                        float d = sqrtf(px * px + py * py) - 0.5f / (float)g;
                        if (d < best) best = d;
                    }
                }
            float d = best > 0.0f ? best : 0.0f;
            if (d > radius) d = radius;
            out[i] = 0.5f + 0.5f * (lit ? d : -d) / radius;
        }
}

/* bake.rs:238's `.powf(b.ao_contrast)` over n values: the host's powf (glibc's on Linux, as Rust's
 * f32::powf calls it), so the array and the scalar restatement round alike */
void orc_powf_n(const float *x, size_t n, float y, float *out) {
    for (size_t i = 0; i < n; i++) out[i] = powf(x[i], y);
}
