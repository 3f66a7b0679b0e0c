/*
 * ORACLE — test infrastructure, not product code.
 *
 * FP64 restatements of crates/lightbaker/src/filter.rs's smoothing and fill over any number of
 * channels, for the adjoint's checker: the bounces of bake.rs:300-330 carried in FP64 with a
 * derivative by every parameter as extra channels (both passes are linear in the values, with
 * weights and masks fixed by the geometry). Surface::same (filter.rs:13-16) is the FP32 test, so
 * the masks are the FP32 bake's.
 *   orc_smooth64  filter.rs:20-53
 *   orc_fill64    filter.rs:58-94
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NONE_U 0xffffffffu

static int same(const uint32_t *chart, const float *normal, size_t i, size_t j) {
    const float *a = normal + i * 3, *b = normal + j * 3;
    return chart[i] == chart[j] && a[0] * b[0] + a[1] * b[1] + a[2] * b[2] >= 0.9f;
}

void orc_smooth64(int w, int h, const uint32_t *chart, const float *normal, const double *v,
                  int C, const uint8_t *valid, int r, double *out) {
    size_t n = (size_t)w * h;
    memcpy(out, v, n * C * sizeof(double));
    if (r == 0) return;
    double rr = (double)r * 0.75;
    double sigma2 = 2.0 * (rr * rr);
    double *sum = malloc((size_t)C * sizeof(double));
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            if (!valid[i]) continue;
            for (int c = 0; c < C; c++) sum[c] = 0.0;
            double wsum = 0.0;
            int y0 = y - r < 0 ? 0 : y - r, y1 = y + r + 1 < h ? y + r + 1 : h;
            int x0 = x - r < 0 ? 0 : x - r, x1 = x + r + 1 < w ? x + r + 1 : w;
            for (int yy = y0; yy < y1; yy++)
                for (int xx = x0; xx < x1; xx++) {
                    size_t j = (size_t)yy * w + xx;
                    if (!valid[j] || !same(chart, normal, i, j)) continue;
                    double dx = (double)(xx - x), dy = (double)(yy - y);
                    double k = exp(-(dx * dx + dy * dy) / sigma2);
                    for (int c = 0; c < C; c++) sum[c] += v[j * C + c] * k;
                    wsum += k;
                }
            for (int c = 0; c < C; c++) out[i * C + c] = sum[c] / wsum;
        }
    free(sum);
}

void orc_fill64(int w, int h, const uint32_t *chart, double *v, int C, uint8_t *valid, int rings) {
    size_t n = (size_t)w * h;
    uint8_t *was = malloc(n ? n : 1);
    double *src = malloc((n ? n : 1) * C * sizeof(double));
    double *own = malloc((size_t)C * sizeof(double)), *any = malloc((size_t)C * sizeof(double));
    for (int ring = 0; ring < rings; ring++) {
        memcpy(was, valid, n);
        memcpy(src, v, n * C * sizeof(double));
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t i = (size_t)y * w + x;
                if (was[i]) continue;
                int no = 0, na = 0;
                for (int c = 0; c < C; c++) own[c] = any[c] = 0.0;
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
                    for (int c = 0; c < C; c++) v[i * C + c] = own[c] / (double)no;
                    valid[i] = 1;
                } else if (na > 0 && chart[i] == NONE_U) {
                    for (int c = 0; c < C; c++) v[i * C + c] = any[c] / (double)na;
                    valid[i] = 1;
                }
            }
    }
    free(was);
    free(src);
    free(own);
    free(any);
}
