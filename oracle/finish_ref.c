/* ORACLE — test infrastructure, not product code.
 *
 * Scalar C restatement of the bake's tail after the rays and of the output codecs:
 *   orc_probe_fill       volume.rs:96-121   probes inside solids from their 6 neighbours, 8 rings
 *   orc_volume_irr       volume.rs:125-147  irradiance from the 8 probes round a point
 *   orc_volume_lit       bake.rs:337-356    texels with no light of their own lit by the volume
 *   orc_atlas            bake.rs:364-421    the atlas from every tile's TileLight, the smaller
 *                                           indirect atlas and its two rings of fill
 *   orc_unorm8           output.rs:101,167  f32 in [0, 1] to a byte
 *   orc_bc4_block/bc5    encode.rs:7-50
 *   orc_rgb9e5(_decode)  encode.rs:53-74
 *   orc_volume_bytes     output.rs:108,165-181
 * Each follows the Rust line by line in FP32 with no contraction (built -ffp-contract=off);
 * Rust's `as` casts from float saturate and send NaN to 0, and f32::max ignores a NaN operand,
 * as restated here.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define VALIDITY_NONE 0
#define VALIDITY_VOLUME 5

static float maxf_rs(float a, float b) { return isnan(a) ? b : (isnan(b) ? a : (a > b ? a : b)); }
static float clampf_rs(float v, float lo, float hi) {  /* f32::clamp: NaN stays NaN */
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
static uint32_t f2u_rs(float v) {  /* `as u32` */
    if (isnan(v) || v <= 0.0f) return 0;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

/* volume.rs:96-121; probes are 16 floats (4 rows of rgb+w), x fastest, then z, then y */
void orc_probe_fill(float *probes, size_t nx, size_t ny, size_t nz) {
    size_t np = nx * ny * nz;
    uint8_t *known = (uint8_t *)malloc(np), *was = (uint8_t *)malloc(np);
    float *src = (float *)malloc(np * 16 * sizeof(float));
    for (size_t k = 0; k < np; k++) known[k] = probes[k * 16 + 7] <= 0.5f;
    static const long d6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int ring = 0; ring < 8; ring++) {
        memcpy(was, known, np);
        memcpy(src, probes, np * 16 * sizeof(float));
        for (size_t k = 0; k < np; k++) {
            if (was[k]) continue;
            long x = (long)(k % nx), z = (long)((k / nx) % nz), y = (long)(k / (nx * nz));
            float sum[16] = {0};
            int c = 0;
            for (int d = 0; d < 6; d++) {
                long xx = x + d6[d][0], yy = y + d6[d][1], zz = z + d6[d][2];
                if (xx < 0 || yy < 0 || zz < 0 || xx >= (long)nx || yy >= (long)ny || zz >= (long)nz)
                    continue;
                size_t j = ((size_t)yy * nz + (size_t)zz) * nx + (size_t)xx;
                if (!was[j]) continue;
                for (int q = 0; q < 16; q++) sum[q] += src[j * 16 + q];
                c++;
            }
            if (c > 0) {
                for (int q = 0; q < 16; q++) probes[k * 16 + q] = sum[q] / (float)c;
                known[k] = 1;
            }
        }
    }
    free(known);
    free(was);
    free(src);
}

/* volume.rs:125-147 */
void orc_volume_irr(const float *probes, const uint32_t size[3], const float origin[3], float step,
                    const float p[3], const float n[3], float e[3], float *sun) {
    size_t nx = size[0], ny = size[1], nz = size[2];
    float f[3];
    for (int a = 0; a < 3; a++) f[a] = (p[a] - origin[a]) / step;
    float fx = clampf_rs(f[0], 0.0f, (float)(nx - 1));
    float fy = clampf_rs(f[1], 0.0f, (float)(ny - 1));
    float fz = clampf_rs(f[2], 0.0f, (float)(nz - 1));
    size_t x0 = (size_t)f2u_rs(fx), y0 = (size_t)f2u_rs(fy), z0 = (size_t)f2u_rs(fz);
    if (x0 > nx - 2) x0 = nx - 2;
    if (y0 > ny - 2) y0 = ny - 2;
    if (z0 > nz - 2) z0 = nz - 2;
    float tx = fx - (float)x0, ty = fy - (float)y0, tz = fz - (float)z0;
    float p4[4][4] = {{0}};
    for (int k = 0; k < 8; k++) {
        int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        float w = (dx ? tx : 1.0f - tx) * (dy ? ty : 1.0f - ty);
        w = w * (dz ? tz : 1.0f - tz);
        const float *q = &probes[(((y0 + dy) * nz + (z0 + dz)) * nx + (x0 + dx)) * 16];
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 4; c++) p4[a][c] += q[a * 4 + c] * w;
    }
    const float PI = 3.14159265358979323846f;
    float a0 = PI * 0.282095f, a1 = 2.0f * PI / 3.0f * 0.488603f;
    for (int c = 0; c < 3; c++) {
        float l1 = p4[1][c] * n[0] + p4[2][c] * n[1];
        l1 = l1 + p4[3][c] * n[2];
        float lobe = a0 * p4[0][c] + a1 * l1;
        e[c] = 0.5f * a0 * p4[0][c] + 0.5f * maxf_rs(lobe, 0.0f);
    }
    *sun = p4[0][3];
}

/* bake.rs:337-356 over `n` texels (16 floats each: pos, normal, dx, dy); direct and indirect rgb */
size_t orc_volume_lit(const float *probes, const uint32_t size[3], const float origin[3],
                      float step, const float sun_dir[3], const float sun_irr[3], size_t n,
                      const float *texels, uint8_t *known, uint8_t *state, float *field, float *ao,
                      float *direct, float *indirect) {
    size_t fixed = 0;
    for (size_t i = 0; i < n; i++) {
        const float *t = &texels[i * 16];
        if (!(t[3] > 0.5f) || known[i] != 0) continue;
        const float *nn = &t[4];
        float p[3];
        for (int a = 0; a < 3; a++) p[a] = t[a] + nn[a] * 0.15f;
        float e[3], s;
        orc_volume_irr(probes, size, origin, step, p, nn, e, &s);
        float d = nn[0] * sun_dir[0] + nn[1] * sun_dir[1];
        d = d + nn[2] * sun_dir[2];
        float c = maxf_rs(d, 0.0f);
        for (int k = 0; k < 3; k++) {
            indirect[i * 3 + k] = e[k];
            direct[i * 3 + k] = sun_irr[k] * c * s;
        }
        field[i] = s > 0.5f ? 1.0f : 0.0f;
        ao[i] = 1.0f;
        known[i] = 1;
        state[i] = VALIDITY_VOLUME;
        fixed++;
    }
    return fixed;
}

/* bake.rs:364-421: tiles (nt x 6 i32: key x, key z, w, h, atlas x, atlas y) and every tile's
 * TileLight, texels tile after tile, row-major in a tile. direct_out aw*ah*2, validity aw*ah,
 * ind_out iw*ih*3 with iw = ceil(aw / ds), ih = ceil(ah / ds). */
void orc_atlas(const int32_t *tiles, size_t nt, const uint8_t *known, const uint8_t *state,
               const float *field, const float *ao, const float *indirect, uint32_t aw,
               uint32_t ah, uint32_t ds, float *direct_out, uint8_t *validity, float *ind_out) {
    size_t iw = (aw + ds - 1) / ds, ih = (ah + ds - 1) / ds;
    for (size_t i = 0; i < (size_t)aw * ah; i++) {
        direct_out[i * 2] = 1.0f;
        direct_out[i * 2 + 1] = 1.0f;
        validity[i] = VALIDITY_NONE;
    }
    float *sum = (float *)calloc(iw * ih * 4, sizeof(float));
    size_t o = 0;
    for (size_t t = 0; t < nt; t++) {
        size_t w = (size_t)tiles[t * 6 + 2], h = (size_t)tiles[t * 6 + 3];
        size_t tx = (size_t)tiles[t * 6 + 4], ty = (size_t)tiles[t * 6 + 5];
        for (size_t y = 0; y < h; y++)
            for (size_t x = 0; x < w; x++) {
                size_t i = o + y * w + x;
                if (known[i] == 0) continue;
                size_t ax = tx + x, ay = ty + y;
                direct_out[(ay * aw + ax) * 2] = field[i];
                direct_out[(ay * aw + ax) * 2 + 1] = ao[i];
                validity[ay * aw + ax] = state[i];
                float k = known[i] == 2 ? 1.0f : 1e-3f;
                float *s = &sum[((ay / ds) * iw + ax / ds) * 4];
                for (int c = 0; c < 3; c++) s[c] += indirect[i * 3 + c] * k;
                s[3] += k;
            }
        o += w * h;
    }
    uint8_t *has = (uint8_t *)malloc(iw * ih), *was = (uint8_t *)malloc(iw * ih);
    float *src = (float *)malloc(iw * ih * 3 * sizeof(float));
    for (size_t i = 0; i < iw * ih; i++) {
        has[i] = sum[i * 4 + 3] > 0.0f;
        for (int c = 0; c < 3; c++) ind_out[i * 3 + c] = has[i] ? sum[i * 4 + c] / sum[i * 4 + 3] : 0.0f;
    }
    for (int ring = 0; ring < 2; ring++) {
        memcpy(was, has, iw * ih);
        memcpy(src, ind_out, iw * ih * 3 * sizeof(float));
        for (size_t y = 0; y < ih; y++)
            for (size_t x = 0; x < iw; x++) {
                size_t i = y * iw + x;
                if (was[i]) continue;
                float acc[3] = {0, 0, 0};
                int cnt = 0;
                for (size_t yy = y > 0 ? y - 1 : 0; yy < (y + 2 < ih ? y + 2 : ih); yy++)
                    for (size_t xx = x > 0 ? x - 1 : 0; xx < (x + 2 < iw ? x + 2 : iw); xx++) {
                        size_t j = yy * iw + xx;
                        if (!was[j]) continue;
                        for (int c = 0; c < 3; c++) acc[c] += src[j * 3 + c];
                        cnt++;
                    }
                if (cnt > 0) {
                    for (int c = 0; c < 3; c++) ind_out[i * 3 + c] = acc[c] / (float)cnt;
                    has[i] = 1;
                }
            }
    }
    free(sum);
    free(has);
    free(was);
    free(src);
}

/* output.rs:101 `q` and :167 `byte`: (v.clamp(0, 1) * 255).round() as u8 */
uint8_t orc_unorm8(float v) {
    float r = roundf(clampf_rs(v, 0.0f, 1.0f) * 255.0f);  /* round: half away from zero */
    uint32_t u = f2u_rs(r);
    return (uint8_t)(u > 255 ? 255 : u);
}

/* encode.rs:7-34 */
void orc_bc4_block(const uint8_t v[16], uint8_t out[8]) {
    uint8_t hi = v[0], lo = v[0];
    for (int k = 1; k < 16; k++) {
        if (v[k] > hi) hi = v[k];
        if (v[k] < lo) lo = v[k];
    }
    memset(out, 0, 8);
    out[0] = hi;
    out[1] = lo;
    if (hi == lo) return;
    float pal[8];
    pal[0] = (float)hi;
    pal[1] = (float)lo;
    for (int i = 2; i < 8; i++) pal[i] = ((float)(8 - i) * (float)hi + (float)(i - 1) * (float)lo) / 7.0f;
    uint64_t bits = 0;
    for (int k = 0; k < 16; k++) {
        float x = (float)v[k];
        int best = 0;
        for (int i = 1; i < 8; i++)
            if (fabsf(pal[i] - x) < fabsf(pal[best] - x)) best = i;
        bits |= (uint64_t)best << (3 * k);
    }
    for (int b = 0; b < 6; b++) out[2 + b] = (uint8_t)(bits >> (8 * b));
}

/* encode.rs:37-50: rg interleaved pairs, w x h (multiples of 4); out w*h bytes */
int orc_bc5(const uint8_t *rg, size_t w, size_t h, uint8_t *out) {
    if (w % 4 || h % 4) return -1;
    size_t o = 0;
    for (size_t by = 0; by < h; by += 4)
        for (size_t bx = 0; bx < w; bx += 4)
            for (int c = 0; c < 2; c++) {
                uint8_t block[16];
                for (int k = 0; k < 16; k++) block[k] = rg[((by + k / 4) * w + bx + k % 4) * 2 + c];
                orc_bc4_block(block, out + o);
                o += 8;
            }
    return 0;
}

/* encode.rs:53-69 */
uint32_t orc_rgb9e5(const float c[3]) {
    const float MAX = 65408.0f;
    float r = clampf_rs(c[0], 0.0f, MAX), g = clampf_rs(c[1], 0.0f, MAX), b = clampf_rs(c[2], 0.0f, MAX);
    float m = maxf_rs(maxf_rs(r, g), b);
    if (m <= 0.0f) return 0;
    float fl = floorf(log2f(m));
    int32_t lg = isnan(fl) ? 0 : (int32_t)fl;  /* `as i32` (m is finite and > 0 here, or NaN) */
    int32_t e = (lg > -16 ? lg : -16) + 1 + 15;
    float scale = ldexpf(1.0f, e - 15 - 9);  /* 2f32.powi: exact for these exponents */
    if (f2u_rs(floorf(m / scale + 0.5f)) == 512) {
        e += 1;
        scale *= 2.0f;
    }
    uint32_t q[3];
    float ch[3] = {r, g, b};
    for (int k = 0; k < 3; k++) {
        uint32_t u = f2u_rs(floorf(ch[k] / scale + 0.5f));
        q[k] = u < 511 ? u : 511;
    }
    return q[0] | (q[1] << 9) | (q[2] << 18) | ((uint32_t)e << 27);
}

/* encode.rs:71-74 */
void orc_rgb9e5_decode(uint32_t p, float out[3]) {
    float s = ldexpf(1.0f, (int)(p >> 27) - 15 - 9);
    out[0] = (float)(p & 511) * s;
    out[1] = (float)((p >> 9) & 511) * s;
    out[2] = (float)((p >> 18) & 511) * s;
}

/* output.rs:108 (the scale: the brightest L0, at least 1e-6) and :165-181 (16 bytes a probe) */
float orc_volume_bytes(const float *probes, size_t n, uint8_t *out) {
    float scale = 1e-6f;
    for (size_t k = 0; k < n; k++)
        for (int c = 0; c < 3; c++) scale = maxf_rs(scale, probes[k * 16 + c]);
    const float SQRT3 = 1.7320508f;
    for (size_t k = 0; k < n; k++) {
        const float *p = &probes[k * 16];
        uint8_t *o = &out[k * 16];
        for (int c = 0; c < 3; c++) o[c] = orc_unorm8(sqrtf(maxf_rs(p[c], 0.0f) / scale));
        o[3] = orc_unorm8(p[3]);
        for (int axis = 1; axis < 4; axis++)
            for (int c = 0; c < 3; c++) {
                float r = p[c] > 1e-6f ? p[axis * 4 + c] / (p[c] * SQRT3) : 0.0f;
                o[4 + (axis - 1) * 3 + c] = orc_unorm8(r * 0.5f + 0.5f);
            }
        o[13] = o[14] = o[15] = 0;
    }
    return scale;
}

/* batch forms for the checkers: output.rs:101 over a direct atlas (n texels of 2 f32), and
 * encode.rs:53-74 over n rgb triples */
void orc_unorm8_n(const float *v, size_t n, uint8_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = orc_unorm8(v[i]);
}
void orc_rgb9e5_n(const float *rgb, size_t n, uint32_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = orc_rgb9e5(&rgb[i * 3]);
}
void orc_rgb9e5_decode_n(const uint32_t *p, size_t n, float *out) {
    for (size_t i = 0; i < n; i++) orc_rgb9e5_decode(p[i], &out[i * 3]);
}
