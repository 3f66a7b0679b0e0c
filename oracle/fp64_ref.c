/*
 * ORACLE — test infrastructure, not product code.
 *
 * An FP64 tracer over the same rays as crates/lightbaker/src/gpu/trace.wgsl: the same BVH
 * (node boxes and triangles as stored, exact in double), the same sample sequences (the u32
 * hashes of pcg/rand/radical, converted to double), every other step in double (libm sin, cos,
 * sqrt; no FP32 rounding anywhere), sums in double. It is the yardstick the FP32 passes'
 * error is reported against: the per-item difference between an FP32 result and this one is the
 * rounding of the FP32 arithmetic (and the rare hit that changes with it).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define PI_D 3.14159265358979323846
#define FAR_D 1e30
#define NONE_U 0xffffffffu
#define STACK 48

typedef struct { double x, y, z; } d3;

static inline d3 mk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 scl(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 cross3(d3 a, d3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline d3 normalize3(d3 v) { return scl(v, 1.0 / sqrt(dot3(v, v))); }
static inline d3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
static uint32_t rev32(uint32_t v) {
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
    v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
    return (v >> 16) | (v << 16);
}

typedef struct {
    const lb_node *nodes;
    const lb_tri *tris;
    const float *materials;
    const float *charts;
    const float *light;
    const lb_params *P;
} ctx_t;

static inline double rnd(const ctx_t *c, uint32_t a, uint32_t b) {
    return (double)pcg(a ^ pcg(b + fbits(c->P->reach[3]))) / 4294967296.0;
}
static inline double radical(uint32_t i) { return (double)rev32(i) / 4294967296.0; }

typedef struct { d3 t, b, n; } frame_t;
static frame_t frame(d3 n) {
    d3 a = fabs(n.y) > 0.9 ? mk(1, 0, 0) : mk(0, 1, 0);
    frame_t f;
    f.t = normalize3(cross3(a, n));
    f.b = cross3(n, f.t);
    f.n = n;
    return f;
}
static d3 cosine(double u, double v, const frame_t *f) {
    double r = sqrt(u), a = v * 2.0 * PI_D;
    d3 l = mk(r * cos(a), r * sin(a), sqrt(fmax(0.0, 1.0 - u)));
    return normalize3(add(add(scl(f->t, l.x), scl(f->b, l.y)), scl(f->n, l.z)));
}
extern float orc_sky_sh[12]; /* trace_ref.c: the sky's L1 term */
static d3 sky(const ctx_t *c, d3 d) {
    d3 s = d.y < 0.0 ? ld3(c->P->ground)
                     : add(scl(ld3(c->P->horizon), 1.0 - d.y), scl(ld3(c->P->zenith), d.y));
    double v[3];
    for (int k = 0; k < 3; k++) {
        const float *q = orc_sky_sh;
        v[k] = 0.282095 * q[k] + 0.488603 * (q[3 + k] * d.x + q[6 + k] * d.y + q[9 + k] * d.z);
    }
    return add(s, mk(v[0], v[1], v[2]));
}
static int stops(const ctx_t *c, uint32_t tri, uint32_t ray) {
    double op = c->materials[fbits(c->tris[tri].v0[3]) * 4 + 3];
    return op >= 1.0 || rnd(c, tri * 9781u, ray) < op;
}
static double boxt(d3 o, d3 inv, const float *mn, const float *mx, double tmax) {
    double ax = (mn[0] - o.x) * inv.x, bx = (mx[0] - o.x) * inv.x;
    double ay = (mn[1] - o.y) * inv.y, by = (mx[1] - o.y) * inv.y;
    double az = (mn[2] - o.z) * inv.z, bz = (mx[2] - o.z) * inv.z;
    double lo = fmax(fmax(fmin(ax, bx), fmin(ay, by)), fmin(az, bz));
    double hi = fmin(fmin(fmax(ax, bx), fmax(ay, by)), fmax(az, bz));
    double l0 = fmax(lo, 0.0);
    /* a margin for the boxes' FP32 corners against the stored triangles */
    return (hi * (1.0 + 1e-9) + 1e-9 >= l0 && lo < tmax) ? l0 : FAR_D;
}

typedef struct { double t; uint32_t tri; int back; } hit_t;

static hit_t trace(const ctx_t *c, d3 o, d3 d, double tmax, int any, uint32_t ray) {
    hit_t hit = {tmax, NONE_U, 0};
    d3 inv = mk(1.0 / (fabs(d.x) < 1e-300 ? 1e-300 : d.x), 1.0 / (fabs(d.y) < 1e-300 ? 1e-300 : d.y),
                1.0 / (fabs(d.z) < 1e-300 ? 1e-300 : d.z));
    uint32_t stack[STACK * 4];
    uint32_t sp = 0, n = 0;
    for (;;) {
        const lb_node *node = &c->nodes[n];
        if (node->count > 0u) {
            for (uint32_t i = node->first; i < node->first + node->count; i++) {
                const lb_tri *tr = &c->tris[i];
                d3 e1 = ld3(tr->e1), e2 = ld3(tr->e2);
                d3 p = cross3(d, e2);
                double det = dot3(e1, p);
                if (fabs(det) < 1e-300) continue;
                double id = 1.0 / det;
                d3 s = sub(o, ld3(tr->v0));
                double u = dot3(s, p) * id;
                if (u < 0.0 || u > 1.0) continue;
                d3 q = cross3(s, e1);
                double v = dot3(d, q) * id;
                if (v < 0.0 || u + v > 1.0) continue;
                double t = dot3(e2, q) * id;
                if (t > 1e-4 && t < hit.t && stops(c, i, ray)) {
                    hit.t = t; hit.tri = i; hit.back = det < 0.0;
                    if (any) return hit;
                }
            }
        } else {
            const lb_node *a = &c->nodes[node->first], *b = &c->nodes[node->first + 1u];
            double ta = boxt(o, inv, a->min, a->max, hit.t), tb = boxt(o, inv, b->min, b->max, hit.t);
            uint32_t near = ta <= tb ? node->first : node->first + 1u;
            uint32_t far = ta <= tb ? node->first + 1u : node->first;
            if (fmin(ta, tb) < FAR_D) {
                if (fmax(ta, tb) < FAR_D && sp < STACK * 4) stack[sp++] = far;
                n = near;
                continue;
            }
        }
        if (sp == 0u) break;
        n = stack[--sp];
    }
    return hit;
}

static d3 bounced(const ctx_t *c, d3 o, d3 d, hit_t h) {
    const lb_tri *tr = &c->tris[h.tri];
    const float *m = &c->materials[fbits(tr->v0[3]) * 4];
    uint32_t ch = fbits(tr->e1[3]);
    d3 e = ld3(c->P->fallback);
    if (ch != NONE_U) {
        d3 p = add(o, scl(d, h.t));
        const float *r0 = &c->charts[ch * 8u], *r1 = &c->charts[ch * 8u + 4u];
        double x = r0[0] * p.x + r0[1] * p.y + r0[2] * p.z + r0[3];
        double y = r1[0] * p.x + r1[1] * p.y + r1[2] * p.z + r1[3];
        int ix = (int)fmin(fmax(floor(x), 0.0), (double)c->P->atlas[0] - 1.0);
        int iy = (int)fmin(fmax(floor(y), 0.0), (double)c->P->atlas[1] - 1.0);
        const float *l = &c->light[((uint32_t)iy * c->P->atlas[0] + (uint32_t)ix) * 4u];
        if (l[3] > 0.5f) e = ld3(l);
    }
    return mk(m[0] / PI_D * e.x, m[1] / PI_D * e.y, m[2] / PI_D * e.z);
}

static int inside(const ctx_t *c, d3 o, d3 n, uint32_t key) {
    frame_t f = frame(n);
    int back = 0;
    for (uint32_t r = 0; r < 16u; r++) {
        d3 d = cosine(((double)r + 0.5) / 16.0, radical(r), &f);
        hit_t h = trace(c, o, d, c->P->ground[3], 0, key + r + 7000u);
        if (h.tri != NONE_U && h.back) back++;
    }
    return (double)back / 16.0 > (double)bitsf(c->P->atlas[3]);
}

static void offset(uint32_t q, double *cx, double *cy) {
    if (q <= 4u) {
        *cx = (q & 1u) ? 0.3 : -0.3;
        *cy = q > 2u ? 0.3 : -0.3;
        return;
    }
    *cx = q == 5u ? -0.45 : (q == 6u ? 0.45 : 0.0);
    *cy = q == 7u ? -0.45 : (q == 8u ? 0.45 : 0.0);
}

static void direct_item(const ctx_t *c, const lb_texel *tx, uint32_t idx, double *out) {
    const lb_params *P = c->P;
    d3 n = ld3(tx->normal);
    uint32_t key = idx * 7919u, g = P->counts[0], gg = g * g;
    d3 sun = ld3(P->sun_dir);
    frame_t sf = frame(sun);
    d3 lift = scl(n, P->reach[2]);
    d3 pos = ld3(tx->pos), dx = ld3(tx->dx), dy = ld3(tx->dy), at = pos;
    uint32_t pick = 0, hard = 0;
    double span = 1.0, soft = 0.0;
    if (inside(c, add(at, lift), n, key)) {
        for (uint32_t q = 1u; q < 9u; q++) {
            double cx, cy;
            offset(q, &cx, &cy);
            d3 p = add(add(pos, scl(dx, cx)), scl(dy, cy));
            if (!inside(c, add(p, lift), n, key + q * 31u)) { at = p; pick = q; span = 0.4; break; }
        }
        if (pick == 0u) {
            d3 p = add(pos, scl(n, 0.03));
            if (!inside(c, add(p, lift), n, key + 331u)) { at = p; pick = 9u; span = 0.4; }
        }
    }
    if (dot3(sun, n) > 0.0) {
        for (uint32_t s = 0u; s < gg; s++) {
            double cx = (((double)(s % g) + 0.5) / g - 0.5) * span;
            double cy = (((double)(s / g) + 0.5) / g - 0.5) * span;
            d3 o = add(add(add(at, scl(dx, cx)), scl(dy, cy)), lift);
            if (trace(c, o, sun, FAR_D, 1, key + s).tri == NONE_U) hard |= 1u << s;
            uint32_t k = (s + (uint32_t)(rnd(c, key, 17u) * gg)) % gg;
            double a = k * 2.399963, r = P->sun_dir[3] * sqrt((k + 0.5) / gg);
            d3 d = normalize3(add(add(sun, scl(sf.t, cos(a) * r)), scl(sf.b, sin(a) * r)));
            if (trace(c, o, d, FAR_D, 1, key + s + 1000u).tri == NONE_U) soft += 1.0;
        }
        soft /= gg;
    }
    frame_t f = frame(n);
    d3 o = add(at, lift);
    uint32_t rays = P->counts[1];
    double s1 = rnd(c, key, 3u), s2 = rnd(c, key, 5u), occ = 0.0, back = 0.0;
    for (uint32_t r = 0u; r < rays; r++) {
        double u0 = (r + 0.5) / rays + s1, v0 = radical(r) + s2;
        d3 d = cosine(u0 - floor(u0), v0 - floor(v0), &f);
        hit_t h = trace(c, o, d, P->reach[0], 0, key + r + 2000u);
        if (h.tri == NONE_U) continue;
        double k = h.t / P->reach[0];
        occ += 1.0 - k * k;
        if (h.back && h.t < P->ground[3]) back += 1.0;
    }
    out[0] = soft;
    out[1] = (double)(hard | (pick << 24u));
    out[2] = occ / rays;
    out[3] = back / rays;
}

static void gather_item(const ctx_t *c, const lb_texel *tx, uint32_t idx, double *out) {
    const lb_params *P = c->P;
    d3 n = ld3(tx->normal);
    frame_t f = frame(n);
    uint32_t key = idx * 7919u, rays = P->counts[2];
    double s1 = rnd(c, key, 11u), s2 = rnd(c, key, 13u), back = 0.0;
    d3 sum = mk(0, 0, 0), pos = ld3(tx->pos), dx = ld3(tx->dx), dy = ld3(tx->dy);
    for (uint32_t r = 0u; r < rays; r++) {
        double jx = rnd(c, key + r, 19u) - 0.5, jy = rnd(c, key + r, 23u) - 0.5;
        d3 o = add(add(add(pos, scl(dx, jx)), scl(dy, jy)), scl(n, P->reach[2]));
        double u0 = (r + 0.5) / rays + s1, v0 = radical(r) + s2;
        d3 d = cosine(u0 - floor(u0), v0 - floor(v0), &f);
        hit_t h = trace(c, o, d, P->reach[1], 0, key + r + 4000u);
        if (h.tri == NONE_U) { sum = add(sum, sky(c, d)); continue; }
        if (h.back) { back += 1.0; continue; }
        sum = add(sum, bounced(c, o, d, h));
    }
    out[0] = sum.x * PI_D / rays;
    out[1] = sum.y * PI_D / rays;
    out[2] = sum.z * PI_D / rays;
    out[3] = back / rays;
}

static void volume_item(const ctx_t *c, const lb_texel *pr, double *out) {
    const lb_params *P = c->P;
    d3 o = ld3(pr->pos);
    uint32_t key = fbits(pr->normal[3]) * 7919u + 77u, rays = P->counts[2];
    double s1 = rnd(c, key, 29u), s2 = rnd(c, key, 31u), back = 0.0;
    double acc[12] = {0};
    for (uint32_t r = 0u; r < rays; r++) {
        double u0 = (r + 0.5) / rays + s1, v0 = radical(r) + s2;
        double u = u0 - floor(u0), v = v0 - floor(v0);
        double z = 1.0 - 2.0 * u, rr = sqrt(fmax(0.0, 1.0 - z * z)), a = v * 2.0 * PI_D;
        d3 d = mk(rr * cos(a), z, rr * sin(a));
        hit_t h = trace(c, o, d, P->reach[1], 0, key + r);
        d3 l = sky(c, d);
        if (h.tri != NONE_U) {
            if (h.back) back += 1.0;
            l = h.back ? mk(0, 0, 0) : bounced(c, o, d, h);
        }
        double ls[3] = {l.x, l.y, l.z};
        for (int k = 0; k < 3; k++) {
            acc[k] += ls[k] * 0.282095;
            acc[3 + k] += ls[k] * 0.488603 * d.x;
            acc[6 + k] += ls[k] * 0.488603 * d.y;
            acc[9 + k] += ls[k] * 0.488603 * d.z;
        }
    }
    double w = 4.0 * PI_D / rays;
    uint32_t g = P->counts[0], gg = g * g;
    d3 sund = ld3(P->sun_dir);
    frame_t sf = frame(sund);
    double sun = 0.0;
    for (uint32_t s = 0u; s < gg; s++) {
        double a = s * 2.399963, r = P->sun_dir[3] * sqrt((s + 0.5) / gg);
        d3 d = normalize3(add(add(sund, scl(sf.t, cos(a) * r)), scl(sf.b, sin(a) * r)));
        if (trace(c, o, d, FAR_D, 1, key + s + 9000u).tri == NONE_U) sun += 1.0;
    }
    for (int b = 0; b < 4; b++)
        for (int k = 0; k < 3; k++) out[b * 4 + k] = acc[b * 3 + k] * w;
    out[3] = sun / gg;
    out[7] = back / rays;
    out[11] = 0.0;
    out[15] = 0.0;
}

/* As orc_run, in double: out holds n * width vec4s of doubles. */
int orc64_run(int pass, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
              const float *materials, const float *charts, const lb_params *params,
              const lb_texel *items, size_t n, uint32_t key_base, const float *light, double *out) {
    static const float zero_light[4] = {0, 0, 0, 0};
    static const float zero_chart[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    lb_node tmp;
    ctx_t c0;
    if (n_nodes == 1 && nodes[0].count == 0) {
        tmp = nodes[0];
        tmp.count = 1;
        nodes = &tmp;
    }
    c0.nodes = nodes; c0.tris = tris; c0.materials = materials;
    c0.charts = charts ? charts : zero_chart;
    c0.light = light ? light : zero_light;
    c0.P = params;
    if (pass < 0 || pass > 2) return 1;
    long nn = (long)n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < nn; i++) {
        if (pass == 0) direct_item(&c0, &items[i], key_base + (uint32_t)i, out + 4 * i);
        else if (pass == 1) gather_item(&c0, &items[i], key_base + (uint32_t)i, out + 4 * i);
        else volume_item(&c0, &items[i], out + 16 * i);
    }
    return 0;
}
