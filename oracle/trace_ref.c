/*
 * ORACLE — test infrastructure, not product code. Only tests/, the smoke check and bench.py's
 * cpu_baseline leg load this library, and only as the checker or the CPU baseline.
 *
 * A scalar FP32 restatement of the reference's ray passes, crates/lightbaker/src/gpu/trace.wgsl,
 * in WGSL evaluation order (left to right, no fused multiply-add: build with -ffp-contract=off).
 * Each function cites the WGSL lines it restates. Choices WGSL leaves to the implementation are
 * fixed here and stated where they are made:
 *   - sin/cos: Cody-Waite reduction by pi/2 (the three-part split of Cephes' sinf, Moshier 1989)
 *     and Taylor polynomials on [-pi/4, pi/4] in Horner form (lb_sincos below);
 *   - sqrt and division: IEEE-754 correctly rounded;
 *   - normalize(v) = v / sqrt(dot(v, v)); mix(a, b, t) = a * (1 - t) + b * t (the WGSL spec forms);
 *   - u32 -> f32: round to nearest even; f32 -> i32 of the atlas index: saturating.
 *
 * Parity pins (tests/test_oracle.py): pcg/rand/radical against integer closed forms; the
 * traversal against a brute-force loop over every triangle; gather/AO/volume against closed forms
 * (open plane under a gradient sky, an infinite ceiling, constant radiance); sun bits against the
 * analytic shadow of a box.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define PI_F 3.14159265f
#define FAR_F 1e30f
#define NONE_U 0xffffffffu
#define PUSH_F 0.03f
#define STACK 48

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 normalize3(v3 v) {
    float l = sqrtf(dot3(v, v));
    return mk(v.x / l, v.y / l, v.z / l);
}
static inline float fminv(float a, float b) { return b < a ? b : a; }
static inline float fmaxv(float a, float b) { return b > a ? b : a; }
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* sin and cos together: x = q * pi/2 + r, |r| <= pi/4, r found with pi/2 split in three parts so
 * that q * part is exact for |q| < 2^15; then sin r = r - r^3/3! + ... - r^11/11! and
 * cos r = 1 - r^2/2! + ... - r^10/10! (Horner in r^2), and the quadrant q mod 4 picks the pair. */
void lb_sincos(float x, float *s, float *c) {
    float q = rintf(x * 0.636619772f);
    float r = x - q * 1.5703125f;
    r = r - q * 4.837512969970703125e-4f;
    r = r - q * 7.54978995489188216e-8f;
    float r2 = r * r;
    float ps = -2.5052108385441718775e-8f;
    ps = ps * r2 + 2.7557319223985890653e-6f;
    ps = ps * r2 + -1.9841269841269841270e-4f;
    ps = ps * r2 + 8.3333333333333333333e-3f;
    ps = ps * r2 + -1.6666666666666666667e-1f;
    float sr = r + (r * r2) * ps;
    float pc = -2.7557319223985890653e-7f;
    pc = pc * r2 + 2.4801587301587301587e-5f;
    pc = pc * r2 + -1.3888888888888888889e-3f;
    pc = pc * r2 + 4.1666666666666666667e-2f;
    pc = pc * r2 + -0.5f;
    float cr = 1.0f + r2 * pc;
    int qi = ((int)q) & 3;
    switch (qi) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

/* trace.wgsl:41-45 */
uint32_t orc_pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

typedef struct {
    const lb_node *nodes;
    const lb_tri *tris;
    const float *materials;
    const float *charts;
    const float *light;
    const lb_params *P;
    orc_counters *cnt;
    const uint32_t *quant; /* the TUNED form's 8-bit node records (lb_tuned_nodes), or NULL */
    /* with `cmp` (and `cmp_quant`): every ray is traced over both node forms, the reference's
     * result is kept, and cmp counts rays, rays whose results differ, and those among them that
     * are no near-tie (a hit against none, another any-hit outcome, or distances more than
     * 2^-20 t apart); an any-hit ray counts as differing only when one form is blocked and the
     * other not */
    unsigned long long *cmp;
    const uint32_t *cmp_quant;
} ctx_t;

/* trace.wgsl:46 */
static inline float rnd(const ctx_t *c, uint32_t a, uint32_t b) {
    return (float)orc_pcg(a ^ orc_pcg(b + fbits(c->P->reach[3]))) / 4294967296.0f;
}
float orc_rand(uint32_t a, uint32_t b, float seed) {
    return (float)orc_pcg(a ^ orc_pcg(b + fbits(seed))) / 4294967296.0f;
}

/* trace.wgsl:49 */
static inline uint32_t rev32(uint32_t v) {
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
    v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
    return (v >> 16) | (v << 16);
}
float orc_radical(uint32_t i) { return (float)rev32(i) / 4294967296.0f; }

/* trace.wgsl:51-55: columns t, cross(n, t), n */
typedef struct { v3 t, b, n; } frame_t;
static frame_t frame(v3 n) {
    v3 a = fabsf(n.y) > 0.9f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
    frame_t f;
    f.t = normalize3(cross3(a, n));
    f.b = cross3(n, f.t);
    f.n = n;
    return f;
}
/* mat3x3f * vec3f: column 0 * x + column 1 * y + column 2 * z */
static inline v3 fmul(const frame_t *f, v3 v) {
    return add(add(scl(f->t, v.x), scl(f->b, v.y)), scl(f->n, v.z));
}

/* trace.wgsl:57-61 */
static v3 cosine(float u, float v, const frame_t *f) {
    float r = sqrtf(u);
    float a = v * 2.0f * PI_F;
    float sa, ca;
    lb_sincos(a, &sa, &ca);
    return normalize3(fmul(f, mk(r * ca, r * sa, sqrtf(fmaxv(0.0f, 1.0f - u)))));
}
void orc_cosine(float u, float v, const float n[3], float out[3]) {
    frame_t f = frame(ld3(n));
    v3 d = cosine(u, v, &f);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

/* trace.wgsl:63-66 */
/* Not the reference: the sky's L1 spherical-harmonic term (orc_set_sky_sh), added where a ray
 * escapes, per channel 0.282095 c0 + 0.488603 ((c1 d.x + c2 d.y) + c3 d.z); all zero: none. */
float orc_sky_sh[12];
int orc_sky_on;
void orc_set_sky_sh(const float *sh) {
    orc_sky_on = 0;
    for (int k = 0; k < 12; k++) {
        orc_sky_sh[k] = sh ? sh[k] : 0.0f;
        orc_sky_on |= orc_sky_sh[k] != 0.0f;
    }
}

static v3 sky(const ctx_t *c, v3 d) {
    const lb_params *P = c->P;
    v3 s;
    if (d.y < 0.0f) {
        s = ld3(P->ground);
    } else {
        float w = 1.0f - d.y;
        s = add(scl(ld3(P->horizon), w), scl(ld3(P->zenith), d.y));
    }
    if (orc_sky_on) {
        float v[3];
        for (int k = 0; k < 3; k++) {
            const float *q = orc_sky_sh;
            v[k] = 0.282095f * q[k] + 0.488603f * ((q[3 + k] * d.x + q[6 + k] * d.y) + q[9 + k] * d.z);
        }
        s = add(s, mk(v[0], v[1], v[2]));
    }
    return s;
}

/* trace.wgsl:69-72 */
static int stops(const ctx_t *c, uint32_t tri, uint32_t ray) {
    if (c->cnt) c->cnt->material_loads++;
    float op = c->materials[fbits(c->tris[tri].v0[3]) * 4 + 3];
    return op >= 1.0f || rnd(c, tri * 9781u, ray) < op;
}

/* trace.wgsl:74-80 */
static float boxt(v3 o, v3 inv, const float *mn, const float *mx, float tmax) {
    v3 a = mul(sub(ld3(mn), o), inv);
    v3 b = mul(sub(ld3(mx), o), inv);
    float lo = fmaxv(fmaxv(fminv(a.x, b.x), fminv(a.y, b.y)), fminv(a.z, b.z));
    float hi = fminv(fminv(fmaxv(a.x, b.x), fmaxv(a.y, b.y)), fmaxv(a.z, b.z));
    float l0 = fmaxv(lo, 0.0f);
    return (hi >= l0 && lo < tmax) ? l0 : FAR_F;
}

typedef struct { float t; uint32_t tri; int back; } hit_t;

/* Not the reference: one inner-node visit over the TUNED form's 8-bit child boxes (a record of
 * lb_tuned_nodes: per axis the origin's bits, their low 9 bits zero, with e + 151 in the low 8,
 * so the word shifted left by 23 reads as 2^24 times the step 2^e; the 12 corner bytes), with
 * the device's arithmetic (lb_traverse.h quant_slabs): the plane of byte q at
 * t = fma(q 2^-24, 2^(e+24) * inv, (origin - o) * inv) (the same rounding as
 * fma(q, 2^e * inv, ...): the scalings are exact), the entry max(0, ...) and the exit min, a child
 * hit when exit >= entry and entry < t, the near child the hit one of the lower entry (a on a tie),
 * the far one pushed when both are hit. Moves to the near child and returns 1, else returns 0. The
 * boxes contain the reference's, so the same triangles can be met; the order and the entries
 * differ, which decides between two hits only when their distances lie within a few ulps. */
static int quant_visit(const uint32_t *q, v3 o, v3 inv, float tmax, uint32_t first, uint32_t *n,
                       uint32_t *stack, uint32_t *sp) {
    float kx = bitsf(q[0] << 23) * inv.x, ky = bitsf(q[1] << 23) * inv.y,
          kz = bitsf(q[2] << 23) * inv.z;
    float cx = (bitsf(q[0] & 0xffffff00u) - o.x) * inv.x;
    float cy = (bitsf(q[1] & 0xffffff00u) - o.y) * inv.y;
    float cz = (bitsf(q[2] & 0xffffff00u) - o.z) * inv.z;
    const uint8_t *b8 = (const uint8_t *)(q + 4); /* a min xyz, a max xyz, b min xyz, b max xyz */
    float b[12];
    for (int k = 0; k < 12; k++) b[k] = (float)b8[k] * 0x1p-24f;
    float x0 = fmaf(b[0], kx, cx), y0 = fmaf(b[1], ky, cy), z0 = fmaf(b[2], kz, cz);
    float x1 = fmaf(b[3], kx, cx), y1 = fmaf(b[4], ky, cy), z1 = fmaf(b[5], kz, cz);
    float la = fmaxv(fmaxv(fmaxv(0.0f, fminv(x0, x1)), fminv(y0, y1)), fminv(z0, z1));
    float ea = fminv(fminv(fmaxv(x0, x1), fmaxv(y0, y1)), fmaxv(z0, z1));
    x0 = fmaf(b[6], kx, cx); y0 = fmaf(b[7], ky, cy); z0 = fmaf(b[8], kz, cz);
    x1 = fmaf(b[9], kx, cx); y1 = fmaf(b[10], ky, cy); z1 = fmaf(b[11], kz, cz);
    float lb = fmaxv(fmaxv(fmaxv(0.0f, fminv(x0, x1)), fminv(y0, y1)), fminv(z0, z1));
    float eb = fminv(fminv(fmaxv(x0, x1), fmaxv(y0, y1)), fmaxv(z0, z1));
    int ha = ea >= la && la < tmax, hb = eb >= lb && lb < tmax;
    if (!(ha || hb)) return 0;
    int an = ha && (!hb || la <= lb);
    if (ha && hb && *sp < STACK) stack[(*sp)++] = an ? first + 1u : first;
    *n = an ? first : first + 1u;
    return 1;
}

/* trace.wgsl:83-130 */
static hit_t trace_nodes(const ctx_t *c, v3 o, v3 d, float tmax, int any, uint32_t ray);

/* the first rays orc_quant_compare finds differing beyond a near-tie: o, d, tmax, any, the two
 * forms' t and triangles (bits), the ray key (bits) */
#define CMP_EXAMPLES 16
static float g_cmp_examples[CMP_EXAMPLES][13];
void orc_quant_examples(float out[CMP_EXAMPLES][13]) { memcpy(out, g_cmp_examples, sizeof g_cmp_examples); }

static hit_t trace(const ctx_t *c, v3 o, v3 d, float tmax, int any, uint32_t ray) {
    if (!c->cmp) return trace_nodes(c, o, d, tmax, any, ray);
    hit_t hf = trace_nodes(c, o, d, tmax, any, ray);
    ctx_t cq = *c;
    cq.quant = c->cmp_quant;
    cq.cnt = NULL;
    hit_t hq = trace_nodes(&cq, o, d, tmax, any, ray);
    __atomic_fetch_add(&c->cmp[0], 1ull, __ATOMIC_RELAXED);
    /* an any-hit ray's result is whether it was blocked; a closest-hit ray's, its hit */
    int differ = any ? (hf.tri == NONE_U) != (hq.tri == NONE_U)
                     : (hf.tri != hq.tri || hf.t != hq.t || hf.back != hq.back);
    if (differ) {
        __atomic_fetch_add(&c->cmp[1], 1ull, __ATOMIC_RELAXED);
        int near_tie = !any && hf.tri != NONE_U && hq.tri != NONE_U &&
                       fabsf(hf.t - hq.t) <= ldexpf(fmaxv(hf.t, hq.t), -20);
        if (!near_tie) {
            unsigned long long k = __atomic_fetch_add(&c->cmp[2], 1ull, __ATOMIC_RELAXED);
            if (k < CMP_EXAMPLES) {
                float *e = g_cmp_examples[k];
                e[0] = o.x; e[1] = o.y; e[2] = o.z; e[3] = d.x; e[4] = d.y; e[5] = d.z;
                e[6] = tmax; e[7] = (float)any; e[8] = hf.t; e[9] = hq.t;
                e[10] = bitsf(hf.tri); e[11] = bitsf(hq.tri); e[12] = bitsf(ray);
            }
        }
    }
    return hf;
}

static hit_t trace_nodes(const ctx_t *c, v3 o, v3 d, float tmax, int any, uint32_t ray) {
    hit_t hit = {tmax, NONE_U, 0};
    v3 dd = mk(fabsf(d.x) < 1e-12f ? 1e-12f : d.x, fabsf(d.y) < 1e-12f ? 1e-12f : d.y,
               fabsf(d.z) < 1e-12f ? 1e-12f : d.z);
    v3 inv = mk(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
    uint32_t stack[STACK];
    uint32_t sp = 0, n = 0;
    if (c->cnt) c->cnt->rays++;
    for (;;) {
        const lb_node *node = &c->nodes[n];
        if (node->count > 0u) {
            if (c->cnt) c->cnt->leaf_visits++;
            for (uint32_t i = node->first; i < node->first + node->count; i++) {
                const lb_tri *tr = &c->tris[i];
                if (c->cnt) c->cnt->tri_tests++;
                v3 e1 = ld3(tr->e1), e2 = ld3(tr->e2);
                v3 p = cross3(d, e2);
                float det = dot3(e1, p);
                if (fabsf(det) < 1e-12f) continue;
                float id = 1.0f / det;
                v3 s = sub(o, ld3(tr->v0));
                float u = dot3(s, p) * id;
                if (u < 0.0f || u > 1.0f) continue;
                v3 q = cross3(s, e1);
                float v = dot3(d, q) * id;
                if (v < 0.0f || u + v > 1.0f) continue;
                float t = dot3(e2, q) * id;
                if (t > 1e-4f && t < hit.t && stops(c, i, ray)) {
                    hit.t = t; hit.tri = i; hit.back = det < 0.0f;
                    if (any) return hit;
                }
            }
        } else if (c->quant) {
            if (c->cnt) c->cnt->inner_visits++;
            if (quant_visit(c->quant + (size_t)n * 8u, o, inv, hit.t, node->first, &n, stack, &sp))
                continue;
        } else {
            const lb_node *a = &c->nodes[node->first];
            const lb_node *b = &c->nodes[node->first + 1u];
            if (c->cnt) c->cnt->inner_visits++;
            float ta = boxt(o, inv, a->min, a->max, hit.t);
            float tb = boxt(o, inv, b->min, b->max, hit.t);
            uint32_t near = ta <= tb ? node->first : node->first + 1u;
            uint32_t far = ta <= tb ? node->first + 1u : node->first;
            float tn = fminv(ta, tb), tf = fmaxv(ta, tb);
            if (tn < FAR_F) {
                if (tf < FAR_F && sp < STACK) stack[sp++] = far;
                n = near;
                continue;
            }
        }
        if (sp == 0u) break;
        sp--;
        n = stack[sp];
    }
    return hit;
}

/* trace.wgsl:133-148 */
static v3 bounced(const ctx_t *c, v3 o, v3 d, hit_t h) {
    const lb_params *P = c->P;
    const lb_tri *tr = &c->tris[h.tri];
    const float *m = &c->materials[fbits(tr->v0[3]) * 4];
    v3 albedo = ld3(m);
    if (c->cnt) { c->cnt->material_loads++; c->cnt->light_gathers++; }
    uint32_t ch = fbits(tr->e1[3]);
    v3 e = ld3(P->fallback);
    if (ch != NONE_U) {
        v3 p = add(o, scl(d, h.t));
        const float *r0 = &c->charts[ch * 8u], *r1 = &c->charts[ch * 8u + 4u];
        float x = r0[0] * p.x + r0[1] * p.y + r0[2] * p.z + r0[3] * 1.0f;
        float y = r1[0] * p.x + r1[1] * p.y + r1[2] * p.z + r1[3] * 1.0f;
        int ix = orc_sat_floor(x, (int)P->atlas[0] - 1);
        int iy = orc_sat_floor(y, (int)P->atlas[1] - 1);
        const float *l = &c->light[((uint32_t)iy * P->atlas[0] + (uint32_t)ix) * 4u];
        if (l[3] > 0.5f) e = ld3(l);
    }
    /* albedo / PI * e: (albedo / PI) first, then times e */
    return mul(mk(albedo.x / PI_F, albedo.y / PI_F, albedo.z / PI_F), e);
}

/* clamp(i32(floor(x)), 0, hi): the conversion saturates, as WGSL's does */
int orc_sat_floor(float x, int hi) {
    float f = floorf(x);
    if (!(f >= 0.0f)) return 0; /* negative or NaN */
    if (f >= (float)hi) return hi;
    return (int)f;
}

/* trace.wgsl:152-156 */
static void offset(uint32_t q, float *cx, float *cy) {
    if (q <= 4u) {
        *cx = (q & 1u) == 1u ? 0.3f : -0.3f;
        *cy = q > 2u ? 0.3f : -0.3f;
        return;
    }
    static const float e[4][2] = {{-0.45f, 0.0f}, {0.45f, 0.0f}, {0.0f, -0.45f}, {0.0f, 0.45f}};
    *cx = e[q - 5u][0];
    *cy = e[q - 5u][1];
}

/* trace.wgsl:159-168 */
static int inside(const ctx_t *c, v3 o, v3 n, uint32_t key) {
    frame_t f = frame(n);
    float back = 0.0f;
    for (uint32_t r = 0; r < 16u; r++) {
        v3 d = cosine(((float)r + 0.5f) / 16.0f, orc_radical(r), &f);
        hit_t h = trace(c, o, d, c->P->ground[3], 0, key + r + 7000u);
        if (h.tri != NONE_U && h.back) back += 1.0f;
    }
    return back / 16.0f > bitsf(c->P->atlas[3]);
}

/* trace.wgsl:170-243, item i with key index `idx` */
static void direct_item(const ctx_t *c, const lb_texel *tx, uint32_t idx, float *out) {
    const lb_params *P = c->P;
    v3 n = ld3(tx->normal);
    uint32_t key = idx * 7919u;
    uint32_t g = P->counts[0];
    v3 sun = ld3(P->sun_dir);
    frame_t sf = frame(sun);
    float soft = 0.0f;
    uint32_t hard = 0u;
    v3 lift = scl(n, P->reach[2]);
    v3 pos = ld3(tx->pos), dx = ld3(tx->dx), dy = ld3(tx->dy);
    v3 at = pos;
    uint32_t pick = 0u;
    float span = 1.0f;
    if (inside(c, add(at, lift), n, key)) {
        for (uint32_t q = 1u; q < 9u; q++) {
            float cx, cy;
            offset(q, &cx, &cy);
            v3 p = add(add(pos, scl(dx, cx)), scl(dy, cy));
            if (!inside(c, add(p, lift), n, key + q * 31u)) {
                at = p; pick = q; span = 0.4f;
                break;
            }
        }
        if (pick == 0u) {
            v3 p = add(pos, scl(n, PUSH_F));
            if (!inside(c, add(p, lift), n, key + 331u)) {
                at = p; pick = 9u; span = 0.4f;
            }
        }
    }
    if (dot3(sun, n) > 0.0f) {
        uint32_t gg = g * g;
        for (uint32_t s = 0u; s < gg; s++) {
            float cx = (((float)(s % g) + 0.5f) / (float)g - 0.5f) * span;
            float cy = (((float)(s / g) + 0.5f) / (float)g - 0.5f) * span;
            v3 o = add(add(add(at, scl(dx, cx)), scl(dy, cy)), lift);
            if (trace(c, o, sun, FAR_F, 1, key + s).tri == NONE_U) hard |= 1u << s;
            uint32_t k = (s + (uint32_t)(rnd(c, key, 17u) * (float)gg)) % gg;
            float a = (float)k * 2.399963f;
            float r = P->sun_dir[3] * sqrtf(((float)k + 0.5f) / (float)gg);
            float sa, ca;
            lb_sincos(a, &sa, &ca);
            v3 d = normalize3(add(add(sun, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
            if (trace(c, o, d, FAR_F, 1, key + s + 1000u).tri == NONE_U) soft += 1.0f;
        }
        soft /= (float)gg;
    }
    frame_t f = frame(n);
    v3 o = add(at, lift);
    uint32_t rays = P->counts[1];
    float s1 = rnd(c, key, 3u), s2 = rnd(c, key, 5u);
    float occ = 0.0f, back = 0.0f;
    for (uint32_t r = 0u; r < rays; r++) {
        float u0 = ((float)r + 0.5f) / (float)rays + s1;
        float v0 = orc_radical(r) + s2;
        v3 d = cosine(u0 - floorf(u0), v0 - floorf(v0), &f);
        hit_t h = trace(c, o, d, P->reach[0], 0, key + r + 2000u);
        if (h.tri == NONE_U) continue;
        float k = h.t / P->reach[0];
        occ += 1.0f - k * k;
        if (h.back && h.t < P->ground[3]) back += 1.0f;
    }
    out[0] = soft;
    out[1] = bitsf(hard | (pick << 24u));
    out[2] = occ / (float)rays;
    out[3] = back / (float)rays;
}

/* trace.wgsl:245-270 */
static void gather_item(const ctx_t *c, const lb_texel *tx, uint32_t idx, float *out) {
    const lb_params *P = c->P;
    v3 n = ld3(tx->normal);
    frame_t f = frame(n);
    uint32_t key = idx * 7919u;
    uint32_t rays = P->counts[2];
    float s1 = rnd(c, key, 11u), s2 = rnd(c, key, 13u);
    v3 sum = mk(0.0f, 0.0f, 0.0f);
    float back = 0.0f;
    v3 pos = ld3(tx->pos), dx = ld3(tx->dx), dy = ld3(tx->dy);
    for (uint32_t r = 0u; r < rays; r++) {
        float jx = rnd(c, key + r, 19u) - 0.5f;
        float jy = rnd(c, key + r, 23u) - 0.5f;
        v3 o = add(add(add(pos, scl(dx, jx)), scl(dy, jy)), scl(n, P->reach[2]));
        float u0 = ((float)r + 0.5f) / (float)rays + s1;
        float v0 = orc_radical(r) + s2;
        v3 d = cosine(u0 - floorf(u0), v0 - floorf(v0), &f);
        hit_t h = trace(c, o, d, P->reach[1], 0, key + r + 4000u);
        if (h.tri == NONE_U) { sum = add(sum, sky(c, d)); continue; }
        if (h.back) { back += 1.0f; continue; }
        sum = add(sum, bounced(c, o, d, h));
    }
    out[0] = sum.x * PI_F / (float)rays;
    out[1] = sum.y * PI_F / (float)rays;
    out[2] = sum.z * PI_F / (float)rays;
    out[3] = back / (float)rays;
}

/* trace.wgsl:272-324 */
static void volume_item(const ctx_t *c, const lb_texel *pr, float *out) {
    const lb_params *P = c->P;
    v3 o = ld3(pr->pos);
    uint32_t key = fbits(pr->normal[3]) * 7919u + 77u;
    uint32_t rays = P->counts[2];
    float s1 = rnd(c, key, 29u), s2 = rnd(c, key, 31u);
    v3 l0 = mk(0, 0, 0), l1x = mk(0, 0, 0), l1y = mk(0, 0, 0), l1z = mk(0, 0, 0);
    float back = 0.0f;
    for (uint32_t r = 0u; r < rays; r++) {
        float u0 = ((float)r + 0.5f) / (float)rays + s1;
        float v0 = orc_radical(r) + s2;
        float u = u0 - floorf(u0), v = v0 - floorf(v0);
        float z = 1.0f - 2.0f * u;
        float rr = sqrtf(fmaxv(0.0f, 1.0f - z * z));
        float a = v * 2.0f * PI_F;
        float sa, ca;
        lb_sincos(a, &sa, &ca);
        v3 d = mk(rr * ca, z, rr * sa);
        hit_t h = trace(c, o, d, P->reach[1], 0, key + r);
        v3 l = sky(c, d);
        if (h.tri != NONE_U) {
            if (h.back) back += 1.0f;
            /* select(bounced(..), 0, back): bounced has no effects, so only a front hit calls it */
            l = h.back ? mk(0.0f, 0.0f, 0.0f) : bounced(c, o, d, h);
        }
        l0 = add(l0, scl(l, 0.282095f));
        l1x = add(l1x, scl(scl(l, 0.488603f), d.x));
        l1y = add(l1y, scl(scl(l, 0.488603f), d.y));
        l1z = add(l1z, scl(scl(l, 0.488603f), d.z));
    }
    float w = 4.0f * PI_F / (float)rays;
    uint32_t g = P->counts[0], gg = g * g;
    v3 sund = ld3(P->sun_dir);
    frame_t sf = frame(sund);
    float sun = 0.0f;
    for (uint32_t s = 0u; s < gg; s++) {
        float a = (float)s * 2.399963f;
        float r = P->sun_dir[3] * sqrtf(((float)s + 0.5f) / (float)gg);
        float sa, ca;
        lb_sincos(a, &sa, &ca);
        v3 d = normalize3(add(add(sund, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
        if (trace(c, o, d, FAR_F, 1, key + s + 9000u).tri == NONE_U) sun += 1.0f;
    }
    out[0] = l0.x * w; out[1] = l0.y * w; out[2] = l0.z * w; out[3] = sun / (float)gg;
    out[4] = l1x.x * w; out[5] = l1x.y * w; out[6] = l1x.z * w; out[7] = back / (float)rays;
    out[8] = l1y.x * w; out[9] = l1y.y * w; out[10] = l1y.z * w; out[11] = 0.0f;
    out[12] = l1z.x * w; out[13] = l1z.y * w; out[14] = l1z.z * w; out[15] = 0.0f;
}

/* Gpu::run (gpu/mod.rs:140-196) without the batches: item i's key index is key_base + i, which
 * is what `first + i` gives over the batches. Parallel over items with OpenMP when built so;
 * counters (optional) are only collected single-threaded. */
/* The reference's empty-scene marker (bvh.rs:76-80: one node, count 0) read literally is an inner
 * node whose children lie past the array (trace.wgsl:111-112); its documented meaning is a box no
 * ray reaches, so it is read here as one leaf over the zero triangle (determinant 0: never hit). */
static const lb_node *empty_fix(const lb_node *nodes, size_t n_nodes, lb_node *tmp) {
    if (n_nodes == 1 && nodes[0].count == 0) {
        *tmp = nodes[0];
        tmp->count = 1;
        return tmp;
    }
    return nodes;
}

static int run_ctx(int pass, const lb_node *nodes, size_t n_nodes, const uint32_t *quant,
                   const lb_tri *tris, const float *materials, const float *charts,
                   const lb_params *params, const lb_texel *items, size_t n, uint32_t key_base,
                   const float *light, float *out, orc_counters *cnt, unsigned long long *cmp,
                   const uint32_t *cmp_quant) {
    static const float zero_light[4] = {0, 0, 0, 0};
    static const float zero_chart[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    lb_node tmp;
    ctx_t c0;
    c0.nodes = empty_fix(nodes, n_nodes, &tmp); c0.tris = tris; c0.materials = materials;
    c0.charts = charts ? charts : zero_chart;
    c0.light = light ? light : zero_light;
    c0.P = params; c0.cnt = cnt; c0.quant = quant;
    c0.cmp = cmp; c0.cmp_quant = cmp_quant;
    if (pass < 0 || pass > 2) return 1;
    long nn = (long)n;
    if (cnt) {
        for (long i = 0; i < nn; i++) {
            if (pass == 0) direct_item(&c0, &items[i], key_base + (uint32_t)i, out + 4 * i);
            else if (pass == 1) gather_item(&c0, &items[i], key_base + (uint32_t)i, out + 4 * i);
            else volume_item(&c0, &items[i], out + 16 * i);
        }
        return 0;
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < nn; i++) {
        if (pass == 0) direct_item(&c0, &items[i], key_base + (uint32_t)i, out + 4 * i);
        else if (pass == 1) gather_item(&c0, &items[i], key_base + (uint32_t)i, out + 4 * i);
        else volume_item(&c0, &items[i], out + 16 * i);
    }
    return 0;
}

int orc_run(int pass, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
            const float *materials, const float *charts, const lb_params *params,
            const lb_texel *items, size_t n, uint32_t key_base, const float *light, float *out,
            orc_counters *cnt) {
    return run_ctx(pass, nodes, n_nodes, NULL, tris, materials, charts, params, items, n,
                   key_base, light, out, cnt, NULL, NULL);
}

/* Not the reference: orc_run over the TUNED form's 8-bit node records (`quant`, n_nodes x 8
 * words from lb_tuned_nodes) in place of the reference's boxes, visited as the device visits
 * them. The yardstick of the quantized form's hits; where it differs from orc_run it differs by
 * a near-tie between two faces (their distances a few ulps apart). */
int orc_run_quant(int pass, const lb_node *nodes, size_t n_nodes, const uint32_t *quant,
                  const lb_tri *tris, const float *materials, const float *charts,
                  const lb_params *params, const lb_texel *items, size_t n, uint32_t key_base,
                  const float *light, float *out, orc_counters *cnt) {
    if (!quant) return 1;
    return run_ctx(pass, nodes, n_nodes, quant, tris, materials, charts, params, items, n,
                   key_base, light, out, cnt, NULL, NULL);
}

/* Not the reference: orc_run (its results in `out`) with every ray also traced over `quant`;
 * counts[0..2]: rays, rays the two node forms end differently, and those among them that are no
 * near-tie (see ctx_t.cmp). */
int orc_quant_compare(int pass, const lb_node *nodes, size_t n_nodes, const uint32_t *quant,
                      const lb_tri *tris, const float *materials, const float *charts,
                      const lb_params *params, const lb_texel *items, size_t n,
                      uint32_t key_base, const float *light, float *out,
                      unsigned long long counts[3]) {
    if (!quant || !counts) return 1;
    counts[0] = counts[1] = counts[2] = 0;
    return run_ctx(pass, nodes, n_nodes, NULL, tris, materials, charts, params, items, n,
                   key_base, light, out, NULL, counts, quant);
}

/* What each gather ray's light depends on (trace.wgsl:248-268), for the adjoint's checker: per
 * item i and ray r, at (i * rays + r) * 2: a word (kind 0: the atlas texel bounced() reads,
 * trace.wgsl:137-145; kind 3: the bits of the direction's y; else 0) and a tag (kind << 30 |
 * material; kinds 0 a front hit on a chart, 1 a front hit off the charts, 2 a back face, 3 the
 * sky). */
int orc_gather_records(const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                       const float *materials, const float *charts, const lb_params *params,
                       const lb_texel *items, size_t n, uint32_t key_base, uint32_t *out) {
    static const float zero_chart[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    lb_node tmp;
    ctx_t c0;
    c0.nodes = empty_fix(nodes, n_nodes, &tmp); c0.tris = tris; c0.materials = materials;
    c0.charts = charts ? charts : zero_chart;
    c0.light = NULL; c0.quant = NULL; c0.cmp = NULL; c0.cmp_quant = NULL;
    c0.P = params; c0.cnt = NULL;
    uint32_t rays = params->counts[2];
    long nn = (long)n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < nn; i++) {
        const ctx_t *c = &c0;
        const lb_texel *tx = &items[i];
        const lb_params *P = c->P;
        v3 nrm = ld3(tx->normal);
        frame_t f = frame(nrm);
        uint32_t key = (key_base + (uint32_t)i) * 7919u;
        float s1 = rnd(c, key, 11u), s2 = rnd(c, key, 13u);
        v3 pos = ld3(tx->pos), dx = ld3(tx->dx), dy = ld3(tx->dy);
        for (uint32_t r = 0u; r < rays; r++) {
            float jx = rnd(c, key + r, 19u) - 0.5f;
            float jy = rnd(c, key + r, 23u) - 0.5f;
            v3 o = add(add(add(pos, scl(dx, jx)), scl(dy, jy)), scl(nrm, P->reach[2]));
            float u0 = ((float)r + 0.5f) / (float)rays + s1;
            float v0 = orc_radical(r) + s2;
            v3 d = cosine(u0 - floorf(u0), v0 - floorf(v0), &f);
            hit_t h = trace(c, o, d, P->reach[1], 0, key + r + 4000u);
            uint32_t *q = out + ((size_t)i * rays + r) * 2u;
            if (h.tri == NONE_U) {
                q[0] = fbits(d.y);
                q[1] = 3u << 30;
            } else if (h.back) {
                q[0] = 0u;
                q[1] = 2u << 30;
            } else {
                const lb_tri *tr = &c->tris[h.tri];
                uint32_t m = fbits(tr->v0[3]), ch = fbits(tr->e1[3]);
                if (ch == NONE_U) {
                    q[0] = 0u;
                    q[1] = (1u << 30) | m;
                } else {
                    v3 p = add(o, scl(d, h.t));
                    const float *r0 = &c->charts[ch * 8u], *r1 = &c->charts[ch * 8u + 4u];
                    float x = r0[0] * p.x + r0[1] * p.y + r0[2] * p.z + r0[3] * 1.0f;
                    float y = r1[0] * p.x + r1[1] * p.y + r1[2] * p.z + r1[3] * 1.0f;
                    int ix = orc_sat_floor(x, (int)P->atlas[0] - 1);
                    int iy = orc_sat_floor(y, (int)P->atlas[1] - 1);
                    q[0] = (uint32_t)iy * P->atlas[0] + (uint32_t)ix;
                    q[1] = m;
                }
            }
        }
    }
    return 0;
}

/* Each gather ray's direction (trace.wgsl:262), ray r of item i at (i * rays + r) * 3: what the
 * sky's L1 term of a ray that escapes reads (a record keeps only d.y). */
int orc_gather_dirs(const lb_params *params, const lb_texel *items, size_t n, uint32_t key_base,
                    float *out) {
    ctx_t c0;
    memset(&c0, 0, sizeof c0);
    c0.P = params;
    uint32_t rays = params->counts[2];
    for (size_t i = 0; i < n; i++) {
        const ctx_t *c = &c0;
        v3 nrm = ld3(items[i].normal);
        frame_t f = frame(nrm);
        uint32_t key = (key_base + (uint32_t)i) * 7919u;
        float s1 = rnd(c, key, 11u), s2 = rnd(c, key, 13u);
        for (uint32_t r = 0u; r < rays; r++) {
            float u0 = ((float)r + 0.5f) / (float)rays + s1;
            float v0 = orc_radical(r) + s2;
            v3 d = cosine(u0 - floorf(u0), v0 - floorf(v0), &f);
            float *q = out + (i * rays + r) * 3u;
            q[0] = d.x; q[1] = d.y; q[2] = d.z;
        }
    }
    return 0;
}

/* Each gather ray's FP32 term as gather_item (trace.wgsl:258-268) adds it: sky(d) on a miss,
 * zero on a back face, bounced() on a front hit; ray r of item i at (i * rays + r) * 3. The exact
 * sum of these terms is what every form's FP32 sum approximates: the yardstick of summation
 * error alone (the hits being equal). */
int orc_gather_terms(const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                     const float *materials, const float *charts, const lb_params *params,
                     const lb_texel *items, size_t n, uint32_t key_base, const float *light,
                     float *out) {
    static const float zero_light[4] = {0, 0, 0, 0};
    static const float zero_chart[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    lb_node tmp;
    ctx_t c0;
    memset(&c0, 0, sizeof c0);
    c0.nodes = empty_fix(nodes, n_nodes, &tmp); c0.tris = tris; c0.materials = materials;
    c0.charts = charts ? charts : zero_chart;
    c0.light = light ? light : zero_light;
    c0.P = params;
    uint32_t rays = params->counts[2];
    long nn = (long)n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < nn; i++) {
        const ctx_t *c = &c0;
        const lb_texel *tx = &items[i];
        const lb_params *P = c->P;
        v3 nrm = ld3(tx->normal);
        frame_t f = frame(nrm);
        uint32_t key = (key_base + (uint32_t)i) * 7919u;
        float s1 = rnd(c, key, 11u), s2 = rnd(c, key, 13u);
        v3 pos = ld3(tx->pos), dx = ld3(tx->dx), dy = ld3(tx->dy);
        for (uint32_t r = 0u; r < rays; r++) {
            float jx = rnd(c, key + r, 19u) - 0.5f;
            float jy = rnd(c, key + r, 23u) - 0.5f;
            v3 o = add(add(add(pos, scl(dx, jx)), scl(dy, jy)), scl(nrm, P->reach[2]));
            float u0 = ((float)r + 0.5f) / (float)rays + s1;
            float v0 = orc_radical(r) + s2;
            v3 d = cosine(u0 - floorf(u0), v0 - floorf(v0), &f);
            hit_t h = trace(c, o, d, P->reach[1], 0, key + r + 4000u);
            v3 t = h.tri == NONE_U ? sky(c, d) : (h.back ? mk(0.0f, 0.0f, 0.0f) : bounced(c, o, d, h));
            float *q = out + ((size_t)i * rays + r) * 3u;
            q[0] = t.x; q[1] = t.y; q[2] = t.z;
        }
    }
    return 0;
}

/* Each gather ray's traversal steps, inner-node visits plus triangle tests (trace.wgsl:83-130,
 * capped at 255), ray r of item i at i * rays + r: the rays of trace.wgsl:248-262 as gather_item
 * makes them, each through trace() with counters of its own. The TUNED gather's first bounce
 * records the same counts to rank each item's rays by; the reference has no such output. */
int orc_gather_steps(const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                     const float *materials, const lb_params *params, const lb_texel *items,
                     size_t n, uint32_t key_base, uint8_t *out) {
    lb_node tmp;
    ctx_t c0;
    memset(&c0, 0, sizeof c0);
    c0.nodes = empty_fix(nodes, n_nodes, &tmp); c0.tris = tris; c0.materials = materials;
    c0.P = params;
    uint32_t rays = params->counts[2];
    long nn = (long)n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < nn; i++) {
        ctx_t c = c0;
        orc_counters cnt;
        c.cnt = &cnt;
        const lb_texel *tx = &items[i];
        const lb_params *P = c.P;
        v3 nrm = ld3(tx->normal);
        frame_t f = frame(nrm);
        uint32_t key = (key_base + (uint32_t)i) * 7919u;
        float s1 = rnd(&c, key, 11u), s2 = rnd(&c, key, 13u);
        v3 pos = ld3(tx->pos), dx = ld3(tx->dx), dy = ld3(tx->dy);
        for (uint32_t r = 0u; r < rays; r++) {
            float jx = rnd(&c, key + r, 19u) - 0.5f;
            float jy = rnd(&c, key + r, 23u) - 0.5f;
            v3 o = add(add(add(pos, scl(dx, jx)), scl(dy, jy)), scl(nrm, P->reach[2]));
            float u0 = ((float)r + 0.5f) / (float)rays + s1;
            float v0 = orc_radical(r) + s2;
            v3 d = cosine(u0 - floorf(u0), v0 - floorf(v0), &f);
            memset(&cnt, 0, sizeof cnt);
            trace(&c, o, d, P->reach[1], 0, key + r + 4000u);
            uint64_t steps = cnt.inner_visits + cnt.tri_tests;
            out[(size_t)i * rays + r] = (uint8_t)(steps < 255u ? steps : 255u);
        }
    }
    return 0;
}

/* One ray through trace() (trace.wgsl:83-130): t, triangle, back. */
void orc_trace(const lb_node *nodes, size_t n_nodes, const lb_tri *tris, const float *materials,
               const lb_params *params, const float o[3], const float d[3], float tmax, int any,
               uint32_t ray, float *t_out, uint32_t *tri_out, int *back_out, orc_counters *cnt) {
    ctx_t c;
    lb_node tmp;
    memset(&c, 0, sizeof c);
    c.nodes = empty_fix(nodes, n_nodes, &tmp); c.tris = tris; c.materials = materials; c.P = params; c.cnt = cnt;
    hit_t h = trace(&c, ld3(o), ld3(d), tmax, any, ray);
    *t_out = h.t; *tri_out = h.tri; *back_out = h.back;
}

/* The same ray against every triangle in order, no BVH (the brute-force check of trace()). */
void orc_trace_brute(const lb_tri *tris, size_t n_tris, const float *materials,
                     const lb_params *params, const float o_[3], const float d_[3], float tmax,
                     uint32_t ray, float *t_out, uint32_t *tri_out, int *back_out) {
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.tris = tris; c.materials = materials; c.P = params;
    v3 o = ld3(o_), d = ld3(d_);
    hit_t hit = {tmax, NONE_U, 0};
    for (uint32_t i = 0; i < (uint32_t)n_tris; i++) {
        const lb_tri *tr = &tris[i];
        v3 e1 = ld3(tr->e1), e2 = ld3(tr->e2);
        v3 p = cross3(d, e2);
        float det = dot3(e1, p);
        if (fabsf(det) < 1e-12f) continue;
        float id = 1.0f / det;
        v3 s = sub(o, ld3(tr->v0));
        float u = dot3(s, p) * id;
        if (u < 0.0f || u > 1.0f) continue;
        v3 q = cross3(s, e1);
        float v = dot3(d, q) * id;
        if (v < 0.0f || u + v > 1.0f) continue;
        float t = dot3(e2, q) * id;
        if (t > 1e-4f && t < hit.t && stops(&c, i, ray)) {
            hit.t = t; hit.tri = i; hit.back = det < 0.0f;
        }
    }
    *t_out = hit.t; *tri_out = hit.tri; *back_out = hit.back;
}
