// Device-side math of the ray passes (crates/lightbaker/src/gpu/trace.wgsl), written for gfx950.
//
// Every function keeps WGSL's evaluation order so that a kernel built with -ffp-contract=off
// reproduces the reference's FP32 results bit for bit: no fused multiply-add, correctly rounded
// division and sqrt, round-to-nearest-even u32->f32. Where WGSL leaves the implementation free
// (sin/cos, normalize, mix, an f32->i32 conversion out of range) one definition is fixed here:
//   sin/cos   Cody-Waite reduction by pi/2 (three-part split, as in Cephes' sinf) and Taylor
//             polynomials of degree 11 / 10 on [-pi/4, pi/4];
//   normalize v / sqrt(dot(v, v));   mix  a * (1 - t) + b * t;   i32(f32)  saturating.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lightbaker.h"

namespace lbk {

constexpr float PI = 3.14159265f;
constexpr float FAR = 1e30f;
constexpr uint32_t NONE = 0xffffffffu;
constexpr float PUSH = 0.03f;
constexpr int STACK = 48;  // trace.wgsl:86; a push beyond it is dropped (trace.wgsl:120)

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 normalize3(f3 v) {
    float l = sqrtf(dot3(v, v));
    return mk(v.x / l, v.y / l, v.z / l);
}
// WGSL's min/max. No NaN reaches them here (finite scene coordinates, |1/d| <= 1e12), and the
// sign of a zero they might pick never changes a comparison or a stored value, so the hardware's
// v_min/v_max (fused into v_min3/v_max3) give the reference's results.
__device__ __forceinline__ float vmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }

// sin and cos of x (|x| < 2^15 * pi/2): x = q pi/2 + r, then Taylor in r, quadrant by q mod 4.
__device__ __forceinline__ void sincos_lb(float x, float *s, float *c) {
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
    float a = (qi & 1) ? cr : sr;
    float b = (qi & 1) ? sr : cr;
    *s = (qi & 2) ? -a : a;
    *c = ((qi + 1) & 2) ? -b : b;
}

// trace.wgsl:41-45
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
// trace.wgsl:46; `seed` is bitcast<u32>(P.reach.w)
__device__ __forceinline__ float rnd(uint32_t a, uint32_t b, uint32_t seed) {
    return (float)pcg(a ^ pcg(b + seed)) * (1.0f / 4294967296.0f);
}
// trace.wgsl:49
__device__ __forceinline__ float radical(uint32_t i) {
    return (float)__builtin_bitreverse32(i) * (1.0f / 4294967296.0f);
}

// trace.wgsl:51-55: columns t, cross(n, t), n
struct frame_t {
    f3 t, b, n;
};
__device__ __forceinline__ frame_t frame(f3 n) {
    f3 a = fabsf(n.y) > 0.9f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
    frame_t f;
    f.t = normalize3(cross3(a, n));
    f.b = cross3(n, f.t);
    f.n = n;
    return f;
}
__device__ __forceinline__ f3 fmul(const frame_t &f, f3 v) {
    return add(add(scl(f.t, v.x), scl(f.b, v.y)), scl(f.n, v.z));
}

// trace.wgsl:57-61
__device__ __forceinline__ f3 cosine(float u, float v, const frame_t &f) {
    float r = sqrtf(u);
    float a = v * 2.0f * PI;
    float sa, ca;
    sincos_lb(a, &sa, &ca);
    return normalize3(fmul(f, mk(r * ca, r * sa, sqrtf(vmax(0.0f, 1.0f - u)))));
}

// The sky's L1 spherical-harmonic radiance, added to the reference's gradient sky (no
// counterpart in the reference, whose sky is the gradient alone): per channel
// 0.282095 c0 + 0.488603 ((c1 d.x + c2 d.y) + c3 d.z), the coefficients in the order of the
// volume's probes (L0, then L1 along x, y, z; volume.rs:25-31). With `on` clear (every
// coefficient zero) the sky is the reference's, bit for bit.
struct SkySH {
    float4 r, g, b;  // (c0, c1, c2, c3) of each channel
    uint32_t on;
};
__device__ __forceinline__ float sh_channel(float4 c, f3 d) {
    return 0.282095f * c.x + 0.488603f * ((c.y * d.x + c.z * d.y) + c.w * d.z);
}

// trace.wgsl:63-66, plus the L1 term
__device__ __forceinline__ f3 sky(const SkySH &K, const lb_params &P, f3 d) {
    f3 s;
    if (d.y < 0.0f) {
        s = mk(P.ground[0], P.ground[1], P.ground[2]);
    } else {
        float w = 1.0f - d.y;
        s = add(scl(mk(P.horizon[0], P.horizon[1], P.horizon[2]), w),
                scl(mk(P.zenith[0], P.zenith[1], P.zenith[2]), d.y));
    }
    if (K.on) s = add(s, mk(sh_channel(K.r, d), sh_channel(K.g, d), sh_channel(K.b, d)));
    return s;
}

// trace.wgsl:74-80 over a box given by its corners
__device__ __forceinline__ float slab(f3 o, f3 inv, f3 mn, f3 mx, float tmax) {
    f3 a = mul(sub(mn, o), inv);
    f3 b = mul(sub(mx, o), inv);
    float lo = vmax(vmax(vmin(a.x, b.x), vmin(a.y, b.y)), vmin(a.z, b.z));
    float hi = vmin(vmin(vmax(a.x, b.x), vmax(a.y, b.y)), vmax(a.z, b.z));
    float l0 = vmax(lo, 0.0f);
    return (hi >= l0 && lo < tmax) ? l0 : FAR;
}

// slab() as its two parts: the entry max(lo, 0) and whether the box counts as hit (slab() is
// the entry when hit, FAR when not)
// With tmax > 0 (a ray's reach, or a hit's t > 1e-4), lo < tmax is l0 < tmax: when lo < 0 both
// hold. The entry is taken in one chain of maxima from 0.
__device__ __forceinline__ float slab_hit(f3 o, f3 inv, f3 mn, f3 mx, float tmax, bool *hit) {
    f3 a = mul(sub(mn, o), inv);
    f3 b = mul(sub(mx, o), inv);
    float l0 = vmax(vmax(vmax(0.0f, vmin(a.x, b.x)), vmin(a.y, b.y)), vmin(a.z, b.z));
    float hi = vmin(vmin(vmax(a.x, b.x), vmax(a.y, b.y)), vmax(a.z, b.z));
    *hit = hi >= l0 && l0 < tmax;
    return l0;
}

// trace.wgsl:85
__device__ __forceinline__ f3 inverse_dir(f3 d) {
    f3 dd = mk(fabsf(d.x) < 1e-12f ? 1e-12f : d.x, fabsf(d.y) < 1e-12f ? 1e-12f : d.y,
               fabsf(d.z) < 1e-12f ? 1e-12f : d.z);
    return mk(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
}

struct Hit {
    float t;
    uint32_t tri;
    bool back;
};

// The scene as the kernels read it (the bindings of trace.wgsl:24-31).
struct SceneView {
    const float4 *nodes;      // lb_node as two float4
    const float4 *tris;       // lb_tri as three float4
    const float4 *materials;  // albedo rgb, opacity
    const float4 *charts;     // two rows a chart
    const float4 *light;      // rgba per atlas texel
    const float4 *wide;       // TUNED: 4 float4 per inner node (both children's boxes, refs)
    const uint4 *quant;       // TUNED: 2 uint4 per inner node (quantized children, refs)
    const uint2 *leaves;      // TUNED: (first triangle, count) at a leaf's node index
    const float4 *tris_t;     // TUNED: the triangles with their material's opacity in e2.w
    uint32_t n_nodes, n_tris;
    uint32_t root;            // TUNED: the root as a child reference (bit 31: a leaf)
    SkySH sh;                 // the sky's L1 term (lb_scene_set_sky)
};

// trace.wgsl:69-72
__device__ __forceinline__ bool stops(const SceneView &S, const lb_params &P, uint32_t tri,
                                      float4 v0, uint32_t ray) {
    float op = S.materials[fbits(v0.w)].w;
    return op >= 1.0f || rnd(tri * 9781u, ray, fbits(P.reach[3])) < op;
}

// Moller-Trumbore as trace.wgsl:93-104; returns t or -1 when rejected, det in *det_out
__device__ __forceinline__ float mt(f3 o, f3 d, float4 v0, float4 e1v, float4 e2v,
                                    float *det_out) {
    f3 e1 = xyz(e1v), e2 = xyz(e2v);
    f3 p = cross3(d, e2);
    float det = dot3(e1, p);
    *det_out = det;
    if (fabsf(det) < 1e-12f) return -1.0f;
    float id = 1.0f / det;
    f3 s = sub(o, xyz(v0));
    float u = dot3(s, p) * id;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    f3 q = cross3(s, e1);
    float v = dot3(d, q) * id;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    return dot3(e2, q) * id;
}

__device__ __forceinline__ float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// The watertight ray-triangle test of Woop, Benthin and Wald ("Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013), without its double-precision fallback: the ray's dominant axis
// becomes z, the triangle is sheared so the ray runs along +z, and the signs of the projected
// triangle's edge functions U, V, W decide the hit, so a ray through a shared edge meets one of
// its two triangles. The vertices are v0, v0 + e1, v0 + e2 (the record keeps edges). Returns t =
// T / det or -1; *det_out is the reference's orientation e1 . (d x e2) (a back face below 0).
__device__ __forceinline__ float wt(f3 o, f3 d, float4 v0v, float4 e1v, float4 e2v,
                                    float *det_out) {
    f3 e1 = xyz(e1v), e2 = xyz(e2v), v0 = xyz(v0v);
    *det_out = dot3(e1, cross3(d, e2));
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.0f) {
        int k = kx;
        kx = ky;
        ky = k;
    }
    float sx = comp(d, kx) / dz, sy = comp(d, ky) / dz, sz = 1.0f / dz;
    f3 A = sub(v0, o), B = sub(add(v0, e1), o), C = sub(add(v0, e2), o);
    float Ax = comp(A, kx) - sx * comp(A, kz), Ay = comp(A, ky) - sy * comp(A, kz);
    float Bx = comp(B, kx) - sx * comp(B, kz), By = comp(B, ky) - sy * comp(B, kz);
    float Cx = comp(C, kx) - sx * comp(C, kz), Cy = comp(C, ky) - sy * comp(C, kz);
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return -1.0f;
    float det = U + V + W;
    if (det == 0.0f) return -1.0f;
    float T = U * (sz * comp(A, kz)) + V * (sz * comp(B, kz)) + W * (sz * comp(C, kz));
    return T / det;
}

// Moller-Trumbore with every product of the three edge vectors taken before any test, so the
// triangle's three loads are consumed together (one memory round trip); the same values and
// rejections as mt().
__device__ __forceinline__ float mt_all(f3 o, f3 d, float4 v0, float4 e1v, float4 e2v,
                                        float *det_out) {
    f3 e1 = xyz(e1v), e2 = xyz(e2v);
    f3 p = cross3(d, e2);
    float det = dot3(e1, p);
    f3 s = sub(o, xyz(v0));
    float du = dot3(s, p);
    f3 q = cross3(s, e1);
    float dv = dot3(d, q);
    float dt = dot3(e2, q);
    *det_out = det;
    if (fabsf(det) < 1e-12f) return -1.0f;
    float id = 1.0f / det;
    float u = du * id;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    float v = dv * id;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    return dt * id;
}

// trace.wgsl:133-148
__device__ __forceinline__ f3 bounced(const SceneView &S, const lb_params &P, f3 o, f3 d,
                                      const Hit &h) {
    float4 v0 = S.tris[h.tri * 3u];
    float4 e1 = S.tris[h.tri * 3u + 1u];
    float4 m = S.materials[fbits(v0.w)];
    uint32_t c = fbits(e1.w);
    f3 e = mk(P.fallback[0], P.fallback[1], P.fallback[2]);
    if (c != NONE) {
        f3 p = add(o, scl(d, h.t));
        float4 r0 = S.charts[c * 2u], r1 = S.charts[c * 2u + 1u];
        float x = r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w * 1.0f;
        float y = r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w * 1.0f;
        float fx = floorf(x), fy = floorf(y);
        int hx = (int)P.atlas[0] - 1, hy = (int)P.atlas[1] - 1;
        int ix = !(fx >= 0.0f) ? 0 : (fx >= (float)hx ? hx : (int)fx);
        int iy = !(fy >= 0.0f) ? 0 : (fy >= (float)hy ? hy : (int)fy);
        float4 l = S.light[(uint32_t)iy * P.atlas[0] + (uint32_t)ix];
        if (l.w > 0.5f) e = mk(l.x, l.y, l.z);
    }
    return mul(mk(m.x / PI, m.y / PI, m.z / PI), e);
}

// trace.wgsl:152-156
__device__ __forceinline__ void offset(uint32_t q, float *cx, float *cy) {
    if (q <= 4u) {
        *cx = (q & 1u) == 1u ? 0.3f : -0.3f;
        *cy = q > 2u ? 0.3f : -0.3f;
        return;
    }
    *cx = q == 5u ? -0.45f : (q == 6u ? 0.45f : 0.0f);
    *cy = q == 7u ? -0.45f : (q == 8u ? 0.45f : 0.0f);
}

// Per-thread traversal counters, flushed to the lb_stats block at the end of a kernel.
template <typename T>
struct CountsT {
    T rays, inner, tris, mats, gathers, leaves;
};
// the launch's traversal counters (flushed to lb_stats)
using Counts = CountsT<uint64_t>;
// a lane's own step count only (the counting gather's ray costs), never flushed: what is not
// read is not computed
using Counts32 = CountsT<uint32_t>;

__device__ __forceinline__ void flush_counts(const Counts &c, unsigned long long *dst) {
    atomicAdd(dst + 0, (unsigned long long)c.rays);
    atomicAdd(dst + 1, (unsigned long long)c.inner);
    atomicAdd(dst + 2, (unsigned long long)c.tris);
    atomicAdd(dst + 3, (unsigned long long)c.mats);
    atomicAdd(dst + 4, (unsigned long long)c.gathers);
    atomicAdd(dst + 5, (unsigned long long)c.leaves);
}

}  // namespace lbk
