// What a gather ray's light depends on, recorded once (trace.wgsl:256-268): the kind of hit, the
// material and the atlas texel it reads, or the sky direction's y. Two u32 a ray, ray r of gather
// item g at r * items + g: (word, tag), tag = kind << 30 | material; word the atlas texel
// (REC_ATLAS) or the bits of d.y (REC_MISS). Replaying records with an atlas gives the gather's
// result for that atlas without tracing (every bounce traces the same rays: the same keys and
// sample points).
#pragma once

#include <cstring>

#include "lb_device.h"

namespace lbk {

enum : uint32_t { REC_ATLAS = 0u, REC_FALLBACK = 1u, REC_BACK = 2u, REC_MISS = 3u };

// the sky's colours and the fallback irradiance (trace.wgsl:63-66, Params.fallback), the sky's
// L1 term, and what a replay needs to remake a sky ray's direction (the L1 term reads all of it;
// a record keeps only d.y): the gather items and the Params' seed
struct SkyP {
    f3 zenith, horizon, ground, fallback;
    SkySH sh;
    const lb_texel *items;
    uint32_t seed;
};

// the sky's L1 term's irradiance on an upward face, per unit coefficient: pi 0.282095 (c0) and
// 2 pi / 3 0.488603 (c2, L1 along y), FP32; the fallback irradiance adds c0 and c2 times these
constexpr float SH_FALLBACK_C0 = (float)(3.14159265358979323846 * 0.282095);
constexpr float SH_FALLBACK_C2 = (float)(2.0 * 3.14159265358979323846 / 3.0 * 0.488603);

// gather ray r of item i (key base 0): its direction as trace.wgsl:248-262 makes it
__device__ __forceinline__ f3 gather_dir(const lb_texel *items, uint32_t i, uint32_t r,
                                         uint32_t rays, uint32_t seed) {
    const float4 *tp = reinterpret_cast<const float4 *>(items + i);
    frame_t f = frame(xyz(tp[1]));
    uint32_t key = i * 7919u;
    float s1 = rnd(key, 11u, seed), s2 = rnd(key, 13u, seed);
    return cosine(fract(((float)r + 0.5f) / (float)rays + s1), fract(radical(r) + s2), f);
}

// the sky a replayed ray r of item i sees, its record's word holding d.y (trace.wgsl:63-66 and
// the L1 term)
__device__ __forceinline__ f3 replay_sky(const SkyP &K, uint32_t word, uint32_t i, uint32_t r,
                                         uint32_t rays) {
    float y = bitsf(word);
    f3 s = y < 0.0f ? K.ground : add(scl(K.horizon, 1.0f - y), scl(K.zenith, y));
    if (K.sh.on) {
        f3 d = gather_dir(K.items, i, r, rays, K.seed);
        s = add(s, mk(sh_channel(K.sh.r, d), sh_channel(K.sh.g, d), sh_channel(K.sh.b, d)));
    }
    return s;
}

// the atlas texel bounced() (trace.wgsl:137-145) reads for a hit at o + d t on chart c
__device__ __forceinline__ uint32_t hit_texel(const SceneView &S, const lb_params &P, f3 o, f3 d,
                                              float t, uint32_t c) {
    f3 p = add(o, scl(d, t));
    float4 r0 = S.charts[c * 2u], r1 = S.charts[c * 2u + 1u];
    float x = r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w * 1.0f;
    float y = r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w * 1.0f;
    float fx = floorf(x), fy = floorf(y);
    int hx = (int)P.atlas[0] - 1, hy = (int)P.atlas[1] - 1;
    int ix = !(fx >= 0.0f) ? 0 : (fx >= (float)hx ? hx : (int)fx);
    int iy = !(fy >= 0.0f) ? 0 : (fy >= (float)hy ? hy : (int)fy);
    return (uint32_t)iy * P.atlas[0] + (uint32_t)ix;
}

// the record of a gather ray o + t d whose trace ended with (tri, back) at t
__device__ __forceinline__ uint2 gather_record(const SceneView &S, const lb_params &P, f3 o, f3 d,
                                               uint32_t tri, bool back, float t) {
    if (tri == NONE) return make_uint2(fbits(d.y), REC_MISS << 30);
    if (back) return make_uint2(0u, REC_BACK << 30);
    uint32_t m = fbits(S.tris[tri * 3u].w), c = fbits(S.tris[tri * 3u + 1u].w);
    return c == NONE ? make_uint2(0u, (REC_FALLBACK << 30) | m)
                     : make_uint2(hit_texel(S, P, o, d, t, c), (REC_ATLAS << 30) | m);
}

// SkyP of a pass's Params, the scene's sky term and the gather items
inline SkyP sky_of(const lb_params &P, const SkySH &sh, const lb_texel *items) {
    SkyP K;
    K.zenith = f3{P.zenith[0], P.zenith[1], P.zenith[2]};
    K.horizon = f3{P.horizon[0], P.horizon[1], P.horizon[2]};
    K.ground = f3{P.ground[0], P.ground[1], P.ground[2]};
    K.fallback = f3{P.fallback[0], P.fallback[1], P.fallback[2]};
    K.sh = sh;
    K.items = items;
    uint32_t seed;
    std::memcpy(&seed, &P.reach[3], 4);
    K.seed = seed;
    return K;
}

// Every gather ray's record, one lane a ray traced as the PLAIN form traces it (item keys from 0).
hipError_t record_gather(const SceneView &S, const lb_params &P, const lb_texel *items,
                         uint32_t n, uint32_t rays, uint2 *rec, hipStream_t stream);

// The gather's result from records and an atlas (trace.wgsl:256-269 in ray order), albedo per
// material in .xyz; out: (E rgb, back share) per item.
hipError_t replay_gather(const uint2 *rec, uint32_t n, uint32_t rays, const float4 *albedo,
                         const float4 *light, const SkyP &K, float4 *out, hipStream_t stream);

}  // namespace lbk
