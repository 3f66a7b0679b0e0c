// The device texel passes between the ray passes (texel_passes.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lightbaker.h"

namespace lbk {

// output.rs:29-40: how a texel got its light
enum : uint8_t {
    VALIDITY_NONE = 0,
    VALIDITY_CENTRE = 1,
    VALIDITY_OFFSET = 2,
    VALIDITY_PUSHED = 3,
    VALIDITY_FILLED = 4,
    VALIDITY_VOLUME = 5,
};

// The dirty tiles' texels back to back, and their per-texel state (bake.rs:61-92 TileLight/Work).
struct TexelView {
    uint32_t n;              // texels
    const uint32_t *tile_of; // texel -> tile
    const uint4 *tiles;      // per tile: first texel, w, h, atlas corner (x | y << 16)
    lb_texel *texels;        // the rasterized texels (sample points move on decode)
    const uint32_t *chart;   // chart, ~0 where empty
    const float4 *normal;    // face normal (0 where empty)
    uint8_t *valid;          // covered and not inside a solid
    uint8_t *known;          // 2 own light, 1 filled, 0 none
    uint8_t *state;          // VALIDITY_*
    uint32_t *hard;          // sun bits (cell s % g, s / g)
    float *soft;             // sun share
    float *ao;
    float *field;
    float4 *direct;          // rgb
    float4 *indirect;        // rgb
};

// filter.rs:13-16: texels i and j lie on one surface (same chart, normals within ~25 degrees)
__device__ __forceinline__ bool same_surface(const TexelView &V, uint32_t i, uint32_t j) {
    float4 a = V.normal[i], b = V.normal[j];
    return V.chart[i] == V.chart[j] && a.x * b.x + a.y * b.y + a.z * b.z >= 0.9f;
}

// A texel's tile and its place in it.
struct Where {
    uint32_t tile, off, w, h, x, y;
};
__device__ __forceinline__ Where where(const TexelView &V, uint32_t g) {
    Where r;
    r.tile = V.tile_of[g];
    uint4 t = V.tiles[r.tile];  // off, w, h, atlas corner (x | y << 16)
    r.off = t.x;
    r.w = t.y;
    r.h = t.z;
    uint32_t i = g - r.off;
    r.x = i % r.w;
    r.y = i / r.w;
    return r;
}

struct DecodeArgs {
    float inside_share, ao_strength, ao_contrast, ao_floor;
};

void tp_decode(const TexelView &V, const float4 *res, const uint32_t *cov, uint32_t n_cov,
               const DecodeArgs &a, hipStream_t s);
void tp_smooth(const TexelView &V, int channels, const float *in, float *out, const uint8_t *valid,
               int r, const float *w, hipStream_t s);
void tp_sun_field(const TexelView &V, const uint8_t *covered, uint32_t g, float radius,
                  hipStream_t s);
void tp_direct_light(const TexelView &V, const float *soft, float3 irr, float3 sun, hipStream_t s);
void tp_fill_ring(const TexelView &V, int channels, float *v, const uint8_t *was, uint8_t *now,
                  hipStream_t s);
// the bounce fill over 3 channels at stride 4: ring r filing each texel's ring (RING_NEVER: never
// filled), and ring r over the texels filed for it
constexpr uint8_t RING_NEVER = 0xff;
void tp_fill_ring_file(const TexelView &V, float *v, const uint8_t *was, uint8_t *now,
                       uint8_t *ring, uint32_t r, hipStream_t s);
void tp_fill_ring_filed(const TexelView &V, float *v, const uint8_t *ring, uint32_t r,
                        hipStream_t s);
// one ring of the occlusion (1 channel), sun field (1) and direct light (3, stride 4) together
void tp_fill_ring_direct(const TexelView &V, float *ao, float *field, float *direct,
                         const uint8_t *was, uint8_t *now, hipStream_t s);
void tp_known(const TexelView &V, const uint8_t *filled, hipStream_t s);
void tp_scatter(const float4 *res, const uint32_t *idx, uint32_t n_items, float4 *e, hipStream_t s);
void tp_write_light(const TexelView &V, float4 *light, uint32_t atlas_w, hipStream_t s);
void tp_collect(const lb_texel *texels, const uint32_t *idx, uint32_t n, lb_texel *items,
                hipStream_t s);
void tp_fill_f32(float *p, float v, uint32_t n, hipStream_t s);
// bake.rs:286-299 on the device: flags (u8 per texel), then the flagged texels' indices in order
// into `out` and their count into `total`; `counts` holds tp_gather_blocks(n) words
void tp_gather_set(const TexelView &V, uint32_t ds, uint8_t *flag, uint32_t *counts,
                   uint32_t *total, uint32_t *out, hipStream_t s);
uint32_t tp_gather_blocks(uint32_t n);
// The indices of the flagged entries of flag[0..n), in order, into `out`, their count into
// `total`; `counts` holds tp_gather_blocks(n) words
void tp_compact(const uint8_t *flag, uint32_t n, uint32_t *counts, uint32_t *total, uint32_t *out,
                hipStream_t s);
void tp_copy_f32(const float *src, float *dst, uint32_t n, hipStream_t s);
void tp_copy_u8(const uint8_t *src, uint8_t *dst, uint32_t n, hipStream_t s);

}  // namespace lbk
