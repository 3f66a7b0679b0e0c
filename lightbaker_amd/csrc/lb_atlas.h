// The bake's tail on the device (bake.rs:337-421): every tile's TileLight, the atlas assembled
// from it, and the output codecs (encode.rs, output.rs). Launchers queue on `s` and return the
// launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbk {

constexpr uint32_t ATLAS_NONE = 0xffffffffu;

// One tile of the layout: its first texel in the all-tile order (tiles in layout order, texels
// row-major), its size, its corner in the atlas, and whether this bake traces it.
struct AtlasTile {
    uint32_t first, w, h, x, y, dirty;
};

// The atlas' 4 x 4 blocks (tiles are 4-aligned: layout.rs:18,216,236,260) -> the tile holding
// each, or ATLAS_NONE.
struct AtlasMap {
    const AtlasTile *tiles;
    const uint32_t *block_tile;
    uint32_t aw, ah, bw;  // atlas size; blocks a row
};

// Every tile's TileLight (bake.rs:61-69), tiles in layout order.
struct TileLightView {
    uint8_t *known, *state;
    float *field, *ao, *direct, *indirect;  // direct and indirect: rgb a texel
};

// The baked tiles' texel state after the passes, and how it maps into TileLightView.
struct BakedView {
    uint32_t n;                       // texels of the baked tiles
    const uint32_t *tile_of;          // texel -> its baked tile (0..baked-1)
    const int64_t *dst_base;          // baked tile -> all-tile index minus baked-list index
    const uint8_t *covered, *known, *state;
    const float *field, *ao;
    const float4 *direct, *indirect;  // rgb (w unused)
    const float4 *pos, *normal;       // texel records' pos (stride 4 float4) and normals
};

// The volume the texels without light of their own take theirs from (volume.rs:125-147).
struct VolumeView {
    const float4 *probes;  // 4 float4 a probe, filled; null: no volume
    uint32_t nx, ny, nz;
    float ox, oy, oz, step;
    float3 sun_dir, sun_irr;
};

// bake.rs:337-362: the baked tiles' TileLight, lit from the volume where they have no light of
// their own, written into the all-tile arrays.
hipError_t atlas_tile_light(const BakedView &B, const VolumeView &V, const TileLightView &T,
                            hipStream_t s);
// bake.rs:435-444 over the clean tiles: the light atlas a rebake's first bounce starts from.
hipError_t atlas_clean_light(const AtlasMap &M, const TileLightView &T, float4 *light,
                             hipStream_t s);
// bake.rs:364-421: direct (2 f32) and validity an atlas texel, the indirect atlas (3 f32 a
// smaller texel, ds in {1, 2, 4}) with its two rings of fill; `tmp` holds iw * ih * 3 floats and
// `has0`, `has1` iw * ih bytes each.
hipError_t atlas_assemble(const AtlasMap &M, const TileLightView &T, uint32_t ds, float2 *direct,
                          uint8_t *validity, float *indirect, float *tmp, uint8_t *has0,
                          uint8_t *has1, hipStream_t s);

// output.rs:101 + encode.rs:37-50 over a float atlas (2 f32 a texel), w and h multiples of 4
hipError_t encode_bc5_f32(const float2 *rg, uint32_t w, uint32_t h, uint8_t *out, hipStream_t s);
// encode.rs:37-50 over byte pairs
hipError_t encode_bc5_u8(const uint8_t *rg, uint32_t w, uint32_t h, uint8_t *out, hipStream_t s);
// encode.rs:53-69 over n rgb triples
hipError_t encode_rgb9e5(const float *rgb, size_t n, uint32_t *out, hipStream_t s);
// output.rs:108,165-181: `scale_bits` (one word of device memory) receives the scale's bits
hipError_t encode_volume(const float4 *probes, size_t n, uint8_t *out, uint32_t *scale_bits,
                         hipStream_t s);

}  // namespace lbk
