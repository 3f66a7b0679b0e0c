// One triangle of a tile as raster.rs:45-70 sets it up: its corners in the tile's texel space,
// the determinant of the barycentric map, its world corners, face normal and texel frame, its
// chart, and the texel box it may claim (inclusive, clamped to the tile).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/lightbaker.h"

namespace lbk {

struct RasterTri {
    float ua[2], ub[2], uc[2];
    float det;
    float pa[3], pb[3], pc[3];
    float n[3], dx[3], dy[3];
    uint32_t chart;
    uint32_t tile;      // the tile's place in the list rasterized
    int32_t lo[2], hi[2];
    uint32_t pad;
};
static_assert(sizeof(RasterTri) == 128, "RasterTri is 32 words");

struct Layout;
// The triangles of `tiles` (layout indices) in raster.rs's order: tile by tile, each tile's charts
// in order, each chart's triangles in order; triangles with a degenerate map are left out.
void raster_records(const lb_scene_desc &s, const Layout &L, const std::vector<uint32_t> &tiles,
                    std::vector<RasterTri> &out);
// The texels of the tiles whose table (first texel, w, h, atlas corner) is `tiles`, into `out`;
// owner: n_texels words of scratch.
hipError_t raster_tiles(const RasterTri *rec, uint32_t n_rec, const uint4 *tiles,
                        const uint32_t *tile_of, uint32_t n_texels, uint32_t *owner,
                        lb_texel *out, hipStream_t stream);
// Each texel's covered flag, chart (0xffffffff where uncovered) and normal (w 0).
hipError_t raster_derive(const lb_texel *tex, uint32_t n, uint32_t *chart, float4 *normal,
                         uint8_t *covered, hipStream_t stream);

}  // namespace lbk
