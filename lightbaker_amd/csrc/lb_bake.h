// Host-side bake substrate: the reference's Rust host (scene, layout, raster, BVH, texel passes,
// bake driver) restated in C++. Scalar FP32 in glam's operation order (glam's Vec2/Vec3 are per
// component), built with -ffp-contract=off.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lightbaker.h"

namespace lbk {

struct V2 {
    float x, y;
};
struct V3 {
    float x, y, z;
    float operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glam's cross: (y*z' - y'*z, z*x' - z'*x, x*y' - x'*y); products commute exactly
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline V3 vmin(V3 a, V3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
// glam try_normalize: v * (1 / |v|) when that reciprocal is finite and positive
inline bool try_normalize(V3 v, V3 *out) {
    float r = 1.0f / std::sqrt(dot(v, v));
    if (!std::isfinite(r) || !(r > 0.0f)) return false;
    *out = v * r;
    return true;
}

// bvh.rs:50-106 over non-indexed triangles (9 floats each) with their material and chart, on
// up to `threads` host threads (0: the host's, at most 16); the same nodes whatever the count.
void build_bvh(const float *positions, size_t n_tris, const uint32_t *material,
               const uint32_t *chart, const uint8_t *casts, std::vector<lb_node> &nodes,
               std::vector<lb_tri> &tris, int threads = 0);

// A scene as the bake reads it (scene.rs:17-128), borrowed from the caller.
struct SceneRef {
    const lb_scene_desc *d;
    size_t n_tris() const;
    V3 vert(uint32_t mesh, uint32_t tri, int k) const {
        const float *p = d->meshes[mesh].positions + ((size_t)tri * 3 + k) * 3;
        return {p[0], p[1], p[2]};
    }
};

struct Chart {
    uint32_t mesh;
    std::vector<uint32_t> tris;  // the mesh's triangles in it, in mesh order
    V3 normal, u, v;
    V2 min, max;
    int32_t tile_x, tile_z;
    float density = 0.0f;  // texels a metre: its tile's (lb_scene_desc.tile_density), else the scene's
    bool turned = false;
    uint32_t w = 0, h = 0, x = 0, y = 0;
};

struct Tile {
    int32_t kx, kz;
    std::vector<uint32_t> charts;
    uint32_t w = 0, h = 0, x = 0, y = 0;
};

struct Layout {
    std::vector<Chart> charts;
    std::vector<Tile> tiles;  // in key order
    uint32_t width = 0, height = 0;
    float density = 0.0f;  // the scene's texels_per_meter (each chart keeps its tile's)
    uint32_t padding = 0;
    std::vector<std::vector<uint32_t>> tri_chart;  // each mesh's triangles' chart (~0u: none)
    std::vector<uint32_t> chart_tile;              // each chart's tile
};

// layout.rs:89-115; false with a message on a tile wider than the atlas (layout.rs:247-249)
bool build_layout(const lb_scene_desc &s, Layout &out, std::string *err);
// layout.rs:70-74, at the chart's density
V2 chart_local(const Chart &c, V3 p, uint32_t padding);
// layout.rs:77-85: two rows of (coefficients, constant), at the chart's density
void chart_affine(const Chart &c, const Tile &t, uint32_t padding, float out[8]);
// the scene's per-tile densities: false with a message on a bad or repeated entry
bool check_tile_density(const lb_scene_desc &s, std::string *err);
V3 face_normal(const lb_scene_desc &s, uint32_t mesh, uint32_t t);
// raster.rs:39-100: the tile's texels, row by row in its w x h rectangle
void rasterize(const lb_scene_desc &s, const Layout &l, const Tile &t, std::vector<lb_texel> &out);

}  // namespace lbk

// the C-ABI's layout handle (lb_layout_build)
struct lb_layout {
    lbk::Layout L;
};
