// The C-ABI of the host substrate (include/lightbaker.h, second half).
#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "../../include/lightbaker.h"
#include "lb_bake.h"

namespace lbk {
int fail(const char *fmt, ...);
}

struct lb_bvh {
    std::vector<lb_node> nodes;
    std::vector<lb_tri> tris;
};

extern "C" {

int lb_bvh_build(const float *positions, size_t n_tris, const uint32_t *material,
                 const uint32_t *chart, const uint8_t *casts, size_t n_materials, lb_bvh **out) {
    return lb_bvh_build_threads(positions, n_tris, material, chart, casts, n_materials, 0, out);
}

int lb_bvh_build_threads(const float *positions, size_t n_tris, const uint32_t *material,
                         const uint32_t *chart, const uint8_t *casts, size_t n_materials,
                         int threads, lb_bvh **out) {
    if (!out) return lbk::fail("lb_bvh_build: null handle");
    if (threads < 0) return lbk::fail("lb_bvh_build: %d threads", threads);
    *out = nullptr;
    if (n_tris && (!positions || !material || !chart)) return lbk::fail("lb_bvh_build: null input");
    if (!casts && n_materials) return lbk::fail("lb_bvh_build: null casts");
    for (size_t t = 0; t < n_tris; t++)
        if (material[t] >= n_materials)
            return lbk::fail("lb_bvh_build: triangle %zu: material %u of %zu", t, material[t],
                             n_materials);
    auto *b = new (std::nothrow) lb_bvh();
    if (!b) return lbk::fail("lb_bvh_build: out of memory");
    lbk::build_bvh(positions, n_tris, material, chart, casts, b->nodes, b->tris, threads);
    *out = b;
    return 0;
}

size_t lb_bvh_nodes(const lb_bvh *bvh, const lb_node **nodes) {
    if (!bvh) return 0;
    if (nodes) *nodes = bvh->nodes.data();
    return bvh->nodes.size();
}

size_t lb_bvh_tris(const lb_bvh *bvh, const lb_tri **tris) {
    if (!bvh) return 0;
    if (tris) *tris = bvh->tris.data();
    return bvh->tris.size();
}

void lb_bvh_destroy(lb_bvh *bvh) { delete bvh; }

static int check_scene(const lb_scene_desc *s) {
    if (!s) return lbk::fail("layout: null scene");
    if (!s->materials || !s->n_materials) return lbk::fail("scene: no materials");
    for (size_t m = 0; m < s->n_meshes; m++) {
        if (s->meshes[m].material >= s->n_materials)
            return lbk::fail("mesh %zu: no material %u", m, s->meshes[m].material);
        if (s->meshes[m].triangles && !s->meshes[m].positions)
            return lbk::fail("mesh %zu: no positions", m);
    }
    if (!(s->texels_per_meter > 0.0f) || !(s->tile_size > 0.0f))
        return lbk::fail("lightmap: texelsPerMeter and tileSize must be positive");
    return 0;
}

int lb_layout_build(const lb_scene_desc *scene, lb_layout **out) {
    if (!out) return lbk::fail("lb_layout_build: null handle");
    *out = nullptr;
    if (check_scene(scene)) return 1;
    auto *l = new (std::nothrow) lb_layout();
    if (!l) return lbk::fail("lb_layout_build: out of memory");
    std::string err;
    if (!lbk::build_layout(*scene, l->L, &err)) {
        delete l;
        return lbk::fail("%s", err.c_str());
    }
    *out = l;
    return 0;
}

void lb_layout_destroy(lb_layout *layout) { delete layout; }

int lb_layout_size(const lb_layout *l, uint32_t *w, uint32_t *h, size_t *charts, size_t *tiles) {
    if (!l) return lbk::fail("lb_layout_size: null handle");
    if (w) *w = l->L.width;
    if (h) *h = l->L.height;
    if (charts) *charts = l->L.charts.size();
    if (tiles) *tiles = l->L.tiles.size();
    return 0;
}

int lb_layout_get(const lb_layout *l, float *affines, uint32_t *chart_tiles, int32_t *tiles,
                  uint32_t *tri_charts) {
    if (!l) return lbk::fail("lb_layout_get: null handle");
    const lbk::Layout &L = l->L;
    for (size_t c = 0; c < L.charts.size(); c++) {
        if (affines)
            lbk::chart_affine(L.charts[c], L.tiles[L.chart_tile[c]], L.padding, affines + c * 8);
        if (chart_tiles) chart_tiles[c] = L.chart_tile[c];
    }
    if (tiles)
        for (size_t t = 0; t < L.tiles.size(); t++) {
            const lbk::Tile &tl = L.tiles[t];
            int32_t v[6] = {tl.kx, tl.kz, (int32_t)tl.w, (int32_t)tl.h, (int32_t)tl.x, (int32_t)tl.y};
            for (int k = 0; k < 6; k++) tiles[t * 6 + k] = v[k];
        }
    if (tri_charts) {
        size_t k = 0;
        for (auto &m : L.tri_chart)
            for (uint32_t c : m) tri_charts[k++] = c;
    }
    return 0;
}

int lb_layout_rasterize(const lb_scene_desc *scene, const lb_layout *l, size_t tile, lb_texel *out,
                        size_t cap) {
    if (!l || !out) return lbk::fail("lb_layout_rasterize: null handle");
    if (check_scene(scene)) return 1;
    if (tile >= l->L.tiles.size()) return lbk::fail("lb_layout_rasterize: no tile %zu", tile);
    const lbk::Tile &t = l->L.tiles[tile];
    if (cap < (size_t)t.w * t.h)
        return lbk::fail("lb_layout_rasterize: %zu texels for a %ux%u tile", cap, t.w, t.h);
    std::vector<lb_texel> tex;
    lbk::rasterize(*scene, l->L, t, tex);
    std::copy(tex.begin(), tex.end(), out);
    return 0;
}

}  // extern "C"
