// The state of a bake between its steps (prepare, execute, finish), shared by the bake driver
// (bake.hip) and the adjoint (adjoint.hip), which replays a bake's bounces.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "lb_bake.h"
#include "lb_host.h"
#include "lb_texel.h"

struct lb_bake {
    lb_gpu *gpu = nullptr;
    lb_bake_options opt{};
    lb_scene_desc sc{};  // valid while the caller's scene lives (prepare .. finish)
    lbk::Layout layout;
    lb_bake_info info{};
    lb_scene *scene = nullptr;
    lb_params P{}, PV{};
    lbk::V3 sun{};
    uint32_t g = 0, aw = 0, ah = 0, ds = 1, rg = 0, rings = 0, grings = 0;
    // the tiles baked (dirty), in layout order; every tile's first texel in the cache order
    std::vector<uint32_t> baked;
    std::vector<size_t> toff_all;
    // host copies of the baked tiles' texels (toff: each baked tile's first)
    std::vector<uint32_t> toff, chart;
    std::vector<lb_texel> texels;
    std::vector<float> normal;
    std::vector<uint8_t> covered;
    std::vector<uint32_t> cov;
    uint32_t vsz[3] = {0, 0, 0};
    size_t T = 0, ncov = 0, ng = 0, np = 0, npd = 0;
    std::vector<uint32_t> gat;
    std::vector<uint32_t> pdirty;  // the probes traced, in index order
    // every tile's TileLight (bake.rs:61-69), clean tiles from the cache, in the cache order
    std::vector<uint8_t> tl_known, tl_state;
    std::vector<float> tl_field, tl_ao, tl_direct, tl_indirect;
    std::vector<float> probes_raw, light0;
    bool any_clean = false;
    // device state
    lbk::DevBuf d_tex0, d_tex, d_tile_of, d_tiles, d_chart, d_normal, d_valid, d_known, d_state,
        d_hard, d_soft, d_ao, d_field, d_direct, d_indirect, d_cov, d_items, d_res, d_light, d_w1,
        d_wg, d_f1, d_f4, d_m1, d_m2, d_covered, d_gat, d_gmask, d_gitems, d_gres, d_filled,
        d_pitems, d_pres, d_counts, d_total, d_light0, d_cost, d_perm, d_rec, d_raster, d_owner;
    lbk::TexelView V{};
    hipEvent_t ev[8] = {};
    hipGraphExec_t graph = nullptr;
    size_t graph_ng = 0;  // the gather item count the graph was captured with
    lbk::Tuning graph_tuning;  // and the tuning
    // results
    std::vector<float> direct, indirect, volume, charts, light;
    std::vector<uint8_t> validity;
    std::vector<uint32_t> tri_charts, chart_tiles;
    std::vector<int32_t> tiles;
    std::vector<lb_node> nodes;
    std::vector<lb_tri> tris;
    std::vector<float> mats;

    ~lb_bake() {
        if (gpu) hipStreamSynchronize(gpu->stream);
        if (graph) hipGraphExecDestroy(graph);
        for (auto &e : ev)
            if (e) hipEventDestroy(e);
        if (scene) lb_scene_destroy(scene);
    }
};

