// The state of a bake between its steps (prepare, execute, finish), shared by the bake driver
// (bake.hip) and the adjoint (adjoint.hip), which replays a bake's bounces.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <vector>

#include "lb_atlas.h"
#include "lb_bake.h"
#include "lb_host.h"
#include "lb_texel.h"

// What a captured bounce graph was recorded with. Its kernels take these by value, so the graph
// replays only while all of them still hold: the gather item count and the gather's tuning, and
// the GPU-wide buffers the TUNED form's launches read (a later launch with another tuning may
// grow, that is free and reallocate, the spill area or the pooled schedule's hits).
// Each buffer is keyed by address and capacity: a regrown buffer may come back at the same
// address, and the capacity tells the two apart.
struct GraphKey {
    size_t ng = 0;
    lbk::Tuning tuning;
    const void *ptr[6] = {};  // spill, pooled hits, pooled aux, cost, perm, records
    size_t cap[6] = {};
    bool operator==(const GraphKey &o) const {
        if (ng != o.ng || !(tuning == o.tuning)) return false;
        for (int i = 0; i < 6; i++)
            if (ptr[i] != o.ptr[i] || cap[i] != o.cap[i]) return false;
        return true;
    }
};

// A bake holds a reference to its lb_gpu (released when the bake is deleted); an adjoint holds one
// to its bake, so lb_bake_destroy defers the delete until the last adjoint over it is gone.
struct lb_bake {
    std::atomic<int> refs{1};
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
    // each baked tile's first texel in the baked list
    std::vector<uint32_t> toff;
    uint32_t vsz[3] = {0, 0, 0};
    size_t T = 0, T_all = 0, ncov = 0, ng = 0, np = 0, npd = 0;
    std::vector<uint32_t> gat;
    std::vector<uint32_t> pdirty;  // the probes traced, in index order
    std::vector<float> probes_raw;  // every probe as traced (the cache's for clean ones)
    bool any_clean = false;
    lbk::AtlasMap amap{};
    // host copies of device arrays, read back when lb_bake_array asks (bit k: array k is held)
    mutable uint64_t held = 0;
    mutable std::vector<uint8_t> h_tl_known, h_tl_state, h_bc5, h_vbytes, h_cost, h_perm;
    bool ranked = false;  // the last execute ranked the gather's rays (d_cost, d_perm hold them)
    // the gather item count d_perm holds a ranking for (0: none yet); every execute traces the
    // same gather rays, so a ranking stays a valid permutation of each item's rays
    size_t perm_ng = 0;
    mutable std::vector<float> h_tl_field, h_tl_ao, h_tl_direct, h_tl_indirect, h_light, h_light0;
    mutable std::vector<uint32_t> h_rgb;
    mutable std::vector<lb_texel> h_texels;
    // device state
    lbk::DevBuf d_tex0, d_tex, d_tile_of, d_tiles, d_chart, d_normal, d_valid, d_known, d_state,
        d_hard, d_soft, d_ao, d_field, d_direct, d_indirect, d_cov, d_items, d_res, d_light, d_w1,
        d_wg, d_f1, d_f4, d_m1, d_m2, d_covered, d_gat, d_gmask, d_gitems, d_gres, d_filled,
        d_pitems, d_pres, d_counts, d_total, d_light0, d_cost, d_perm, d_rec, d_raster, d_owner,
        d_ring;
    // every tile's TileLight (bake.rs:61-69) in the cache order, the atlas' tile map, and the
    // finish's outputs
    lbk::DevBuf d_tl_known, d_tl_state, d_tl_field, d_tl_ao, d_tl_direct, d_tl_indirect, d_atiles,
        d_bmap, d_dst_base, d_probes, d_adirect, d_avalid, d_aind, d_atmp, d_has0, d_has1, d_bc5,
        d_rgb, d_vbytes, d_word;
    lbk::TexelView V{};
    hipEvent_t ev[10] = {};
    hipGraphExec_t graph = nullptr;
    GraphKey graph_key;        // what the graph was captured with
    bool graph_ranks = false;  // the graph's first bounce counts and ranks the gather's rays
    uint32_t graph_captures = 0;
    // results
    std::vector<float> direct, indirect, volume, charts;
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
        lbk::gpu_release(gpu);
    }
};

