// The bake (crates/lightbaker/src/bake.rs:94-432 and volume.rs:34-148) in three steps:
//   prepare  the host's part before any ray: layout, BVH, Params, the tiles' texels, the probes,
//            all uploaded (the reference's bake.rs:98-149, volume.rs:36-75);
//   execute  every device pass on resident data, in one stream: the direct pass, its texel
//            passes, the gather set, the bounces (optionally replayed from a captured hipGraph)
//            and the volume pass; repeatable (it starts from the prepared state each time);
//   finish   after the rays: probes inside solids on the host (volume.rs:86-117), then on the
//            device the volume's light for texels with none, every tile's TileLight, the atlas
//            and the shipped encodings (bake.rs:337-421, output.rs:95-110), read back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <set>
#include <thread>
#include <type_traits>
#include <vector>

#include "lb_atlas.h"
#include "lb_bake_state.h"
#include "lb_raster.h"
#include "lb_records.h"

using namespace lbk;

namespace lbk {
namespace {

constexpr float LIFT = 0.01f;          // bake.rs:24
constexpr float INSIDE_REACH = 0.3f;   // bake.rs:26
constexpr float VOLUME_INSIDE = 0.5f;  // volume.rs:20

using clk = std::chrono::steady_clock;
double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

template <typename T>
int put(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    return b.upload(v.data(), v.size() * sizeof(T), s);
}

TileLightView tile_light_view(const lb_bake *B) {
    return TileLightView{(uint8_t *)B->d_tl_known.ptr, (uint8_t *)B->d_tl_state.ptr,
                         (float *)B->d_tl_field.ptr,   (float *)B->d_tl_ao.ptr,
                         (float *)B->d_tl_direct.ptr,  (float *)B->d_tl_indirect.ptr};
}

#define TRYH(expr)                                        \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)
#define TRY(expr)           \
    do {                    \
        if (expr) return 1; \
    } while (0)

// filter.rs:34: the Gaussian's weights over the (2r + 1)^2 window, with the host's expf
std::vector<float> gauss(int r) {
    std::vector<float> w((size_t)(2 * r + 1) * (2 * r + 1), 0.0f);
    float rf = (float)r * 0.75f;
    float sigma2 = 2.0f * (rf * rf);
    for (int dy = -r; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++) {
            float fx = (float)dx, fy = (float)dy;
            w[(size_t)(dy + r) * (2 * r + 1) + (dx + r)] = std::exp(-(fx * fx + fy * fy) / sigma2);
        }
    return w;
}


}  // namespace
}  // namespace lbk

namespace lbk {

int bake_prepare(lb_gpu *gpu, const lb_scene_desc &sc, const lb_bake_options &opt, lb_bake *B) {
    auto t_start = clk::now();
    hipStream_t st = gpu->stream;
    hipSetDevice(gpu->device);
    B->gpu = gpu;
    gpu_retain(gpu);  // released by ~lb_bake
    B->opt = opt;
    B->sc = sc;
    lb_bake_info &I = B->info;
    if (!sc.materials || !sc.n_materials) return fail("scene: no materials");
    for (size_t m = 0; m < sc.n_meshes; m++) {
        if (sc.meshes[m].material >= sc.n_materials)
            return fail("mesh %zu: no material %u", m, sc.meshes[m].material);
        if (sc.meshes[m].triangles && !sc.meshes[m].positions)
            return fail("mesh %zu: no positions", m);
    }
    // scene.rs:188-202
    if (!(sc.texels_per_meter > 0.0f) || !(sc.tile_size > 0.0f))
        return fail("lightmap: texelsPerMeter and tileSize must be positive");
    uint32_t g = (uint32_t)std::sqrt((float)sc.sun_samples);
    if (g * g != sc.sun_samples || g < 1) return fail("bake: sunSamples must be a square number");
    if (sc.sun_samples > 24) return fail("bake: sunSamples above 24 would reach the pick bits");
    if (sc.ao_rays < 1 || sc.gi_rays < 1) return fail("bake: aoRays and giRays at least one");
    if (sc.indirect_downscale != 1 && sc.indirect_downscale != 2 && sc.indirect_downscale != 4)
        return fail("lightmap: indirectDownscale 1, 2 or 4 (the tiles' grid)");
    if (sc.atlas_width < 64 || sc.atlas_width % 4 || sc.atlas_width > 65535)
        return fail("lightmap: atlasWidth must be a multiple of 4 in 64..65535");
    if (sc.has_volume) {
        if (!(sc.volume_step > 0.0f)) return fail("volume: bad bounds or step");
        for (int a = 0; a < 3; a++)
            if (!(sc.volume_max[a] >= sc.volume_min[a])) return fail("volume: bad bounds or step");
        if (sc.volume_rays < 1) return fail("volume: rays at least one");
    }
    B->g = g;

    // layout, charts, BVH (bake.rs:98-104)
    auto t0 = clk::now();
    Layout &L = B->layout;
    std::string err;
    if (!build_layout(sc, L, &err)) return fail("%s", err.c_str());
    if (L.height > 65535) return fail("layout: atlas %u texels high exceeds 65535", L.height);
    I.s_layout = secs(t0);
    uint32_t aw = L.width, ah = L.height;
    B->aw = aw;
    B->ah = ah;
    size_t n_charts = L.charts.size();
    B->charts.assign(n_charts * 8, 0.0f);
    for (size_t c = 0; c < n_charts; c++)
        chart_affine(L.charts[c], L.tiles[L.chart_tile[c]], L.padding, &B->charts[c * 8]);
    t0 = clk::now();
    size_t n_tris = 0;
    for (size_t m = 0; m < sc.n_meshes; m++) n_tris += sc.meshes[m].triangles;
    std::vector<float> pos(n_tris * 9);
    std::vector<uint32_t> tmat(n_tris), tchart(n_tris);
    {
        size_t k = 0;
        for (size_t m = 0; m < sc.n_meshes; m++) {
            const lb_mesh_desc &md = sc.meshes[m];
            if (md.triangles)
                std::memcpy(&pos[k * 9], md.positions, (size_t)md.triangles * 9 * sizeof(float));
            for (uint32_t t = 0; t < md.triangles; t++, k++) {
                tmat[k] = md.material;
                tchart[k] = L.tri_chart[m][t];
            }
        }
    }
    B->tri_charts = tchart;
    std::vector<uint8_t> casts(sc.n_materials);
    B->mats.assign(sc.n_materials * 4, 0.0f);
    for (size_t m = 0; m < sc.n_materials; m++) {
        casts[m] = sc.materials[m].casts ? 1 : 0;
        for (int c = 0; c < 3; c++) B->mats[m * 4 + c] = sc.materials[m].albedo[c];
        B->mats[m * 4 + 3] = sc.materials[m].opacity;
    }
    build_bvh(pos.data(), n_tris, tmat.data(), tchart.data(), casts.data(), B->nodes, B->tris,
              opt.threads);
    I.bvh_nodes = (uint32_t)B->nodes.size();
    I.bvh_tris = (uint32_t)B->tris.size();
    I.s_bvh = secs(t0);
    if (lb_scene_create(gpu, B->nodes.data(), B->nodes.size(), B->tris.data(), B->tris.size(),
                        B->mats.data(), sc.n_materials, n_charts ? B->charts.data() : nullptr,
                        n_charts, &B->scene) ||
        lb_scene_set_sky(B->scene, sc.sky_sh))
        return 1;

    // Params (bake.rs:105-120)
    V3 sun{sc.sun_direction[0], sc.sun_direction[1], sc.sun_direction[2]};
    sun = sun * (1.0f / std::sqrt(dot(sun, sun)));
    B->sun = sun;
    const float PIF = 3.14159265358979323846f;
    lb_params &P = B->P;
    P = lb_params{};
    P.sun_dir[0] = sun.x;
    P.sun_dir[1] = sun.y;
    P.sun_dir[2] = sun.z;
    P.sun_dir[3] = std::tan(sc.sun_radius_deg * (PIF / 180.0f));
    for (int c = 0; c < 3; c++) {
        P.sun_irr[c] = sc.sun_irradiance[c];
        P.zenith[c] = sc.zenith[c];
        P.horizon[c] = sc.horizon[c];
        P.ground[c] = sc.ground[c];
        P.fallback[c] = (sc.zenith[c] + sc.horizon[c]) * 0.5f * PIF;
        // the sky's L1 term (no counterpart in the reference): its irradiance on an upward face,
        // pi a0 c0 + 2 pi / 3 a1 c2 (the volume's cosine lobe, volume.rs:125-147); 0 without one
        P.fallback[c] += SH_FALLBACK_C0 * sc.sky_sh[c] + SH_FALLBACK_C2 * sc.sky_sh[6 + c];
    }
    P.ground[3] = std::fmin(INSIDE_REACH, sc.ao_reach);
    P.counts[0] = g;
    P.counts[1] = sc.ao_rays;
    P.counts[2] = sc.gi_rays;
    P.reach[0] = sc.ao_reach;
    P.reach[1] = sc.gi_reach;
    P.reach[2] = LIFT;
    P.atlas[0] = aw;
    P.atlas[1] = ah;
    std::memcpy(&P.atlas[3], &sc.inside_share, 4);
    B->PV = P;
    B->PV.counts[2] = sc.volume_rays;  // volume.rs:76

    // which tiles to bake, and the clean ones' light from the cache (bake.rs:123-136)
    t0 = clk::now();
    size_t nt = L.tiles.size();
    B->toff_all.assign(nt + 1, 0);
    for (size_t t = 0; t < nt; t++)
        B->toff_all[t + 1] = B->toff_all[t] + (size_t)L.tiles[t].w * L.tiles[t].h;
    size_t T_all = B->toff_all[nt];
    if (T_all > 0x7fffffffu) return fail("bake: %zu texels exceed 31-bit indices", T_all);
    B->T_all = T_all;
    const lb_bake_cache *cache = opt.cache;
    for (size_t t = 0; t < nt; t++)
        if (!cache || !cache->tile_dirty || cache->tile_dirty[t]) B->baked.push_back((uint32_t)t);
    B->any_clean = B->baked.size() < nt;
    if (B->any_clean) {
        const lb_bake_cache &c = *cache;
        if (!c.known || !c.state || !c.field || !c.ao || !c.direct || !c.indirect)
            return fail("bake cache: clean tiles need known, state, field, ao, direct, indirect");
        if (c.n_texels != T_all)
            return fail("bake cache: %zu texels, the layout's tiles hold %zu", c.n_texels, T_all);
    }
    // every tile's TileLight stays on the device (what the cache receives): a rebake's clean
    // tiles from the cache, the baked tiles' written by finish
    {
        size_t Ta = std::max<size_t>(T_all, 1);
        TRY(B->d_tl_known.reserve(Ta));
        TRY(B->d_tl_state.reserve(Ta));
        TRY(B->d_tl_field.reserve(Ta * 4));
        TRY(B->d_tl_ao.reserve(Ta * 4));
        TRY(B->d_tl_direct.reserve(Ta * 12));
        TRY(B->d_tl_indirect.reserve(Ta * 12));
    }
    // the atlas' 4 x 4 blocks -> their tile (tiles are 4-aligned: layout.rs:18,216,236,260)
    {
        std::vector<AtlasTile> at(std::max<size_t>(nt, 1));
        uint32_t bw = (aw + 3) / 4, bh = (ah + 3) / 4;
        std::vector<uint32_t> bmap((size_t)bw * bh, ATLAS_NONE);
        size_t k = 0;
        for (size_t t = 0; t < nt; t++) {
            const Tile &tl = L.tiles[t];
            bool dirty = k < B->baked.size() && B->baked[k] == t;
            if (dirty) k++;
            if ((tl.x | tl.y | tl.w | tl.h) & 3u) return fail("layout: tile %zu is not 4-aligned", t);
            at[t] = AtlasTile{(uint32_t)B->toff_all[t], tl.w, tl.h, tl.x, tl.y, dirty ? 1u : 0u};
            for (uint32_t by = tl.y / 4; by < (tl.y + tl.h) / 4; by++)
                for (uint32_t bx = tl.x / 4; bx < (tl.x + tl.w) / 4; bx++)
                    bmap[(size_t)by * bw + bx] = (uint32_t)t;
        }
        TRY(put(B->d_atiles, at, st));
        TRY(put(B->d_bmap, bmap, st));
        B->amap = AtlasMap{(const AtlasTile *)B->d_atiles.ptr, (const uint32_t *)B->d_bmap.ptr,
                           aw, ah, bw};
    }
    if (B->any_clean) {
        // the clean tiles' TileLight from the cache, and their light (bake.rs:435-444): the
        // atlas the first bounce reads
        const lb_bake_cache &c = *cache;
        TRYH(hipMemcpyAsync(B->d_tl_known.ptr, c.known, T_all, hipMemcpyHostToDevice, st));
        TRYH(hipMemcpyAsync(B->d_tl_state.ptr, c.state, T_all, hipMemcpyHostToDevice, st));
        TRYH(hipMemcpyAsync(B->d_tl_field.ptr, c.field, T_all * 4, hipMemcpyHostToDevice, st));
        TRYH(hipMemcpyAsync(B->d_tl_ao.ptr, c.ao, T_all * 4, hipMemcpyHostToDevice, st));
        TRYH(hipMemcpyAsync(B->d_tl_direct.ptr, c.direct, T_all * 12, hipMemcpyHostToDevice, st));
        TRYH(hipMemcpyAsync(B->d_tl_indirect.ptr, c.indirect, T_all * 12, hipMemcpyHostToDevice, st));
        TRY(B->d_light0.reserve((size_t)aw * ah * 16));
        TRYH(atlas_clean_light(B->amap, tile_light_view(B), (float4 *)B->d_light0.ptr, st));
        TRYH(hipStreamSynchronize(st));  // the cache is borrowed until prepare returns
    }

    // the dirty tiles' texels (bake.rs:139-148), rasterized on the device (raster.rs:39-114)
    size_t nd = B->baked.size();
    I.tiles = (uint32_t)nt;
    I.tiles_baked = (uint32_t)nd;
    I.charts = (uint32_t)n_charts;
    B->chart_tiles = L.chart_tile;
    B->tiles.resize(nt * 6);
    for (size_t t = 0; t < nt; t++) {
        const Tile &tl = L.tiles[t];
        int32_t v[6] = {tl.kx, tl.kz, (int32_t)tl.w, (int32_t)tl.h, (int32_t)tl.x, (int32_t)tl.y};
        std::memcpy(&B->tiles[t * 6], v, sizeof v);
    }
    B->toff.assign(nd + 1, 0);
    for (size_t k = 0; k < nd; k++) {
        const Tile &tl = L.tiles[B->baked[k]];
        B->toff[k + 1] = B->toff[k] + tl.w * tl.h;
    }
    size_t T = B->toff[nd];
    if (T > 0x7fffffffu) return fail("bake: %zu texels exceed 31-bit indices", T);
    B->T = T;
    std::vector<uint32_t> tile_of(T);
    std::vector<uint4> tiles(std::max<size_t>(nd, 1));
    for (size_t t = 0; t < nd; t++) {
        const Tile &tl = L.tiles[B->baked[t]];
        uint32_t o = B->toff[t];
        tiles[t] = make_uint4(o, tl.w, tl.h, tl.x | (tl.y << 16));
        std::fill(tile_of.begin() + o, tile_of.begin() + o + (size_t)tl.w * tl.h, (uint32_t)t);
    }
    {
        std::vector<RasterTri> recs;
        raster_records(sc, L, B->baked, recs);
        if (recs.size() > 0x7fffffffu) return fail("bake: %zu triangles to rasterize", recs.size());
        size_t Tn0 = std::max<size_t>(T, 1);
        TRY(put(B->d_tiles, tiles, st));
        TRY(put(B->d_tile_of, tile_of, st));
        TRY(put(B->d_raster, recs, st));
        TRY(B->d_owner.reserve(Tn0 * 4));
        TRY(B->d_tex0.reserve(Tn0 * sizeof(lb_texel)));
        TRYH(raster_tiles((const RasterTri *)B->d_raster.ptr, (uint32_t)recs.size(),
                          (const uint4 *)B->d_tiles.ptr, (const uint32_t *)B->d_tile_of.ptr,
                          (uint32_t)T, (uint32_t *)B->d_owner.ptr, (lb_texel *)B->d_tex0.ptr, st));
    }
    // each texel's covered flag, chart and normal, and the covered texels' list (bake.rs:149),
    // on the device
    {
        size_t Tn = std::max<size_t>(T, 1);
        TRY(B->d_chart.reserve(Tn * 4));
        TRY(B->d_normal.reserve(Tn * 16));
        TRY(B->d_covered.reserve(Tn));
        TRY(B->d_cov.reserve(Tn * 4));
        TRY(B->d_counts.reserve((size_t)tp_gather_blocks((uint32_t)Tn) * 4 + 16));
        TRY(B->d_total.reserve(16));
        TRYH(raster_derive((const lb_texel *)B->d_tex0.ptr, (uint32_t)T, (uint32_t *)B->d_chart.ptr,
                           (float4 *)B->d_normal.ptr, (uint8_t *)B->d_covered.ptr, st));
        tp_compact((const uint8_t *)B->d_covered.ptr, (uint32_t)T, (uint32_t *)B->d_counts.ptr,
                   (uint32_t *)B->d_total.ptr, (uint32_t *)B->d_cov.ptr, st);
        TRYH(hipGetLastError());
        uint32_t ncov = 0;
        TRYH(hipMemcpyAsync(&ncov, B->d_total.ptr, 4, hipMemcpyDeviceToHost, st));
        TRYH(hipStreamSynchronize(st));
        B->ncov = ncov;
    }
    {
        // baked texel g of baked tile k is all-tile texel g + dst_base[k]
        std::vector<int64_t> dst(std::max<size_t>(nd, 1), 0);
        for (size_t k = 0; k < nd; k++)
            dst[k] = (int64_t)B->toff_all[B->baked[k]] - (int64_t)B->toff[k];
        TRY(put(B->d_dst_base, dst, st));
    }
    I.texels = T;
    I.texels_baked = B->ncov;
    I.s_raster = secs(t0);

    // the probes (volume.rs:36-75): the dirty ones in index order (a probe's key is its index, so
    // the reference's chunk order gives the same rays); the clean ones from the cache
    std::vector<lb_texel> pitems;
    if (sc.has_volume) {
        for (int a = 0; a < 3; a++)
            B->vsz[a] = std::max<uint32_t>(
                (uint32_t)std::round((sc.volume_max[a] - sc.volume_min[a]) / sc.volume_step) + 1, 2);
        size_t nx = B->vsz[0], ny = B->vsz[1], nz = B->vsz[2];
        B->np = nx * ny * nz;
        if (B->np > 0x7fffffffu) return fail("volume: %zu probes exceed 31-bit indices", B->np);
        B->probes_raw.assign(B->np * 16, 0.0f);
        bool some_clean = false;
        for (size_t k = 0; k < B->np; k++) {
            if (cache && cache->probe_dirty && !cache->probe_dirty[k]) {
                some_clean = true;
                continue;
            }
            B->pdirty.push_back((uint32_t)k);
        }
        if (some_clean) {
            if (!cache->probes) return fail("bake cache: clean probes need their values");
            if (cache->n_probes != B->np)
                return fail("bake cache: %zu probes, the volume holds %zu", cache->n_probes, B->np);
            for (size_t k = 0, d = 0; k < B->np; k++) {
                if (d < B->pdirty.size() && B->pdirty[d] == k) {
                    d++;
                    continue;
                }
                std::memcpy(&B->probes_raw[k * 16], &cache->probes[k * 16], 64);
            }
        }
        B->npd = B->pdirty.size();
        pitems.resize(B->npd);
        for (size_t d = 0; d < B->npd; d++) {
            size_t k = B->pdirty[d];
            float px = sc.volume_min[0] + (float)(k % nx) * sc.volume_step;
            float py = sc.volume_min[1] + (float)(k / (nx * nz)) * sc.volume_step;
            float pz = sc.volume_min[2] + (float)((k / nx) % nz) * sc.volume_step;
            uint32_t kk = (uint32_t)k;
            float kf;
            std::memcpy(&kf, &kk, 4);
            pitems[d] = lb_texel{{px, py, pz, 1.0f}, {0.0f, 1.0f, 0.0f, kf}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        }
        I.probes = B->np;
        I.probes_baked = B->npd;
        for (int a = 0; a < 3; a++) {
            I.volume_size[a] = B->vsz[a];
            I.volume_origin[a] = sc.volume_min[a];
        }
        I.volume_step = sc.volume_step;
    }

    // device state
    size_t Tn = std::max<size_t>(T, 1);
    TRY(B->d_tex.reserve(Tn * sizeof(lb_texel)));
    TRY(B->d_valid.reserve(Tn));
    TRY(B->d_known.reserve(Tn));
    TRY(B->d_state.reserve(Tn));
    TRY(B->d_filled.reserve(Tn));
    TRY(B->d_gmask.reserve(Tn));
    TRY(B->d_hard.reserve(Tn * 4));
    TRY(B->d_soft.reserve(Tn * 4));
    TRY(B->d_ao.reserve(Tn * 4));
    TRY(B->d_field.reserve(Tn * 4));
    TRY(B->d_direct.reserve(Tn * 16));
    TRY(B->d_indirect.reserve(Tn * 16));
    TRY(B->d_f1.reserve(Tn * 4));
    TRY(B->d_f4.reserve(Tn * 16));
    TRY(B->d_m1.reserve(Tn));
    TRY(B->d_m2.reserve(Tn));
    TRY(B->d_ring.reserve(Tn));
    TRY(B->d_items.reserve(std::max<size_t>(B->ncov, 1) * sizeof(lb_texel)));
    TRY(B->d_res.reserve(std::max<size_t>(B->ncov, 1) * 16));
    TRY(B->d_light.reserve((size_t)aw * ah * 16));
    TRY(B->d_gat.reserve(Tn * 4));
    TRY(B->d_gitems.reserve(Tn * sizeof(lb_texel)));
    TRY(B->d_gres.reserve(Tn * 16));
    TRY(B->d_counts.reserve((size_t)tp_gather_blocks((uint32_t)Tn) * 4 + 16));
    TRY(B->d_total.reserve(16));
    TRY(put(B->d_w1, gauss(1), st));
    B->ds = sc.indirect_downscale;
    B->rg = sc.gi_filter * sc.indirect_downscale;
    TRY(put(B->d_wg, gauss((int)B->rg), st));
    B->rings = L.padding + 4;                  // bake.rs:243
    B->grings = L.padding + 2 * B->ds + 2;     // bake.rs:320
    if (B->grings >= lbk::RING_NEVER) return fail("bake: too many fill rings for the bounce fill");
    if (B->npd) {
        TRY(put(B->d_pitems, pitems, st));
        TRY(B->d_pres.reserve(B->npd * 64));
    }
    TexelView &V = B->V;
    V.n = (uint32_t)T;
    V.tile_of = (const uint32_t *)B->d_tile_of.ptr;
    V.tiles = (const uint4 *)B->d_tiles.ptr;
    V.texels = (lb_texel *)B->d_tex.ptr;
    V.chart = (const uint32_t *)B->d_chart.ptr;
    V.normal = (const float4 *)B->d_normal.ptr;
    V.valid = (uint8_t *)B->d_valid.ptr;
    V.known = (uint8_t *)B->d_known.ptr;
    V.state = (uint8_t *)B->d_state.ptr;
    V.hard = (uint32_t *)B->d_hard.ptr;
    V.soft = (float *)B->d_soft.ptr;
    V.ao = (float *)B->d_ao.ptr;
    V.field = (float *)B->d_field.ptr;
    V.direct = (float4 *)B->d_direct.ptr;
    V.indirect = (float4 *)B->d_indirect.ptr;
    for (auto &e : B->ev) TRYH(hipEventCreate(&e));
    if (opt.form == LB_FORM_TUNED) {
        // the tuned form's spill area, sized for the largest launch, before any capture
        uint32_t most = (uint32_t)std::max<size_t>({B->ncov, T, B->npd, 1});
        for (int pass = 0; pass < 3; pass++)
            if (tuned_prepare(gpu, pass, pass == LB_PASS_VOLUME ? B->PV : P, most) != hipSuccess)
                return fail("GPU: tuned spill area");
    }
    TRYH(hipStreamSynchronize(st));
    I.s_total = secs(t_start);
    return 0;
}

namespace {

// filter.rs:58-94 over grings rings into the indirect light from the gather set, as bake.rs:320
// fills it after each bounce. Every bounce starts from the same mask, so the rings fill the same
// texels: the first bounce files each texel's ring in d_ring, the later ones take each ring over
// the texels filed for it (the same values bit for bit, without a pass over the unfilled texels).
void fill_bounce(lb_bake *B, uint32_t k) {
    hipStream_t st = B->gpu->stream;
    float *v = (float *)B->V.indirect;
    uint8_t *ring = (uint8_t *)B->d_ring.ptr;
    if (k > 0) {
        for (uint32_t r = 0; r < B->grings; r++) tp_fill_ring_filed(B->V, v, ring, r, st);
        return;
    }
    uint8_t *ma = (uint8_t *)B->d_m1.ptr, *mb = (uint8_t *)B->d_m2.ptr;
    tp_copy_u8((const uint8_t *)B->d_gmask.ptr, ma, (uint32_t)B->T, st);
    for (uint32_t r = 0; r < B->grings; r++) {
        tp_fill_ring_file(B->V, v, ma, mb, ring, r, st);
        std::swap(ma, mb);
    }
}

// the direct phase's three fills (bake.rs:253-255): the occlusion, the sun field and the direct
// light fill from the same mask over the same rings, so each ring is one pass over all three
void fill_direct(lb_bake *B, uint32_t rings, const uint8_t *valid0, uint8_t *final_mask) {
    hipStream_t st = B->gpu->stream;
    TexelView &V = B->V;
    uint32_t T = (uint32_t)B->T;
    uint8_t *ma = (uint8_t *)B->d_m1.ptr, *mb = (uint8_t *)B->d_m2.ptr;
    tp_copy_u8(valid0, ma, T, st);
    for (uint32_t r = 0; r < rings; r++) {
        tp_fill_ring_direct(V, V.ao, V.field, (float *)V.direct, ma, mb, st);
        std::swap(ma, mb);
    }
    tp_copy_u8(ma, final_mask, T, st);
}

int pass_stats(lb_bake *B, int pass, const lb_params *P, const lb_texel *items, size_t n,
               const float *light, size_t n_light, float *out, lb_stats *acc) {
    if (!acc) return run_pass(B->gpu, pass, B->opt.form, B->scene, P, items, n, 0, light, n_light,
                              out, false);
    lb_stats s{};
    if (lb_run_stats(B->gpu, pass, B->opt.form, B->scene, P, items, n, 0, light, n_light, out, &s))
        return 1;
    acc->rays += s.rays;
    acc->inner_visits += s.inner_visits;
    acc->tri_tests += s.tri_tests;
    acc->material_loads += s.material_loads;
    acc->light_gathers += s.light_gathers;
    acc->leaf_visits += s.leaf_visits;
    return 0;
}

// Whether the TUNED lock-step gather takes each item's rays cheapest first: every bounce traces
// the same rays (the same keys and sample points; only the light they read changes), so the
// first bounce's step counts rank the rays for the others, and the lanes of a round then trace
// rays of like length. The ranking changes only which lane traces which ray when.
bool ordered_gather(const lb_bake *B) {
    const Tuning &t = B->gpu->tuning[LB_PASS_GATHER];
    return B->opt.form == LB_FORM_TUNED && t.schedule == 0 && t.gather_order != 2 &&
           B->P.counts[2] <= 256u &&
           B->sc.gi_bounces > 1 && B->ng > 0 && B->d_perm.ptr;
}

// one bounce (bake.rs:301-328): gather, the light at its texels, smooth, fill, the atlas; k its
// index
int bounce(lb_bake *B, lb_stats *acc, uint32_t k) {
    hipStream_t st = B->gpu->stream;
    TexelView &V = B->V;
    uint32_t T = (uint32_t)B->T;
    bool replay = !acc && B->opt.replay && B->sc.gi_bounces > 1 && B->d_rec.ptr;
    if (replay && k > 0) {
        // the rays of the first bounce, met again: their records with this bounce's atlas
        hipError_t e = replay_gather((const uint2 *)B->d_rec.ptr, (uint32_t)B->ng,
                                     B->P.counts[2], (const float4 *)B->scene->materials.ptr,
                                     (const float4 *)B->d_light.ptr,
                                     sky_of(B->P, sky_sh_of(B->scene->sky_sh),
                                            (const lb_texel *)B->d_gitems.ptr),
                                     (float4 *)B->d_gres.ptr, st);
        if (e != hipSuccess) return hip_fail(e, "replay_gather");
    } else if (replay) {
        const Tuning &t = B->gpu->tuning[LB_PASS_GATHER];
        if (B->opt.form == LB_FORM_TUNED && t.schedule == 0) {
            // the lock-step gather writes each ray's record as it shades it
            GatherOrder o;
            o.rec = (uint2 *)B->d_rec.ptr;
            TRY(run_pass(B->gpu, LB_PASS_GATHER, B->opt.form, B->scene, &B->P,
                         (const lb_texel *)B->d_gitems.ptr, B->ng, 0,
                         (const float *)B->d_light.ptr, (size_t)B->aw * B->ah,
                         (float *)B->d_gres.ptr, false, &o));
        } else {
            TRY(pass_stats(B, LB_PASS_GATHER, &B->P, (const lb_texel *)B->d_gitems.ptr, B->ng,
                           (const float *)B->d_light.ptr, (size_t)B->aw * B->ah,
                           (float *)B->d_gres.ptr, nullptr));
            hipError_t e = record_gather(scene_view(B->scene, nullptr), B->P,
                                         (const lb_texel *)B->d_gitems.ptr, (uint32_t)B->ng,
                                         B->P.counts[2], (uint2 *)B->d_rec.ptr, st);
            if (e != hipSuccess) return hip_fail(e, "record_gather");
        }
    } else if (!acc && ordered_gather(B)) {
        // the first bounce counts each ray's steps, the rest take the rays ranked by them
        GatherOrder o;
        o.cost = k == 0 ? (uint8_t *)B->d_cost.ptr : nullptr;
        o.perm = k == 0 ? nullptr : (const uint8_t *)B->d_perm.ptr;
        TRY(run_pass(B->gpu, LB_PASS_GATHER, B->opt.form, B->scene, &B->P,
                     (const lb_texel *)B->d_gitems.ptr, B->ng, 0, (const float *)B->d_light.ptr,
                     (size_t)B->aw * B->ah, (float *)B->d_gres.ptr, false, &o));
        if (k == 0) {
            hipError_t e = rank_slots((const uint8_t *)B->d_cost.ptr, (uint32_t)B->ng,
                                      B->P.counts[2], (uint8_t *)B->d_perm.ptr, st);
            if (e != hipSuccess) return hip_fail(e, "rank_slots");
            B->ranked = true;
        }
    } else {
        TRY(pass_stats(B, LB_PASS_GATHER, &B->P, (const lb_texel *)B->d_gitems.ptr, B->ng,
                       (const float *)B->d_light.ptr, (size_t)B->aw * B->ah,
                       (float *)B->d_gres.ptr, acc));
    }
    // the gathered irradiance at its texels (zero elsewhere) in d_f4, smoothed into the
    // indirect light
    float4 *irr = (float4 *)B->d_f4.ptr;
    tp_fill_f32((float *)irr, 0.0f, T * 4u, st);
    tp_scatter((const float4 *)B->d_gres.ptr, (const uint32_t *)B->d_gat.ptr, (uint32_t)B->ng,
               irr, st);
    tp_smooth(V, 4, (const float *)irr, (float *)V.indirect, (const uint8_t *)B->d_gmask.ptr,
              (int)B->rg, (const float *)B->d_wg.ptr, st);
    fill_bounce(B, k);
    tp_write_light(V, (float4 *)B->d_light.ptr, B->aw, st);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "bounce");
}

}  // namespace

int bake_execute(lb_bake *B, lb_stats *per_pass) {
    lb_gpu *gpu = B->gpu;
    hipStream_t st = gpu->stream;
    hipSetDevice(gpu->device);
    const lb_scene_desc &sc = B->sc;
    lb_bake_info &I = B->info;
    TexelView &V = B->V;
    uint32_t T = (uint32_t)B->T;
    size_t Tn = std::max<size_t>(B->T, 1);
    lb_stats *sd = per_pass ? &per_pass[0] : nullptr;
    lb_stats *sg = per_pass ? &per_pass[1] : nullptr;
    lb_stats *sv = per_pass ? &per_pass[2] : nullptr;
    auto t_exec = clk::now();
    B->held = 0;  // the host copies of device arrays are stale from here
    B->ranked = false;

    // the prepared state (bake.rs:204-205: soft 0, ao 1, hard 0, no state; valid = covered)
    TRYH(hipEventRecord(B->ev[0], st));
    TRYH(hipMemcpyAsync(V.texels, B->d_tex0.ptr, B->T * sizeof(lb_texel), hipMemcpyDeviceToDevice, st));
    if (B->any_clean)
        TRYH(hipMemcpyAsync(B->d_light.ptr, B->d_light0.ptr, (size_t)B->aw * B->ah * 16,
                            hipMemcpyDeviceToDevice, st));
    else
        TRYH(hipMemsetAsync(B->d_light.ptr, 0, (size_t)B->aw * B->ah * 16, st));
    TRYH(hipMemcpyAsync(V.valid, B->d_covered.ptr, B->T, hipMemcpyDeviceToDevice, st));
    TRYH(hipMemsetAsync(V.state, 0, Tn, st));
    TRYH(hipMemsetAsync(V.hard, 0, Tn * 4, st));
    TRYH(hipMemsetAsync(V.soft, 0, Tn * 4, st));
    TRYH(hipMemsetAsync(V.indirect, 0, Tn * 16, st));
    tp_fill_f32(V.ao, 1.0f, T, st);

    // the direct pass (bake.rs:149-151)
    tp_collect(V.texels, (const uint32_t *)B->d_cov.ptr, (uint32_t)B->ncov,
               (lb_texel *)B->d_items.ptr, st);
    TRY(pass_stats(B, LB_PASS_DIRECT, &B->P, (const lb_texel *)B->d_items.ptr, B->ncov, nullptr, 0,
                   (float *)B->d_res.ptr, sd));
    TRYH(hipEventRecord(B->ev[1], st));

    // decode, smooth, field, direct light, fill (bake.rs:196-272)
    DecodeArgs da{sc.inside_share, sc.ao_strength, sc.ao_contrast, sc.ao_floor};
    tp_decode(V, (const float4 *)B->d_res.ptr, (const uint32_t *)B->d_cov.ptr, (uint32_t)B->ncov,
              da, st);
    tp_smooth(V, 1, V.ao, (float *)B->d_f1.ptr, V.valid, 1, (const float *)B->d_w1.ptr, st);
    tp_copy_f32((const float *)B->d_f1.ptr, V.ao, T, st);
    tp_smooth(V, 1, V.soft, (float *)B->d_f1.ptr, V.valid, 1, (const float *)B->d_w1.ptr, st);
    tp_sun_field(V, V.valid, B->g, sc.sdf_radius, st);
    tp_direct_light(V, (const float *)B->d_f1.ptr,
                    make_float3(sc.sun_irradiance[0], sc.sun_irradiance[1], sc.sun_irradiance[2]),
                    make_float3(B->sun.x, B->sun.y, B->sun.z), st);
    fill_direct(B, B->rings, V.valid, (uint8_t *)B->d_filled.ptr);
    tp_known(V, (const uint8_t *)B->d_filled.ptr, st);
    tp_write_light(V, (float4 *)B->d_light.ptr, B->aw, st);  // bake.rs:277-281
    TRYH(hipGetLastError());
    TRYH(hipEventRecord(B->ev[2], st));

    // the gather set (bake.rs:282-299) on the device; only its count comes back
    auto t0 = clk::now();
    tp_gather_set(V, B->ds, (uint8_t *)B->d_gmask.ptr, (uint32_t *)B->d_counts.ptr,
                  (uint32_t *)B->d_total.ptr, (uint32_t *)B->d_gat.ptr, st);
    uint32_t ng32 = 0;
    TRYH(hipMemcpyAsync(&ng32, B->d_total.ptr, 4, hipMemcpyDeviceToHost, st));
    TRYH(hipStreamSynchronize(st));
    B->ng = ng32;
    I.gather_items = B->ng;
    I.s_gather_set = secs(t0);
    tp_collect(V.texels, (const uint32_t *)B->d_gat.ptr, (uint32_t)B->ng,
               (lb_texel *)B->d_gitems.ptr, st);

    // the bounces (bake.rs:300-330)
    TRYH(hipEventRecord(B->ev[3], st));
    if (sc.gi_bounces > 0 && B->ng > 0) {
        // the ordered gather's step counts and ranks, one byte a ray (not with replayed
        // bounces: those trace one bounce only)
        size_t nr = B->ng * (size_t)B->P.counts[2];
        if (B->opt.form == LB_FORM_TUNED && B->P.counts[2] <= 256u && sc.gi_bounces > 1 &&
            !B->opt.replay) {
            if (nr > B->d_perm.cap) B->perm_ng = 0;  // a new area holds no ranking
            if (B->d_cost.reserve(nr) || B->d_perm.reserve(nr)) return 1;
        }
        if (B->opt.replay && sc.gi_bounces > 1 && B->d_rec.reserve(nr * 8)) return 1;
        if (B->opt.graph && !per_pass) {
            // the bounces captured once and replayed. The capture holds the item count, the
            // gather's tuning and every buffer's address by value: any change captures again.
            // The spill area is sized for this tuning before the key is taken, so the key sees
            // the buffers the kernels will use.
            if (B->opt.form == LB_FORM_TUNED &&
                tuned_prepare(gpu, LB_PASS_GATHER, B->P, (uint32_t)B->ng) != hipSuccess)
                return fail("GPU: tuned spill area");
            GraphKey key;
            key.ng = B->ng;
            key.tuning = gpu->tuning[LB_PASS_GATHER];
            const DevBuf *bufs[6] = {&gpu->spill, &gpu->stream_hits, &gpu->stream_aux,
                                     &B->d_cost, &B->d_perm, &B->d_rec};
            for (int i = 0; i < 6; i++) {
                key.ptr[i] = bufs[i]->ptr;
                key.cap[i] = bufs[i]->cap;
            }
            if (!B->graph || !(key == B->graph_key)) {
                if (B->graph) hipGraphExecDestroy(B->graph);
                B->graph = nullptr;
                hipGraph_t gr;
                TRYH(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                int rc = 0;
                for (uint32_t b = 0; b < sc.gi_bounces && !rc; b++) rc = bounce(B, nullptr, b);
                hipError_t ce = hipStreamEndCapture(st, &gr);
                if (rc) {
                    if (ce == hipSuccess) hipGraphDestroy(gr);
                    return 1;
                }
                if (ce != hipSuccess) return hip_fail(ce, "graph capture");
                hipError_t ie = hipGraphInstantiate(&B->graph, gr, nullptr, nullptr, 0);
                hipGraphDestroy(gr);
                if (ie != hipSuccess) return hip_fail(ie, "graph instantiate");
                B->graph_key = key;
                B->graph_ranks = B->ranked;
                B->graph_captures++;
            }
            TRYH(hipGraphLaunch(B->graph, st));
            B->ranked = B->graph_ranks;
            I.graph_captures = B->graph_captures;
        } else {
            for (uint32_t b = 0; b < sc.gi_bounces; b++) TRY(bounce(B, sg, b));
        }
        if (B->ranked) B->perm_ng = B->ng;
    }
    TRYH(hipEventRecord(B->ev[4], st));

    // the volume (volume.rs:77)
    if (B->npd)
        TRY(pass_stats(B, LB_PASS_VOLUME, &B->PV, (const lb_texel *)B->d_pitems.ptr, B->npd,
                       (const float *)B->d_light.ptr, (size_t)B->aw * B->ah,
                       (float *)B->d_pres.ptr, sv));
    TRYH(hipEventRecord(B->ev[5], st));
    TRYH(hipStreamSynchronize(st));
    TRYH(hipEventElapsedTime(&I.ms_direct, B->ev[0], B->ev[1]));
    TRYH(hipEventElapsedTime(&I.ms_direct_texel, B->ev[1], B->ev[2]));
    TRYH(hipEventElapsedTime(&I.ms_bounces, B->ev[3], B->ev[4]));
    TRYH(hipEventElapsedTime(&I.ms_volume, B->ev[4], B->ev[5]));
    I.s_execute = secs(t_exec);
    return 0;
}

int bake_finish(lb_bake *B) {
    lb_gpu *gpu = B->gpu;
    hipStream_t st = gpu->stream;
    hipSetDevice(gpu->device);
    const lb_scene_desc &sc = B->sc;
    lb_bake_info &I = B->info;
    auto t0 = clk::now();
    B->held = 0;
    uint32_t aw = B->aw, ah = B->ah, ds = B->ds;
    uint32_t iw = (aw + ds - 1) / ds, ih = (ah + ds - 1) / ds;
    size_t na = (size_t)aw * ah, ni = (size_t)iw * ih;
    // the traced probes beside the cached ones: what volume.rs:81-84 caches
    if (B->npd) {
        std::vector<float> pres(B->npd * 16);
        TRYH(hipMemcpyAsync(pres.data(), B->d_pres.ptr, B->npd * 64, hipMemcpyDeviceToHost, st));
        TRYH(hipStreamSynchronize(st));
        for (size_t d = 0; d < B->npd; d++)
            std::memcpy(&B->probes_raw[(size_t)B->pdirty[d] * 16], &pres[d * 16], 64);
    }
    std::vector<float> probes = B->probes_raw;

    // probes inside solids from their neighbours, 8 rings (volume.rs:86-117)
    size_t nx = B->vsz[0], ny = B->vsz[1], nz = B->vsz[2], np = B->np;
    if (np) {
        std::vector<uint8_t> pk(np);
        for (size_t k = 0; k < np; k++) pk[k] = probes[k * 16 + 7] <= VOLUME_INSIDE;
        for (int ring = 0; ring < 8; ring++) {
            std::vector<uint8_t> was = pk;
            std::vector<float> src = probes;
            for (size_t k = 0; k < np; k++) {
                if (was[k]) continue;
                long x = (long)(k % nx), z = (long)((k / nx) % nz), y = (long)(k / (nx * nz));
                float sum[16] = {};
                int c = 0;
                const long d6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
                for (auto &d : d6) {
                    long xx = x + d[0], yy = y + d[1], zz = z + d[2];
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= (long)nx || yy >= (long)ny || zz >= (long)nz)
                        continue;
                    size_t j = ((size_t)yy * nz + (size_t)zz) * nx + (size_t)xx;
                    if (!was[j]) continue;
                    for (int q = 0; q < 16; q++) sum[q] += src[j * 16 + q];
                    c++;
                }
                if (c > 0) {
                    for (int q = 0; q < 16; q++) probes[k * 16 + q] = sum[q] / (float)c;
                    pk[k] = 1;
                }
            }
        }
    }

    // on the device: the baked tiles' TileLight with the volume's light for texels with none
    // (bake.rs:337-362), the atlas from every tile (bake.rs:364-421), the shipped encodings
    TRYH(hipEventRecord(B->ev[8], st));
    VolumeView vol{};
    if (np) {
        TRY(put(B->d_probes, probes, st));
        vol.probes = (const float4 *)B->d_probes.ptr;
        vol.nx = (uint32_t)nx;
        vol.ny = (uint32_t)ny;
        vol.nz = (uint32_t)nz;
        vol.ox = sc.volume_min[0];
        vol.oy = sc.volume_min[1];
        vol.oz = sc.volume_min[2];
        vol.step = sc.volume_step;
        vol.sun_dir = make_float3(B->sun.x, B->sun.y, B->sun.z);
        vol.sun_irr = make_float3(sc.sun_irradiance[0], sc.sun_irradiance[1], sc.sun_irradiance[2]);
    }
    BakedView bv{};
    bv.n = (uint32_t)B->T;
    bv.tile_of = (const uint32_t *)B->d_tile_of.ptr;
    bv.dst_base = (const int64_t *)B->d_dst_base.ptr;
    bv.covered = (const uint8_t *)B->d_covered.ptr;
    bv.known = (const uint8_t *)B->d_known.ptr;
    bv.state = (const uint8_t *)B->d_state.ptr;
    bv.field = (const float *)B->d_field.ptr;
    bv.ao = (const float *)B->d_ao.ptr;
    bv.direct = (const float4 *)B->d_direct.ptr;
    bv.indirect = (const float4 *)B->d_indirect.ptr;
    bv.pos = (const float4 *)B->d_tex0.ptr;
    bv.normal = (const float4 *)B->d_normal.ptr;
    TileLightView tl = tile_light_view(B);
    TRYH(atlas_tile_light(bv, vol, tl, st));
    TRY(B->d_adirect.reserve(std::max<size_t>(na, 1) * 8));
    TRY(B->d_avalid.reserve(std::max<size_t>(na, 1)));
    TRY(B->d_bc5.reserve(std::max<size_t>(na, 16)));
    TRY(B->d_aind.reserve(std::max<size_t>(ni, 1) * 12));
    TRY(B->d_atmp.reserve(std::max<size_t>(ni, 1) * 12));
    TRY(B->d_has0.reserve(std::max<size_t>(ni, 1)));
    TRY(B->d_has1.reserve(std::max<size_t>(ni, 1)));
    TRY(B->d_rgb.reserve(std::max<size_t>(ni, 1) * 4));
    TRY(B->d_vbytes.reserve(std::max<size_t>(np, 1) * 16));
    TRY(B->d_word.reserve(16));
    TRYH(atlas_assemble(B->amap, tl, ds, (float2 *)B->d_adirect.ptr, (uint8_t *)B->d_avalid.ptr,
                        (float *)B->d_aind.ptr, (float *)B->d_atmp.ptr, (uint8_t *)B->d_has0.ptr,
                        (uint8_t *)B->d_has1.ptr, st));
    // output.rs:99-110 (the atlas width and height are multiples of 4: ALIGN and atlasWidth)
    TRYH(encode_bc5_f32((const float2 *)B->d_adirect.ptr, aw, ah, (uint8_t *)B->d_bc5.ptr, st));
    TRYH(encode_rgb9e5((const float *)B->d_aind.ptr, ni, (uint32_t *)B->d_rgb.ptr, st));
    if (np)
        TRYH(encode_volume((const float4 *)B->d_probes.ptr, np, (uint8_t *)B->d_vbytes.ptr,
                           (uint32_t *)B->d_word.ptr, st));
    TRYH(hipEventRecord(B->ev[9], st));

    // the bake's outputs (bake.rs:36-47): the atlases and the volume
    B->direct.resize(na * 2);
    B->validity.resize(na);
    B->indirect.resize(ni * 3);
    uint32_t scale_bits = 0;
    if (na) {
        TRYH(hipMemcpyAsync(B->direct.data(), B->d_adirect.ptr, na * 8, hipMemcpyDeviceToHost, st));
        TRYH(hipMemcpyAsync(B->validity.data(), B->d_avalid.ptr, na, hipMemcpyDeviceToHost, st));
        TRYH(hipMemcpyAsync(B->indirect.data(), B->d_aind.ptr, ni * 12, hipMemcpyDeviceToHost, st));
    }
    if (np) TRYH(hipMemcpyAsync(&scale_bits, B->d_word.ptr, 4, hipMemcpyDeviceToHost, st));
    TRYH(hipStreamSynchronize(st));
    TRYH(hipEventElapsedTime(&I.ms_finish, B->ev[8], B->ev[9]));
    B->volume = std::move(probes);
    float scale = 1.0f;  // output.rs:108: 1 without a volume
    if (np) std::memcpy(&scale, &scale_bits, 4);
    I.volume_scale = scale;
    I.atlas_width = aw;
    I.atlas_height = ah;
    I.indirect_width = iw;
    I.indirect_height = ih;
    I.s_finish = secs(t0);
    return 0;
}

// Times `reps` launches of one pass over the bake's own items (resident), with HIP events.
int bake_time_pass(lb_bake *B, int pass, int reps, float *ms_avg) {
    lb_gpu *gpu = B->gpu;
    hipStream_t st = gpu->stream;
    const lb_texel *items;
    size_t n;
    const lb_params *P = &B->P;
    float *out;
    if (pass == LB_PASS_DIRECT) {
        items = (const lb_texel *)B->d_items.ptr;
        n = B->ncov;
        out = (float *)B->d_res.ptr;
    } else if (pass == LB_PASS_GATHER) {
        items = (const lb_texel *)B->d_gitems.ptr;
        n = B->ng;
        out = (float *)B->d_gres.ptr;
    } else {
        items = (const lb_texel *)B->d_pitems.ptr;
        n = B->npd;
        P = &B->PV;
        out = (float *)B->d_pres.ptr;
    }
    if (!n) return fail("bake: pass %d has no items (execute first)", pass);
    if (reps < 1) return fail("bake: reps %d", reps);
    // the gather as the bake's later bounces run it (its rays ranked by the first bounce's cost)
    GatherOrder o;
    o.perm = pass == LB_PASS_GATHER && ordered_gather(B) && B->perm_ng == B->ng
                 ? (const uint8_t *)B->d_perm.ptr
                 : nullptr;
    TRYH(hipEventRecord(B->ev[6], st));
    for (int r = 0; r < reps; r++)
        TRY(run_pass(gpu, pass, B->opt.form, B->scene, P, items, n, 0,
                     pass == LB_PASS_DIRECT ? nullptr : (const float *)B->d_light.ptr,
                     (size_t)B->aw * B->ah, out, false, &o));
    TRYH(hipEventRecord(B->ev[7], st));
    TRYH(hipEventSynchronize(B->ev[7]));
    float ms = 0.0f;
    TRYH(hipEventElapsedTime(&ms, B->ev[6], B->ev[7]));
    *ms_avg = ms / (float)reps;
    return 0;
}

}  // namespace lbk

extern "C" {

int lb_bake_prepare(lb_gpu *gpu, const lb_scene_desc *scene, const lb_bake_options *options,
                    lb_bake **out) {
    if (!gpu || !scene || !out) return fail("lb_bake_prepare: null handle");
    *out = nullptr;
    lb_bake_options opt{LB_FORM_PLAIN, 0, 0, nullptr, 0};
    if (options) opt = *options;
    if (opt.form != LB_FORM_PLAIN && opt.form != LB_FORM_TUNED && opt.form != LB_FORM_WATERTIGHT)
        return fail("lb_bake_prepare: no form %d", opt.form);
    auto *b = new lb_bake();
    if (bake_prepare(gpu, *scene, opt, b)) {
        delete b;
        return 1;
    }
    *out = b;
    return 0;
}

int lb_bake_execute(lb_bake *bake, lb_stats *per_pass) {
    if (!bake) return fail("lb_bake_execute: null handle");
    if (per_pass) std::memset(per_pass, 0, 3 * sizeof(lb_stats));
    return bake_execute(bake, per_pass);
}

int lb_bake_finish(lb_bake *bake) {
    if (!bake) return fail("lb_bake_finish: null handle");
    return bake_finish(bake);
}

int lb_bake_time_pass(lb_bake *bake, int pass, int reps, float *ms_avg) {
    if (!bake || !ms_avg) return fail("lb_bake_time_pass: null handle");
    if (pass < LB_PASS_DIRECT || pass > LB_PASS_VOLUME) return fail("lb_bake_time_pass: no pass %d", pass);
    return bake_time_pass(bake, pass, reps, ms_avg);
}

int lb_bake_run(lb_gpu *gpu, const lb_scene_desc *scene, const lb_bake_options *options,
                lb_bake **out) {
    if (lb_bake_prepare(gpu, scene, options, out)) return 1;
    if (bake_execute(*out, nullptr) || bake_finish(*out)) {
        lb_bake_destroy(*out);
        *out = nullptr;
        return 1;
    }
    (*out)->info.s_total += (*out)->info.s_execute + (*out)->info.s_finish;
    return 0;
}

void lb_bake_destroy(lb_bake *bake) {
    if (bake && bake->refs.fetch_sub(1) == 1) delete bake;
}

int lb_bake_info_get(const lb_bake *bake, lb_bake_info *info) {
    if (!bake || !info) return fail("lb_bake_info_get: null handle");
    *info = bake->info;
    return 0;
}

size_t lb_bake_array(const lb_bake *b, int which, const void **data) {
    if (data) *data = nullptr;
    if (!b) return 0;
    const void *p = nullptr;
    size_t n = 0;
    // an array the bake keeps on the device, read back on the first request
    auto lazy = [&](auto &host, const DevBuf &dev, size_t count) -> int {
        using T = typename std::remove_reference_t<decltype(host)>::value_type;
        uint64_t bit = 1ull << which;
        if (!(b->held & bit)) {
            host.resize(count);
            if (count) {
                hipSetDevice(b->gpu->device);
                hipStream_t st = b->gpu->stream;
                hipError_t e = hipMemcpyAsync(host.data(), dev.ptr, count * sizeof(T),
                                              hipMemcpyDeviceToHost, st);
                if (e == hipSuccess) e = hipStreamSynchronize(st);
                if (e != hipSuccess) {
                    host.clear();
                    return hip_fail(e, "bake: reading an array back");
                }
            }
            b->held |= bit;
        }
        p = host.data();
        n = host.size() * sizeof(T);
        return 0;
    };
    size_t na = (size_t)b->aw * b->ah, ni = b->indirect.size() / 3, T = b->T_all;
    switch (which) {
    case LB_BAKE_DIRECT: p = b->direct.data(); n = b->direct.size() * 4; break;
    case LB_BAKE_INDIRECT: p = b->indirect.data(); n = b->indirect.size() * 4; break;
    case LB_BAKE_VALIDITY: p = b->validity.data(); n = b->validity.size(); break;
    case LB_BAKE_VOLUME: p = b->volume.data(); n = b->volume.size() * 4; break;
    case LB_BAKE_CHARTS: p = b->charts.data(); n = b->charts.size() * 4; break;
    case LB_BAKE_TRI_CHARTS: p = b->tri_charts.data(); n = b->tri_charts.size() * 4; break;
    case LB_BAKE_LIGHT: lazy(b->h_light, b->d_light, na * 4); break;
    case LB_BAKE_CHART_TILES: p = b->chart_tiles.data(); n = b->chart_tiles.size() * 4; break;
    case LB_BAKE_TILES: p = b->tiles.data(); n = b->tiles.size() * 4; break;
    case LB_BAKE_NODES: p = b->nodes.data(); n = b->nodes.size() * sizeof(lb_node); break;
    case LB_BAKE_TRIS: p = b->tris.data(); n = b->tris.size() * sizeof(lb_tri); break;
    case LB_BAKE_MATERIALS: p = b->mats.data(); n = b->mats.size() * 4; break;
    case LB_BAKE_PARAMS: p = &b->P; n = sizeof(lb_params); break;
    case LB_BAKE_TEXELS: lazy(b->h_texels, b->d_tex0, b->T); break;
    case LB_BAKE_TILE_KNOWN: lazy(b->h_tl_known, b->d_tl_known, T); break;
    case LB_BAKE_TILE_STATE: lazy(b->h_tl_state, b->d_tl_state, T); break;
    case LB_BAKE_TILE_FIELD: lazy(b->h_tl_field, b->d_tl_field, T); break;
    case LB_BAKE_TILE_AO: lazy(b->h_tl_ao, b->d_tl_ao, T); break;
    case LB_BAKE_TILE_DIRECT: lazy(b->h_tl_direct, b->d_tl_direct, T * 3); break;
    case LB_BAKE_TILE_INDIRECT: lazy(b->h_tl_indirect, b->d_tl_indirect, T * 3); break;
    case LB_BAKE_PROBES: p = b->probes_raw.data(); n = b->probes_raw.size() * 4; break;
    case LB_BAKE_BAKED_TILES: p = b->baked.data(); n = b->baked.size() * 4; break;
    case LB_BAKE_LIGHT0: lazy(b->h_light0, b->d_light0, b->any_clean ? na * 4 : 0); break;
    case LB_BAKE_DIRECT_BC5: lazy(b->h_bc5, b->d_bc5, b->direct.empty() ? 0 : na); break;
    case LB_BAKE_INDIRECT_RGB9E5: lazy(b->h_rgb, b->d_rgb, ni); break;
    case LB_BAKE_VOLUME_BYTES: lazy(b->h_vbytes, b->d_vbytes, b->volume.empty() ? 0 : b->np * 16); break;
    case LB_BAKE_GATHER_COST:
        lazy(b->h_cost, b->d_cost, b->ranked ? b->ng * b->P.counts[2] : 0);
        break;
    case LB_BAKE_GATHER_ORDER:
        lazy(b->h_perm, b->d_perm, b->ranked ? b->ng * b->P.counts[2] : 0);
        break;
    case LB_BAKE_GATHER_INDEX:
        if (b->gat.size() != b->ng) {
            auto *m = const_cast<lb_bake *>(b);
            m->gat.resize(b->ng);
            if (b->ng)
                hipMemcpy(m->gat.data(), b->d_gat.ptr, b->ng * 4, hipMemcpyDeviceToHost);
        }
        p = b->gat.data();
        n = b->gat.size() * 4;
        break;
    default: break;
    }
    if (!n) p = nullptr;
    if (data) *data = p;
    return n;
}

}  // extern "C"
