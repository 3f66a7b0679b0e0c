// The bake (crates/lightbaker/src/bake.rs:94-432 and volume.rs:34-148): the host keeps the
// layout, the BVH build and the rasterizer, as the reference does; every ray pass and every texel
// pass between them runs on the device, in one stream, with the bounce loop (gather, scatter,
// smooth, fill, light write) optionally replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <set>
#include <thread>
#include <unordered_set>
#include <vector>

#include "lb_bake.h"
#include "lb_host.h"
#include "lb_texel.h"

using namespace lbk;

struct lb_bake {
    Layout layout;
    lb_bake_info info{};
    std::vector<float> direct;     // [aw * ah][2]
    std::vector<float> indirect;   // [iw * ih][3]
    std::vector<uint8_t> validity;
    std::vector<float> volume;     // [probes][16]
    std::vector<float> charts;     // [charts][8]
    std::vector<uint32_t> tri_charts;
    std::vector<float> light;      // [aw * ah][4]
    std::vector<uint32_t> chart_tiles;
    std::vector<int32_t> tiles;    // [tiles][6]
};

namespace lbk {

namespace {

constexpr float LIFT = 0.01f;          // bake.rs:24
constexpr float INSIDE_REACH = 0.3f;   // bake.rs:26
constexpr float VOLUME_INSIDE = 0.5f;  // volume.rs:20

using clk = std::chrono::steady_clock;
double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

struct Dev {
    std::vector<DevBuf *> all;
    DevBuf *make() {
        all.push_back(new DevBuf());
        return all.back();
    }
    ~Dev() {
        for (auto *b : all) delete b;
    }
};

template <typename T>
int put(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    return b.upload(v.data(), v.size() * sizeof(T), s);
}

#define TRYH(expr)                                      \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)
#define TRY(expr)        \
    do {                 \
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

// volume.rs:125-147
void volume_irradiance(const std::vector<float> &probes, const uint32_t size[3], const float origin[3],
                       float step, V3 p, V3 n, float e[3], float *sun) {
    size_t nx = size[0], ny = size[1], nz = size[2];
    V3 f = p - V3{origin[0], origin[1], origin[2]};
    f = {f.x / step, f.y / step, f.z / step};
    auto cl = [](float v, size_t k) { return std::fmin(std::fmax(v, 0.0f), (float)(k - 1)); };
    float fx = cl(f.x, nx), fy = cl(f.y, ny), fz = cl(f.z, nz);
    size_t x0 = std::min((size_t)fx, nx - 2), y0 = std::min((size_t)fy, ny - 2),
           z0 = std::min((size_t)fz, nz - 2);
    float tx = fx - (float)x0, ty = fy - (float)y0, tz = fz - (float)z0;
    float p4[4][4] = {};
    for (int k = 0; k < 8; k++) {
        int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        float w = (dx ? tx : 1.0f - tx) * (dy ? ty : 1.0f - ty) * (dz ? tz : 1.0f - tz);
        const float *q = &probes[(((y0 + dy) * nz + (z0 + dz)) * nx + (x0 + dx)) * 16];
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 4; c++) p4[a][c] += q[a * 4 + c] * w;
    }
    const float PIF = 3.14159265358979323846f;
    float a0 = PIF * 0.282095f, a1 = 2.0f * PIF / 3.0f * 0.488603f;
    for (int c = 0; c < 3; c++)
        e[c] = 0.5f * a0 * p4[0][c] +
               0.5f * std::fmax(a0 * p4[0][c] + a1 * (p4[1][c] * n.x + p4[2][c] * n.y + p4[3][c] * n.z), 0.0f);
    *sun = p4[0][3];
}

}  // namespace

int bake_run(lb_gpu *gpu, const lb_scene_desc &sc, const lb_bake_options &opt, lb_bake *B) {
    auto t_start = clk::now();
    hipStream_t st = gpu->stream;
    hipSetDevice(gpu->device);
    lb_bake_info &I = B->info;
    if (!sc.materials || !sc.n_materials) return fail("scene: no materials");
    for (size_t m = 0; m < sc.n_meshes; m++) {
        if (sc.meshes[m].material >= sc.n_materials)
            return fail("mesh %zu: no material %u", m, sc.meshes[m].material);
        if (sc.meshes[m].triangles && !sc.meshes[m].positions) return fail("mesh %zu: no positions", m);
    }
    uint32_t g = (uint32_t)std::sqrt((float)sc.sun_samples);
    if (g * g != sc.sun_samples || g < 1) return fail("bake: sunSamples must be a square number");
    if (sc.sun_samples > 24) return fail("bake: sunSamples above 24 would reach the pick bits");
    if (sc.ao_rays < 1 || sc.gi_rays < 1) return fail("bake: aoRays and giRays at least one");
    if (sc.indirect_downscale != 1 && sc.indirect_downscale != 2 && sc.indirect_downscale != 4)
        return fail("lightmap: indirectDownscale 1, 2 or 4 (the tiles' grid)");
    if (sc.atlas_width < 64 || sc.atlas_width % 4 || sc.atlas_width > 65535)
        return fail("lightmap: atlasWidth must be a multiple of 4 in 64..65535");

    // layout, charts, BVH (bake.rs:98-104)
    auto t0 = clk::now();
    Layout &L = B->layout;
    std::string err;
    if (!build_layout(sc, L, &err)) return fail("%s", err.c_str());
    if (L.height > 65535) return fail("layout: atlas %u texels high exceeds 65535", L.height);
    I.s_layout = secs(t0);
    uint32_t aw = L.width, ah = L.height;
    size_t n_charts = L.charts.size();
    B->charts.assign(n_charts * 8, 0.0f);
    for (size_t c = 0; c < n_charts; c++)
        chart_affine(L.charts[c], L.tiles[L.chart_tile[c]], L.density, L.padding, &B->charts[c * 8]);
    t0 = clk::now();
    size_t n_tris = 0;
    for (size_t m = 0; m < sc.n_meshes; m++) n_tris += sc.meshes[m].triangles;
    std::vector<float> pos(n_tris * 9);
    std::vector<uint32_t> tmat(n_tris), tchart(n_tris);
    B->tri_charts.resize(n_tris);
    {
        size_t k = 0;
        for (size_t m = 0; m < sc.n_meshes; m++) {
            const lb_mesh_desc &md = sc.meshes[m];
            std::memcpy(&pos[k * 9], md.positions, (size_t)md.triangles * 9 * sizeof(float));
            for (uint32_t t = 0; t < md.triangles; t++, k++) {
                tmat[k] = md.material;
                tchart[k] = L.tri_chart[m][t];
            }
        }
    }
    B->tri_charts = tchart;
    std::vector<uint8_t> casts(sc.n_materials);
    std::vector<float> mats(sc.n_materials * 4);
    for (size_t m = 0; m < sc.n_materials; m++) {
        casts[m] = sc.materials[m].casts ? 1 : 0;
        for (int c = 0; c < 3; c++) mats[m * 4 + c] = sc.materials[m].albedo[c];
        mats[m * 4 + 3] = sc.materials[m].opacity;
    }
    std::vector<lb_node> nodes;
    std::vector<lb_tri> tris;
    build_bvh(pos.data(), n_tris, tmat.data(), tchart.data(), casts.data(), nodes, tris);
    I.bvh_nodes = (uint32_t)nodes.size();
    I.bvh_tris = (uint32_t)tris.size();
    I.s_bvh = secs(t0);
    lb_scene *scene = nullptr;
    if (lb_scene_create(gpu, nodes.data(), nodes.size(), tris.data(), tris.size(), mats.data(),
                        sc.n_materials, n_charts ? B->charts.data() : nullptr, n_charts, &scene))
        return 1;
    struct SceneGuard {
        lb_scene *s;
        ~SceneGuard() { lb_scene_destroy(s); }
    } sg{scene};

    // Params (bake.rs:105-120)
    V3 sun{sc.sun_direction[0], sc.sun_direction[1], sc.sun_direction[2]};
    sun = sun * (1.0f / std::sqrt(dot(sun, sun)));
    const float PIF = 3.14159265358979323846f;
    lb_params P{};
    P.sun_dir[0] = sun.x; P.sun_dir[1] = sun.y; P.sun_dir[2] = sun.z;
    P.sun_dir[3] = std::tan(sc.sun_radius_deg * (PIF / 180.0f));
    for (int c = 0; c < 3; c++) {
        P.sun_irr[c] = sc.sun_irradiance[c];
        P.zenith[c] = sc.zenith[c];
        P.horizon[c] = sc.horizon[c];
        P.ground[c] = sc.ground[c];
        P.fallback[c] = (sc.zenith[c] + sc.horizon[c]) * 0.5f * PIF;
    }
    P.ground[3] = std::fmin(INSIDE_REACH, sc.ao_reach);
    P.counts[0] = g; P.counts[1] = sc.ao_rays; P.counts[2] = sc.gi_rays; P.counts[3] = 0;
    P.reach[0] = sc.ao_reach; P.reach[1] = sc.gi_reach; P.reach[2] = LIFT; P.reach[3] = 0.0f;
    P.atlas[0] = aw; P.atlas[1] = ah; P.atlas[2] = 0;
    std::memcpy(&P.atlas[3], &sc.inside_share, 4);

    // the tiles' texels (bake.rs:139-148): every tile is dirty
    t0 = clk::now();
    size_t nt = L.tiles.size();
    I.tiles = I.tiles_baked = (uint32_t)nt;
    I.charts = (uint32_t)n_charts;
    std::vector<std::vector<lb_texel>> tex(nt);
    {
        unsigned hw = opt.threads > 0 ? (unsigned)opt.threads : std::max(1u, std::thread::hardware_concurrency());
        std::vector<std::thread> pool;
        std::atomic<size_t> next{0};
        for (unsigned w = 0; w < std::min<size_t>(hw, std::max<size_t>(nt, 1)); w++)
            pool.emplace_back([&] {
                for (size_t t; (t = next.fetch_add(1)) < nt;) rasterize(sc, L, L.tiles[t], tex[t]);
            });
        for (auto &th : pool) th.join();
    }
    std::vector<uint32_t> toff(nt + 1, 0);
    for (size_t t = 0; t < nt; t++) toff[t + 1] = toff[t] + L.tiles[t].w * L.tiles[t].h;
    size_t T = toff[nt];
    if (T > 0x7fffffffu) return fail("bake: %zu texels exceed 31-bit indices", T);
    std::vector<lb_texel> texels(T);
    std::vector<uint32_t> tile_of(T), chart(T), cov;
    std::vector<float> normal(T * 4, 0.0f);
    std::vector<uint8_t> covered(T);
    std::vector<uint4> tiles(std::max<size_t>(nt, 1));
    for (size_t t = 0; t < nt; t++) {
        const Tile &tl = L.tiles[t];
        tiles[t] = make_uint4(toff[t], tl.w, tl.h, tl.x | (tl.y << 16));
        std::memcpy(&texels[toff[t]], tex[t].data(), tex[t].size() * sizeof(lb_texel));
        for (uint32_t i = 0; i < tl.w * tl.h; i++) {
            size_t gi = toff[t] + i;
            const lb_texel &x = tex[t][i];
            tile_of[gi] = (uint32_t)t;
            bool cv = x.pos[3] > 0.5f;
            covered[gi] = cv;
            uint32_t c;
            std::memcpy(&c, &x.normal[3], 4);
            chart[gi] = cv ? c : 0xffffffffu;
            for (int k = 0; k < 3; k++) normal[gi * 4 + k] = x.normal[k];
            if (cv) cov.push_back((uint32_t)gi);
        }
    }
    tex.clear();
    I.texels = T;
    I.texels_baked = cov.size();
    I.s_raster = secs(t0);

    // device state
    Dev D;
    DevBuf &d_tex = *D.make(), &d_tile_of = *D.make(), &d_tiles = *D.make(), &d_chart = *D.make(),
           &d_normal = *D.make(), &d_valid = *D.make(), &d_known = *D.make(), &d_state = *D.make(),
           &d_hard = *D.make(), &d_soft = *D.make(), &d_ao = *D.make(), &d_field = *D.make(),
           &d_direct = *D.make(), &d_indirect = *D.make(), &d_cov = *D.make(),
           &d_items = *D.make(), &d_res = *D.make(), &d_light = *D.make(), &d_w1 = *D.make(),
           &d_wg = *D.make(), &d_f1 = *D.make(), &d_f4 = *D.make(), &d_m1 = *D.make(),
           &d_m2 = *D.make(), &d_covered = *D.make(), &d_gat = *D.make(), &d_gitems = *D.make(),
           &d_gres = *D.make();
    size_t Tn = std::max<size_t>(T, 1);
    TRY(put(d_tex, texels, st));
    TRY(put(d_tile_of, tile_of, st));
    TRY(put(d_tiles, tiles, st));
    TRY(put(d_chart, chart, st));
    TRY(put(d_normal, normal, st));
    TRY(put(d_covered, covered, st));
    TRY(d_valid.reserve(Tn));
    TRY(d_known.reserve(Tn));
    TRY(d_state.reserve(Tn));
    TRY(d_hard.reserve(Tn * 4));
    TRY(d_soft.reserve(Tn * 4));
    TRY(d_ao.reserve(Tn * 4));
    TRY(d_field.reserve(Tn * 4));
    TRY(d_direct.reserve(Tn * 16));
    TRY(d_indirect.reserve(Tn * 16));
    TRY(d_f1.reserve(Tn * 4));
    TRY(d_f4.reserve(Tn * 16));
    TRY(d_m1.reserve(Tn));
    TRY(d_m2.reserve(Tn));
    TRY(put(d_cov, cov, st));
    size_t ncov = cov.size();
    TRY(d_items.reserve(std::max<size_t>(ncov, 1) * sizeof(lb_texel)));
    TRY(d_res.reserve(std::max<size_t>(ncov, 1) * 16));
    TRY(d_light.reserve((size_t)aw * ah * 16));
    std::vector<float> w1 = gauss(1);
    TRY(put(d_w1, w1, st));
    int rg = (int)(sc.gi_filter * sc.indirect_downscale);
    std::vector<float> wg = gauss(std::max(rg, 0));
    TRY(put(d_wg, wg, st));

    TexelView V;
    V.n = (uint32_t)T;
    V.tile_of = (const uint32_t *)d_tile_of.ptr;
    V.tiles = (const uint4 *)d_tiles.ptr;
    V.texels = (lb_texel *)d_tex.ptr;
    V.chart = (const uint32_t *)d_chart.ptr;
    V.normal = (const float4 *)d_normal.ptr;
    V.valid = (uint8_t *)d_valid.ptr;
    V.known = (uint8_t *)d_known.ptr;
    V.state = (uint8_t *)d_state.ptr;
    V.hard = (uint32_t *)d_hard.ptr;
    V.soft = (float *)d_soft.ptr;
    V.ao = (float *)d_ao.ptr;
    V.field = (float *)d_field.ptr;
    V.direct = (float4 *)d_direct.ptr;
    V.indirect = (float4 *)d_indirect.ptr;

    // initial per-texel state (bake.rs:204-205: soft 0, ao 1, hard 0, state none; valid = covered)
    TRYH(hipMemsetAsync(d_light.ptr, 0, (size_t)aw * ah * 16, st));
    TRYH(hipMemcpyAsync(d_valid.ptr, d_covered.ptr, T, hipMemcpyDeviceToDevice, st));
    TRYH(hipMemsetAsync(d_state.ptr, 0, Tn, st));
    TRYH(hipMemsetAsync(d_hard.ptr, 0, Tn * 4, st));
    TRYH(hipMemsetAsync(d_soft.ptr, 0, Tn * 4, st));
    TRYH(hipMemsetAsync(d_indirect.ptr, 0, Tn * 16, st));
    tp_fill_f32(V.ao, 1.0f, (uint32_t)T, st);

    hipEvent_t ev[8];
    for (auto &e : ev) TRYH(hipEventCreate(&e));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 8; i++) hipEventDestroy(e[i]);
        }
    } eg{ev};

    // the direct pass (bake.rs:149-151)
    TRYH(hipEventRecord(ev[0], st));
    tp_collect(V.texels, (const uint32_t *)d_cov.ptr, (uint32_t)ncov, (lb_texel *)d_items.ptr, st);
    TRY(run_pass(gpu, LB_PASS_DIRECT, opt.form, scene, &P, (const lb_texel *)d_items.ptr, ncov, 0,
                 nullptr, 0, (float *)d_res.ptr, false));
    TRYH(hipEventRecord(ev[1], st));

    // decode, smooth, field, direct light, fill (bake.rs:196-272)
    DecodeArgs da{sc.inside_share, sc.ao_strength, sc.ao_contrast, sc.ao_floor};
    tp_decode(V, (const float4 *)d_res.ptr, (const uint32_t *)d_cov.ptr, (uint32_t)ncov, da, st);
    // ao and soft smoothed over the valid texels, radius 1
    tp_smooth(V, 1, V.ao, (float *)d_f1.ptr, V.valid, 1, (const float *)d_w1.ptr, st);
    tp_copy_f32((const float *)d_f1.ptr, V.ao, (uint32_t)T, st);
    tp_smooth(V, 1, V.soft, (float *)d_f1.ptr, V.valid, 1, (const float *)d_w1.ptr, st);
    tp_sun_field(V, V.valid, g, sc.sdf_radius, st);
    tp_direct_light(V, (const float *)d_f1.ptr, make_float3(sc.sun_irradiance[0], sc.sun_irradiance[1], sc.sun_irradiance[2]),
                    make_float3(sun.x, sun.y, sun.z), st);
    // fill ao, field, direct from valid, padding + 4 rings (bake.rs:243,253-256)
    uint32_t rings = L.padding + 4;
    auto fill = [&](int channels, float *v, float *tmp, uint32_t nrings, const uint8_t *valid0,
                    uint8_t *final_mask) -> int {
        uint8_t *ma = (uint8_t *)d_m1.ptr, *mb = (uint8_t *)d_m2.ptr;
        tp_copy_u8(valid0, ma, (uint32_t)T, st);
        float *a = v, *b = tmp;
        for (uint32_t r = 0; r < nrings; r++) {
            tp_fill_ring(V, channels, a, b, ma, mb, st);
            std::swap(a, b);
            std::swap(ma, mb);
        }
        if (a != v) tp_copy_f32(a, v, (uint32_t)(T * (channels == 1 ? 1 : 4)), st);
        if (final_mask) tp_copy_u8(ma, final_mask, (uint32_t)T, st);
        return 0;
    };
    TRY(fill(1, V.ao, (float *)d_f1.ptr, rings, V.valid, nullptr));
    TRY(fill(1, V.field, (float *)d_f1.ptr, rings, V.valid, nullptr));
    DevBuf &d_filled = *D.make();
    TRY(d_filled.reserve(Tn));
    uint8_t *filled = (uint8_t *)d_filled.ptr;
    TRY(fill(3, (float *)V.direct, (float *)d_f4.ptr, rings, V.valid, filled));
    tp_known(V, filled, st);
    // the light atlas every tile left (bake.rs:277-281)
    tp_write_light(V, (float4 *)d_light.ptr, aw, st);
    TRYH(hipEventRecord(ev[2], st));

    // the gather set (bake.rs:282-299), on the host from the valid texels
    std::vector<uint8_t> valid(T);
    TRYH(hipMemcpyAsync(valid.data(), d_valid.ptr, T, hipMemcpyDeviceToHost, st));
    TRYH(hipStreamSynchronize(st));
    t0 = clk::now();
    uint32_t ds = sc.indirect_downscale;
    std::vector<uint8_t> gathered(T, 0);
    std::vector<uint32_t> gat;
    for (size_t t = 0; t < nt; t++) {
        uint32_t tw = L.tiles[t].w, size = L.tiles[t].w * L.tiles[t].h, o = toff[t];
        uint32_t bw = (tw + ds - 1) / ds;
        auto block = [&](uint32_t i) { return (uint64_t)((i / tw) / ds) * bw + (i % tw) / ds; };
        std::set<std::pair<uint64_t, uint32_t>> seen;
        for (uint32_t i = 0; i < size; i++) {
            bool gt = valid[o + i] && (i % tw) % ds == 0 && (i / tw) % ds == 0;
            gathered[o + i] = gt;
            if (gt) seen.insert({block(i), chart[o + i]});
        }
        for (uint32_t i = 0; i < size; i++)
            if (valid[o + i] && seen.insert({block(i), chart[o + i]}).second) gathered[o + i] = 1;
        for (uint32_t i = 0; i < size; i++)
            if (gathered[o + i]) gat.push_back(o + i);
    }
    I.gather_items = gat.size();
    I.s_gather_set = secs(t0);
    size_t ng = gat.size();
    TRY(put(d_gat, gat, st));
    DevBuf &d_gmask = *D.make();
    TRY(put(d_gmask, gathered, st));
    TRY(d_gitems.reserve(std::max<size_t>(ng, 1) * sizeof(lb_texel)));
    TRY(d_gres.reserve(std::max<size_t>(ng, 1) * 16));
    tp_collect(V.texels, (const uint32_t *)d_gat.ptr, (uint32_t)ng, (lb_texel *)d_gitems.ptr, st);

    // the bounces (bake.rs:300-330)
    uint32_t grings = L.padding + 2 * ds + 2;
    auto bounce = [&]() -> int {
        TRY(run_pass(gpu, LB_PASS_GATHER, opt.form, scene, &P, (const lb_texel *)d_gitems.ptr, ng,
                     0, (const float *)d_light.ptr, (size_t)aw * ah, (float *)d_gres.ptr, false));
        // e: the gathered items' light at their texels, 0 elsewhere (into `indirect`)
        tp_fill_f32((float *)V.indirect, 0.0f, (uint32_t)(T * 4), st);
        tp_scatter((const float4 *)d_gres.ptr, (const uint32_t *)d_gat.ptr, (uint32_t)ng, V.indirect, st);
        tp_smooth(V, 4, (const float *)V.indirect, (float *)d_f4.ptr, (const uint8_t *)d_gmask.ptr, rg,
                  (const float *)d_wg.ptr, st);
        float *e = (float *)d_f4.ptr, *tmp = (float *)V.indirect;
        // fill into d_f4, ping-ponging through `indirect`, then the result to `indirect`
        uint8_t *ma = (uint8_t *)d_m1.ptr, *mb = (uint8_t *)d_m2.ptr;
        tp_copy_u8((const uint8_t *)d_gmask.ptr, ma, (uint32_t)T, st);
        float *a = e, *b = tmp;
        for (uint32_t r = 0; r < grings; r++) {
            tp_fill_ring(V, 3, a, b, ma, mb, st);
            std::swap(a, b);
            std::swap(ma, mb);
        }
        if (a != (float *)V.indirect) tp_copy_f32(a, (float *)V.indirect, (uint32_t)(T * 4), st);
        tp_write_light(V, (float4 *)d_light.ptr, aw, st);
        return hipGetLastError() == hipSuccess ? 0 : fail("GPU: bounce launch failed");
    };
    TRYH(hipEventRecord(ev[3], st));
    if (sc.gi_bounces > 0 && ng > 0) {
        if (opt.graph) {
            // one bounce captured once, replayed gi_bounces times: the same buffers each time
            if (opt.form == LB_FORM_TUNED) {
                // size the tuned form's spill area outside the capture (no allocation inside)
                if (tuned_prepare(gpu, LB_PASS_GATHER, P, (uint32_t)ng) != hipSuccess)
                    return fail("GPU: tuned spill area");
            }
            hipGraph_t graph;
            hipGraphExec_t exec;
            TRYH(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            int rc = bounce();
            hipError_t ce = hipStreamEndCapture(st, &graph);
            if (rc) return 1;
            if (ce != hipSuccess) return hip_fail(ce, "graph capture");
            TRYH(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            for (uint32_t b = 0; b < sc.gi_bounces; b++) TRYH(hipGraphLaunch(exec, st));
            TRYH(hipStreamSynchronize(st));
            hipGraphExecDestroy(exec);
            hipGraphDestroy(graph);
        } else {
            for (uint32_t b = 0; b < sc.gi_bounces; b++) TRY(bounce());
        }
    }
    TRYH(hipEventRecord(ev[4], st));

    // the volume (volume.rs:34-120)
    std::vector<float> probes;
    uint32_t vsz[3] = {0, 0, 0};
    if (sc.has_volume) {
        if (!(sc.volume_step > 0.0f)) return fail("volume: bad bounds or step");
        for (int a = 0; a < 3; a++) {
            if (sc.volume_max[a] < sc.volume_min[a]) return fail("volume: bad bounds or step");
            vsz[a] = std::max<uint32_t>((uint32_t)std::round((sc.volume_max[a] - sc.volume_min[a]) / sc.volume_step) + 1, 2);
        }
        size_t nx = vsz[0], ny = vsz[1], nz = vsz[2], np = nx * ny * nz;
        std::vector<lb_texel> pitems(np);
        for (size_t k = 0; k < np; k++) {
            float px = sc.volume_min[0] + (float)(k % nx) * sc.volume_step;
            float py = sc.volume_min[1] + (float)(k / (nx * nz)) * sc.volume_step;
            float pz = sc.volume_min[2] + (float)((k / nx) % nz) * sc.volume_step;
            uint32_t kk = (uint32_t)k;
            float kf;
            std::memcpy(&kf, &kk, 4);
            pitems[k] = lb_texel{{px, py, pz, 1.0f}, {0.0f, 1.0f, 0.0f, kf}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        }
        DevBuf &d_pitems = *D.make(), &d_pres = *D.make();
        TRY(put(d_pitems, pitems, st));
        TRY(d_pres.reserve(np * 64));
        lb_params PV = P;
        PV.counts[2] = sc.volume_rays;  // volume.rs:76
        if (PV.counts[2] < 1) return fail("volume: rays at least one");
        TRY(run_pass(gpu, LB_PASS_VOLUME, opt.form, scene, &PV, (const lb_texel *)d_pitems.ptr, np,
                     0, (const float *)d_light.ptr, (size_t)aw * ah, (float *)d_pres.ptr, false));
        probes.resize(np * 16);
        TRYH(hipMemcpyAsync(probes.data(), d_pres.ptr, np * 64, hipMemcpyDeviceToHost, st));
        I.probes = np;
    }
    TRYH(hipEventRecord(ev[5], st));

    // read back the tiles' light
    std::vector<uint8_t> known(T), state(T);
    std::vector<float> field(T), ao(T), dir4(T * 4), ind4(T * 4);
    B->light.resize((size_t)aw * ah * 4);
    TRYH(hipMemcpyAsync(known.data(), d_known.ptr, T, hipMemcpyDeviceToHost, st));
    TRYH(hipMemcpyAsync(state.data(), d_state.ptr, T, hipMemcpyDeviceToHost, st));
    TRYH(hipMemcpyAsync(field.data(), d_field.ptr, T * 4, hipMemcpyDeviceToHost, st));
    TRYH(hipMemcpyAsync(ao.data(), d_ao.ptr, T * 4, hipMemcpyDeviceToHost, st));
    TRYH(hipMemcpyAsync(dir4.data(), d_direct.ptr, T * 16, hipMemcpyDeviceToHost, st));
    TRYH(hipMemcpyAsync(ind4.data(), d_indirect.ptr, T * 16, hipMemcpyDeviceToHost, st));
    TRYH(hipMemcpyAsync(B->light.data(), d_light.ptr, (size_t)aw * ah * 16, hipMemcpyDeviceToHost, st));
    TRYH(hipStreamSynchronize(st));
    TRYH(hipEventElapsedTime(&I.ms_direct, ev[0], ev[1]));
    TRYH(hipEventElapsedTime(&I.ms_direct_texel, ev[1], ev[2]));
    TRYH(hipEventElapsedTime(&I.ms_bounces, ev[3], ev[4]));
    TRYH(hipEventElapsedTime(&I.ms_volume, ev[4], ev[5]));
    t0 = clk::now();

    // probes inside solids from their neighbours, 8 rings (volume.rs:86-117)
    if (sc.has_volume) {
        size_t nx = vsz[0], ny = vsz[1], nz = vsz[2], np = nx * ny * nz;
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
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= (long)nx || yy >= (long)ny || zz >= (long)nz) continue;
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
        for (int a = 0; a < 3; a++) {
            I.volume_size[a] = vsz[a];
            I.volume_origin[a] = sc.volume_min[a];
        }
        I.volume_step = sc.volume_step;
    }

    // texels with no light of their own take the volume's (bake.rs:337-359)
    std::vector<uint8_t> validity_t(state);
    if (sc.has_volume) {
        for (size_t gi = 0; gi < T; gi++) {
            if (!covered[gi] || known[gi] != 0) continue;
            V3 n{normal[gi * 4], normal[gi * 4 + 1], normal[gi * 4 + 2]};
            V3 p{texels[gi].pos[0], texels[gi].pos[1], texels[gi].pos[2]};
            float e[3], s;
            volume_irradiance(probes, vsz, sc.volume_min, sc.volume_step, p + n * 0.15f, n, e, &s);
            float c = std::fmax(dot(n, sun), 0.0f);
            for (int k = 0; k < 3; k++) {
                ind4[gi * 4 + k] = e[k];
                dir4[gi * 4 + k] = sc.sun_irradiance[k] * c * s;
            }
            field[gi] = s > 0.5f ? 1.0f : 0.0f;
            ao[gi] = 1.0f;
            known[gi] = 1;
            validity_t[gi] = VALIDITY_VOLUME;
        }
    }

    // the atlas from every tile (bake.rs:364-421)
    uint32_t iw = (aw + ds - 1) / ds, ih = (ah + ds - 1) / ds;
    B->direct.assign((size_t)aw * ah * 2, 1.0f);
    B->validity.assign((size_t)aw * ah, VALIDITY_NONE);
    std::vector<float> sum((size_t)iw * ih * 4, 0.0f);
    for (size_t t = 0; t < nt; t++) {
        const Tile &tl = L.tiles[t];
        for (uint32_t y = 0; y < tl.h; y++)
            for (uint32_t x = 0; x < tl.w; x++) {
                size_t gi = toff[t] + (size_t)y * tl.w + x;
                if (known[gi] == 0) continue;
                size_t ax = tl.x + x, ay = tl.y + y;
                B->direct[(ay * aw + ax) * 2] = field[gi];
                B->direct[(ay * aw + ax) * 2 + 1] = ao[gi];
                B->validity[ay * aw + ax] = validity_t[gi];
                float k = known[gi] == 2 ? 1.0f : 1e-3f;
                float *s = &sum[((ay / ds) * iw + ax / ds) * 4];
                for (int c = 0; c < 3; c++) s[c] += ind4[gi * 4 + c] * k;
                s[3] += k;
            }
    }
    B->indirect.assign((size_t)iw * ih * 3, 0.0f);
    std::vector<uint8_t> has((size_t)iw * ih);
    for (size_t i = 0; i < (size_t)iw * ih; i++) {
        has[i] = sum[i * 4 + 3] > 0.0f;
        if (has[i])
            for (int c = 0; c < 3; c++) B->indirect[i * 3 + c] = sum[i * 4 + c] / sum[i * 4 + 3];
    }
    for (int ring = 0; ring < 2; ring++) {
        std::vector<uint8_t> was = has;
        std::vector<float> src = B->indirect;
        for (uint32_t y = 0; y < ih; y++)
            for (uint32_t x = 0; x < iw; x++) {
                size_t i = (size_t)y * iw + x;
                if (was[i]) continue;
                float acc[3] = {0, 0, 0};
                int n = 0;
                for (uint32_t yy = y > 0 ? y - 1 : 0; yy < std::min(y + 2, ih); yy++)
                    for (uint32_t xx = x > 0 ? x - 1 : 0; xx < std::min(x + 2, iw); xx++) {
                        size_t j = (size_t)yy * iw + xx;
                        if (!was[j]) continue;
                        for (int c = 0; c < 3; c++) acc[c] += src[j * 3 + c];
                        n++;
                    }
                if (n > 0) {
                    for (int c = 0; c < 3; c++) B->indirect[i * 3 + c] = acc[c] / (float)n;
                    has[i] = 1;
                }
            }
    }
    B->volume = std::move(probes);
    B->chart_tiles = L.chart_tile;
    B->tiles.resize(nt * 6);
    for (size_t t = 0; t < nt; t++) {
        const Tile &tl = L.tiles[t];
        int32_t v[6] = {tl.kx, tl.kz, (int32_t)tl.w, (int32_t)tl.h, (int32_t)tl.x, (int32_t)tl.y};
        std::memcpy(&B->tiles[t * 6], v, sizeof v);
    }
    I.atlas_width = aw;
    I.atlas_height = ah;
    I.indirect_width = iw;
    I.indirect_height = ih;
    I.s_finish = secs(t0);
    I.s_total = secs(t_start);
    return 0;
}

}  // namespace lbk

extern "C" {

int lb_bake_run(lb_gpu *gpu, const lb_scene_desc *scene, const lb_bake_options *options,
                lb_bake **out) {
    if (!gpu || !scene || !out) return fail("lb_bake_run: null handle");
    *out = nullptr;
    lb_bake_options opt{LB_FORM_PLAIN, 0, 0};
    if (options) opt = *options;
    if (opt.form != LB_FORM_PLAIN && opt.form != LB_FORM_TUNED) return fail("lb_bake_run: no form %d", opt.form);
    auto *b = new lb_bake();
    if (bake_run(gpu, *scene, opt, b)) {
        hipStreamSynchronize(gpu->stream);
        delete b;
        return 1;
    }
    *out = b;
    return 0;
}

void lb_bake_destroy(lb_bake *bake) { delete bake; }

int lb_bake_info_get(const lb_bake *bake, lb_bake_info *info) {
    if (!bake || !info) return fail("lb_bake_info_get: null handle");
    *info = bake->info;
    return 0;
}

size_t lb_bake_array(const lb_bake *b, int which, const void **data) {
    if (!b) return 0;
    const void *p = nullptr;
    size_t n = 0;
    switch (which) {
    case LB_BAKE_DIRECT: p = b->direct.data(); n = b->direct.size() * 4; break;
    case LB_BAKE_INDIRECT: p = b->indirect.data(); n = b->indirect.size() * 4; break;
    case LB_BAKE_VALIDITY: p = b->validity.data(); n = b->validity.size(); break;
    case LB_BAKE_VOLUME: p = b->volume.data(); n = b->volume.size() * 4; break;
    case LB_BAKE_CHARTS: p = b->charts.data(); n = b->charts.size() * 4; break;
    case LB_BAKE_TRI_CHARTS: p = b->tri_charts.data(); n = b->tri_charts.size() * 4; break;
    case LB_BAKE_LIGHT: p = b->light.data(); n = b->light.size() * 4; break;
    case LB_BAKE_CHART_TILES: p = b->chart_tiles.data(); n = b->chart_tiles.size() * 4; break;
    case LB_BAKE_TILES: p = b->tiles.data(); n = b->tiles.size() * 4; break;
    default: break;
    }
    if (data) *data = p;
    return n;
}

}  // extern "C"
