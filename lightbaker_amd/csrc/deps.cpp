// The tiles' and probe chunks' dependency hashes (crates/lightbaker/src/deps.rs:72-102,
// bake.rs:124-128,447-477, volume.rs:36-56): the host assembles each hash's byte stream exactly as
// the reference feeds blake3 — the global hash, the seed, then for every casting mesh with a
// triangle in reach its index, its material and each such triangle's index and corners — and the
// device (lb_deps_swept) decides which triangles are in reach of every zone in one launch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/lightbaker.h"
#include "lb_bake.h"
#include "lb_blake3.h"

namespace lbk {
int fail(const char *fmt, ...);

namespace {

struct Box {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    // deps.rs:24-27 (glam's min/max: a NaN operand is ignored)
    void add(const float *p) {
        for (int a = 0; a < 3; a++) {
            mn[a] = std::fmin(mn[a], p[a]);
            mx[a] = std::fmax(mx[a], p[a]);
        }
    }
};

// Every triangle's corners (meshes in order), each one's casting flag, each mesh's first triangle,
// and deps.rs:47-49's `top`: the highest box top over every triangle of every mesh.
struct Tris {
    std::vector<float> pos;
    std::vector<uint8_t> casts;
    std::vector<size_t> first;
    float top = -INFINITY;
};

int gather_tris(const lb_scene_desc &s, Tris &T) {
    size_t n = 0;
    for (size_t m = 0; m < s.n_meshes; m++) {
        if (s.meshes[m].material >= s.n_materials) return fail("deps: mesh %zu's material", m);
        if (s.meshes[m].triangles && !s.meshes[m].positions) return fail("deps: mesh %zu", m);
        n += s.meshes[m].triangles;
    }
    T.pos.resize(n * 9);
    T.casts.resize(n);
    T.first.assign(s.n_meshes + 1, 0);
    size_t k = 0;
    for (size_t m = 0; m < s.n_meshes; m++) {
        const lb_mesh_desc &md = s.meshes[m];
        T.first[m] = k;
        if (md.triangles) std::memcpy(&T.pos[k * 9], md.positions, (size_t)md.triangles * 36);
        uint8_t c = s.materials[md.material].casts ? 1 : 0;
        for (uint32_t t = 0; t < md.triangles; t++, k++) {
            T.casts[k] = c;
            for (int v = 0; v < 3; v++) T.top = std::fmax(T.top, T.pos[k * 9 + v * 3 + 1]);
        }
    }
    T.first[s.n_meshes] = k;
    return 0;
}

// glam's normalize: v * (1 / |v|)
void sun_of(const lb_scene_desc &s, float out[3]) {
    const float *d = s.sun_direction;
    float r = 1.0f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int a = 0; a < 3; a++) out[a] = d[a] * r;
}

void put_u32(Blake3 &h, uint32_t v) {
    uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
    h.update(b, 4);
}

// deps.rs:72-102 for every zone: the hash from the global hash, each zone's seed and the hit bits
int hash_zones(lb_gpu *gpu, const lb_scene_desc &s, const Tris &T, const uint8_t global[32],
               const std::vector<Box> &own, float reach, const std::vector<std::string> &seeds,
               uint8_t *out) {
    size_t nz = own.size(), n = T.casts.size(), words = (n + 31) / 32;
    std::vector<float> zones(nz * 6);
    for (size_t z = 0; z < nz; z++)
        for (int a = 0; a < 3; a++) {
            zones[z * 6 + a] = own[z].mn[a] - reach;  // Aabb::grow (deps.rs:29-31)
            zones[z * 6 + 3 + a] = own[z].mx[a] + reach;
        }
    float sun[3];
    sun_of(s, sun);
    std::vector<uint32_t> bits(nz * words);
    if (nz && n &&
        lb_deps_swept(gpu, T.pos.data(), n, T.casts.data(), zones.data(), nz, sun, T.top, bits.data()))
        return 1;
    for (size_t z = 0; z < nz; z++) {
        Blake3 h;
        h.update(global, 32);
        h.update(seeds[z].data(), seeds[z].size());
        const uint32_t *row = &bits[z * words];
        for (size_t m = 0; m < s.n_meshes; m++) {
            if (!s.materials[s.meshes[m].material].casts) continue;
            size_t a = T.first[m], b = T.first[m + 1];
            bool any = false;
            for (size_t t = a; t < b; t++) {
                if (!(row[t >> 5] >> (t & 31) & 1u)) continue;
                if (!any) {
                    put_u32(h, (uint32_t)m);
                    put_u32(h, s.meshes[m].material);
                    any = true;
                }
                put_u32(h, (uint32_t)(t - a));
                h.update(&T.pos[t * 9], 36);  // little-endian f32 corners, as bytemuck gives them
            }
        }
        h.finalize(out + z * 32);
    }
    return 0;
}

// layout.rs:347-349 (`as i32` saturates, NaN to 0)
int32_t key_of(float v, float size) {
    float f = std::floor(v / size);
    if (!(f == f)) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483647.0f) return INT32_MAX;
    return (int32_t)f;
}

std::string le_i32(int32_t v) {
    uint32_t u = (uint32_t)v;
    char b[4] = {(char)u, (char)(u >> 8), (char)(u >> 16), (char)(u >> 24)};
    return std::string(b, 4);
}

}  // namespace
}  // namespace lbk

extern "C" {

int lb_deps_tile_hashes(lb_gpu *gpu, const lb_scene_desc *scene, const lb_layout *layout,
                        const uint8_t global[32], uint8_t *hashes) {
    using namespace lbk;
    if (!gpu || !scene || !global) return fail("deps: null argument");
    const lb_scene_desc &s = *scene;
    lb_layout own_layout;
    if (!layout) {
        std::string err;
        if (!build_layout(s, own_layout.L, &err)) return fail("%s", err.c_str());
        layout = &own_layout;
    }
    const Layout &L = layout->L;
    if (L.tiles.size() && !hashes) return fail("deps: null output");
    Tris T;
    if (gather_tris(s, T)) return 1;
    size_t nt = L.tiles.size();
    std::vector<Box> own(nt);
    std::vector<std::string> seeds(nt);
    for (size_t t = 0; t < nt; t++) {
        const Tile &tl = L.tiles[t];
        // bake.rs:447-458 tile_box and bake.rs:461-477 tile_seed over the tile's charts
        Blake3 h;
        std::string key = le_i32(tl.kx) + le_i32(tl.kz);
        h.update(key.data(), key.size());
        for (uint32_t c : tl.charts) {
            const Chart &ch = L.charts[c];
            if (ch.mesh >= s.n_meshes) return fail("deps: chart %u's mesh", c);
            size_t base = T.first[ch.mesh];
            put_u32(h, ch.mesh);
            put_u32(h, s.meshes[ch.mesh].material);
            for (uint32_t tri : ch.tris) {
                const float *p = &T.pos[(base + tri) * 9];
                for (int v = 0; v < 3; v++) own[t].add(p + v * 3);
                put_u32(h, tri);
                h.update(p, 36);
            }
        }
        // a tile laid out at a density of its own (no counterpart in the reference): the
        // density's bytes end its seed, so changing it rebakes the tile
        for (size_t k = 0; k < s.n_tile_density; k++)
            if (s.tile_density[k].x == tl.kx && s.tile_density[k].z == tl.kz)
                h.update(&s.tile_density[k].texels_per_meter, 4);
        uint8_t seed[32];
        h.finalize(seed);
        seeds[t].assign((const char *)seed, 32);
    }
    // bake.rs:127: occlusion's reach, or the bounces' where they reach further
    float reach = s.gi_bounces > 0 ? std::fmax(s.gi_reach, s.ao_reach) : s.ao_reach;
    return hash_zones(gpu, s, T, global, own, reach, seeds, hashes);
}

int lb_deps_probe_hashes(lb_gpu *gpu, const lb_scene_desc *scene, const uint8_t global[32],
                         const char *volume_json, size_t json_len, int32_t *keys, uint8_t *hashes,
                         size_t cap, size_t *n_chunks) {
    using namespace lbk;
    if (!gpu || !scene || !global || !n_chunks || (json_len && !volume_json))
        return fail("deps: null argument");
    const lb_scene_desc &s = *scene;
    *n_chunks = 0;
    if (!s.has_volume) return 0;
    if (!(s.volume_step > 0.0f)) return fail("volume: bad bounds or step");
    // volume.rs:36-45: the probes by ground-plan tile, keys in order
    uint32_t n[3];
    for (int a = 0; a < 3; a++)
        n[a] = std::max<uint32_t>(
            (uint32_t)std::round((s.volume_max[a] - s.volume_min[a]) / s.volume_step) + 1, 2);
    size_t nx = n[0], ny = n[1], nz = n[2], np = nx * ny * nz;
    std::map<std::pair<int32_t, int32_t>, Box> chunks;
    for (size_t k = 0; k < np; k++) {
        float p[3] = {s.volume_min[0] + (float)(k % nx) * s.volume_step,
                      s.volume_min[1] + (float)(k / (nx * nz)) * s.volume_step,
                      s.volume_min[2] + (float)((k / nx) % nz) * s.volume_step};
        chunks[{key_of(p[0], s.tile_size), key_of(p[2], s.tile_size)}].add(p);
    }
    *n_chunks = chunks.size();
    if (chunks.size() > cap) return fail("deps: %zu probe chunks, room for %zu", chunks.size(), cap);
    if (!keys || !hashes) return fail("deps: null output");
    Tris T;
    if (gather_tris(s, T)) return 1;
    std::vector<Box> own;
    std::vector<std::string> seeds;
    std::string json(volume_json ? volume_json : "", json_len);
    size_t i = 0;
    for (auto &kv : chunks) {
        keys[i * 2] = kv.first.first;
        keys[i * 2 + 1] = kv.first.second;
        own.push_back(kv.second);
        seeds.push_back(le_i32(kv.first.first) + le_i32(kv.first.second) + json);  // volume.rs:50-54
        i++;
    }
    return hash_zones(gpu, s, T, global, own, s.gi_reach, seeds, hashes);
}

}  // extern "C"
