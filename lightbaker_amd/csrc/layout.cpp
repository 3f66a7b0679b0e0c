// Where every lit surface sits in the lightmap (crates/lightbaker/src/layout.rs), host-kept:
// charts grown by angle within a tile, each tile skyline-packed on its own, tiles shelved into the
// atlas. Everything in a fixed order, so a scene always gives the same layout.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <map>
#include <unordered_map>

#include "lb_bake.h"

namespace lbk {

namespace {

constexpr uint32_t ALIGN = 4;  // layout.rs:18

uint32_t round_up(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }

V3 vert(const lb_scene_desc &s, uint32_t mesh, uint32_t t, int k) {
    const float *p = s.meshes[mesh].positions + ((size_t)t * 3 + k) * 3;
    return {p[0], p[1], p[2]};
}

// layout.rs:342-345
V3 box_centre(const lb_scene_desc &s, uint32_t mesh, uint32_t t) {
    V3 a = vert(s, mesh, t, 0), b = vert(s, mesh, t, 1), c = vert(s, mesh, t, 2);
    return (vmin(vmin(a, b), c) + vmax(vmax(a, b), c)) * 0.5f;
}

// layout.rs:347-349 (`as i32` saturates)
int32_t key_of(float v, float size) {
    float f = std::floor(v / size);
    if (!(f == f)) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483647.0f) return INT32_MAX;
    return (int32_t)f;
}

struct KeyHash {
    size_t operator()(const std::array<int64_t, 3> &k) const {
        uint64_t h = (uint64_t)k[0] * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)k[1] + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)k[2] + 0x85EBCA77C2B2AE63ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

int64_t vkey(float v) {
    // (v * 1e4).round() as i64: half away from zero, saturating
    double r = std::round((double)(v * 1e4f));
    if (r >= 9.2e18) return INT64_MAX;
    if (r <= -9.2e18) return INT64_MIN;
    return (int64_t)r;
}

// layout.rs:131-171
std::vector<std::vector<uint32_t>> grow(const lb_scene_desc &s, uint32_t mesh, float cosang,
                                        float tile) {
    uint32_t n = s.meshes[mesh].triangles;
    std::vector<V3> normals(n);
    std::vector<std::pair<int32_t, int32_t>> keys(n);
    for (uint32_t t = 0; t < n; t++) {
        normals[t] = face_normal(s, mesh, t);
        V3 c = box_centre(s, mesh, t);
        keys[t] = {key_of(c.x, tile), key_of(c.z, tile)};
    }
    std::unordered_map<std::array<int64_t, 3>, std::vector<uint32_t>, KeyHash> at;
    auto key = [&](V3 v) { return std::array<int64_t, 3>{vkey(v.x), vkey(v.y), vkey(v.z)}; };
    for (uint32_t t = 0; t < n; t++)
        for (int k = 0; k < 3; k++) at[key(vert(s, mesh, t, k))].push_back(t);
    std::vector<uint8_t> taken(n, 0);
    std::vector<std::vector<uint32_t>> out;
    for (uint32_t seed = 0; seed < n; seed++) {
        if (taken[seed]) continue;
        taken[seed] = 1;
        std::vector<uint32_t> group{seed};
        V3 sn = normals[seed];
        auto sk = keys[seed];
        for (size_t i = 0; i < group.size(); i++) {
            uint32_t t = group[i];
            for (int k = 0; k < 3; k++) {
                for (uint32_t o : at[key(vert(s, mesh, t, k))]) {
                    if (taken[o] || keys[o] != sk || dot(normals[o], sn) < cosang) continue;
                    taken[o] = 1;
                    group.push_back(o);
                }
            }
        }
        std::sort(group.begin(), group.end());
        out.push_back(std::move(group));
    }
    return out;
}

// layout.rs:173-196
Chart make_chart(const lb_scene_desc &s, uint32_t mesh, std::vector<uint32_t> tris, float tile,
                 const std::map<std::pair<int32_t, int32_t>, float> &densities, float density,
                 uint32_t padding) {
    V3 sum{0, 0, 0};
    for (uint32_t t : tris) {
        V3 p0 = vert(s, mesh, t, 0), p1 = vert(s, mesh, t, 1), p2 = vert(s, mesh, t, 2);
        sum = sum + cross(p2 - p1, p0 - p1);
    }
    V3 n;
    if (!try_normalize(sum, &n)) n = face_normal(s, mesh, tris[0]);
    V3 u = std::fabs(n.y) > 0.9f ? V3{1, 0, 0} : cross(V3{0, 1, 0}, n);
    V3 uu = u - n * dot(u, n);
    u = uu * (1.0f / std::sqrt(dot(uu, uu)));
    V3 v = cross(n, u);
    V2 mn{INFINITY, INFINITY}, mx{-INFINITY, -INFINITY};
    for (uint32_t t : tris)
        for (int k = 0; k < 3; k++) {
            V3 p = vert(s, mesh, t, k);
            float qx = dot(p, u), qy = dot(p, v);
            mn = {std::fmin(mn.x, qx), std::fmin(mn.y, qy)};
            mx = {std::fmax(mx.x, qx), std::fmax(mx.y, qy)};
        }
    V3 c = box_centre(s, mesh, tris[0]);
    int32_t kx = key_of(c.x, tile), kz = key_of(c.z, tile);
    auto d = densities.find({kx, kz});
    if (d != densities.end()) density = d->second;
    auto size = [&](float e) { return (uint32_t)std::ceil(e * density) + 1 + 2 * padding; };
    Chart ch;
    ch.mesh = mesh;
    ch.normal = n;
    ch.u = u;
    ch.v = v;
    ch.min = mn;
    ch.max = mx;
    ch.tile_x = kx;
    ch.tile_z = kz;
    ch.density = density;
    ch.w = size(mx.x - mn.x);
    ch.h = size(mx.y - mn.y);
    ch.tris = std::move(tris);
    return ch;
}

// layout.rs:263-331
struct Skyline {
    struct Seg {
        uint32_t x, w, top;
    };
    std::vector<Seg> segs;
    uint32_t width;
    explicit Skyline(uint32_t w) : segs{{0, w, 0}}, width(w) {}
    // the lowest place (then the leftmost) for a w-wide rectangle
    bool find(uint32_t w, uint32_t *ox, uint32_t *oy) const {
        if (w > width) return false;
        bool have = false;
        uint32_t bx = 0, by = 0;
        for (size_t i = 0; i < segs.size(); i++) {
            uint32_t x = segs[i].x;
            if (x + w > width) break;
            uint32_t top = 0, reach = 0;
            for (size_t j = i; j < segs.size(); j++) {
                top = std::max(top, segs[j].top);
                reach += segs[j].w;
                if (reach >= w) break;
            }
            if (!have || top < by) {
                have = true;
                bx = x;
                by = top;
            }
        }
        *ox = bx;
        *oy = by;
        return have;
    }
    void place(uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
        std::vector<Seg> next;
        next.reserve(segs.size() + 2);
        for (auto &s : segs) {
            uint32_t end = s.x + s.w;
            if (end <= x || s.x >= x + w) {
                next.push_back(s);
                continue;
            }
            if (s.x < x) next.push_back({s.x, x - s.x, s.top});
            if (end > x + w) next.push_back({x + w, end - x - w, s.top});
        }
        next.push_back({x, w, y + h});
        std::stable_sort(next.begin(), next.end(), [](const Seg &a, const Seg &b) { return a.x < b.x; });
        std::vector<Seg> merged;
        for (auto &s : next) {
            if (!merged.empty() && merged.back().top == s.top && merged.back().x + merged.back().w == s.x)
                merged.back().w += s.w;
            else
                merged.push_back(s);
        }
        segs = std::move(merged);
    }
    uint32_t top() const {
        uint32_t t = 0;
        for (auto &s : segs) t = std::max(t, s.top);
        return t;
    }
};

// layout.rs:200-238
Tile pack_tile(int32_t kx, int32_t kz, const std::vector<uint32_t> &ids, std::vector<Chart> &charts) {
    for (uint32_t i : ids) {
        Chart &c = charts[i];
        if (c.h > c.w) {
            c.turned = true;
            std::swap(c.w, c.h);
        }
    }
    std::vector<uint32_t> order = ids;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const Chart &ca = charts[a], &cb = charts[b];
        if (ca.h != cb.h) return ca.h > cb.h;
        if (ca.w != cb.w) return ca.w > cb.w;
        return a < b;
    });
    uint64_t area = 0;
    uint32_t widest = 1;
    bool any = false;
    for (uint32_t i : ids) {
        area += (uint64_t)charts[i].w * charts[i].h;
        widest = any ? std::max(widest, charts[i].w) : charts[i].w;
        any = true;
    }
    uint32_t sq = (uint32_t)std::ceil(std::sqrt((double)area * 1.1));
    uint32_t width = round_up(std::max(widest, sq), ALIGN);
    Skyline sky(width);
    for (uint32_t i : order) {
        Chart &c = charts[i];
        uint32_t fx = 0, fy = 0, ux = 0, uy = 0;
        bool flat = sky.find(c.w, &fx, &fy);
        bool up = (c.h <= width && c.w != c.h) ? sky.find(c.h, &ux, &uy) : false;
        uint32_t x, y;
        bool turn;
        if (flat && up && uy + c.w < fy + c.h) {
            x = ux; y = uy; turn = true;
        } else if (flat) {
            x = fx; y = fy; turn = false;
        } else {
            x = ux; y = uy; turn = true;  // a chart no wider than its tile always fits
        }
        if (turn) {
            c.turned = !c.turned;
            std::swap(c.w, c.h);
        }
        c.x = x;
        c.y = y;
        sky.place(x, y, c.w, c.h);
    }
    Tile t;
    t.kx = kx;
    t.kz = kz;
    t.charts = ids;
    t.w = width;
    t.h = round_up(std::max(sky.top(), 1u), ALIGN);
    return t;
}

}  // namespace

V3 face_normal(const lb_scene_desc &s, uint32_t mesh, uint32_t t) {
    // layout.rs:333-335
    V3 p0 = vert(s, mesh, t, 0), p1 = vert(s, mesh, t, 1), p2 = vert(s, mesh, t, 2);
    V3 n;
    if (!try_normalize(cross(p2 - p1, p0 - p1), &n)) return V3{0, 1, 0};
    return n;
}

V2 chart_local(const Chart &c, V3 p, uint32_t padding) {
    float density = c.density;
    V2 a{(dot(p, c.u) - c.min.x) * density, (dot(p, c.v) - c.min.y) * density};
    if (c.turned) a = {a.y, a.x};
    float off = (float)padding + 0.5f;
    return {((float)c.x + off) + a.x, ((float)c.y + off) + a.y};
}

void chart_affine(const Chart &c, const Tile &t, uint32_t padding, float out[8]) {
    float density = c.density;
    float off = (float)padding + 0.5f;
    V2 o{(float)(t.x + c.x) + off, (float)(t.y + c.y) + off};
    V3 ax = c.turned ? c.v : c.u, ay = c.turned ? c.u : c.v;
    float mx = c.turned ? c.min.y : c.min.x, my = c.turned ? c.min.x : c.min.y;
    V3 cx = ax * density, cy = ay * density;
    out[0] = cx.x; out[1] = cx.y; out[2] = cx.z; out[3] = o.x - mx * density;
    out[4] = cy.x; out[5] = cy.y; out[6] = cy.z; out[7] = o.y - my * density;
}

bool check_tile_density(const lb_scene_desc &s, std::string *err) {
    if (s.n_tile_density && !s.tile_density) {
        *err = "lightmap: per-tile densities: null array";
        return false;
    }
    std::map<std::pair<int32_t, int32_t>, float> seen;
    for (size_t i = 0; i < s.n_tile_density; i++) {
        const lb_tile_density &d = s.tile_density[i];
        char buf[160];
        if (!(d.texels_per_meter > 0.0f) || !std::isfinite(d.texels_per_meter)) {
            snprintf(buf, sizeof buf, "lightmap: tile (%d, %d): texelsPerMeter %g is not positive",
                     d.x, d.z, (double)d.texels_per_meter);
            *err = buf;
            return false;
        }
        if (!seen.emplace(std::make_pair(d.x, d.z), d.texels_per_meter).second) {
            snprintf(buf, sizeof buf, "lightmap: tile (%d, %d) has two densities", d.x, d.z);
            *err = buf;
            return false;
        }
    }
    return true;
}

bool build_layout(const lb_scene_desc &s, Layout &L, std::string *err) {
    if (!check_tile_density(s, err)) return false;
    std::map<std::pair<int32_t, int32_t>, float> densities;
    for (size_t i = 0; i < s.n_tile_density; i++)
        densities[{s.tile_density[i].x, s.tile_density[i].z}] = s.tile_density[i].texels_per_meter;
    float density = s.texels_per_meter;
    uint32_t padding = s.padding;
    float cosang = std::cos(s.chart_angle_deg * (3.14159265358979323846f / 180.0f));
    L = Layout();
    L.density = density;
    L.padding = padding;
    L.tri_chart.resize(s.n_meshes);
    for (size_t mi = 0; mi < s.n_meshes; mi++) {
        L.tri_chart[mi].assign(s.meshes[mi].triangles, 0xffffffffu);
        if (!s.materials[s.meshes[mi].material].receives) continue;
        for (auto &group : grow(s, (uint32_t)mi, cosang, s.tile_size)) {
            Chart c = make_chart(s, (uint32_t)mi, std::move(group), s.tile_size, densities,
                                 density, padding);
            for (uint32_t t : c.tris) L.tri_chart[mi][t] = (uint32_t)L.charts.size();
            L.charts.push_back(std::move(c));
        }
    }
    std::map<std::pair<int32_t, int32_t>, std::vector<uint32_t>> by_tile;
    for (size_t i = 0; i < L.charts.size(); i++)
        by_tile[{L.charts[i].tile_x, L.charts[i].tile_z}].push_back((uint32_t)i);
    for (auto &kv : by_tile) L.tiles.push_back(pack_tile(kv.first.first, kv.first.second, kv.second, L.charts));
    // layout.rs:241-261: shelve, the tallest first
    std::vector<size_t> order(L.tiles.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const Tile &ta = L.tiles[a], &tb = L.tiles[b];
        if (ta.h != tb.h) return ta.h > tb.h;
        return std::make_pair(ta.kx, ta.kz) < std::make_pair(tb.kx, tb.kz);
    });
    uint32_t x = 0, y = 0, row = 0, width = s.atlas_width;
    for (size_t i : order) {
        Tile &t = L.tiles[i];
        if (t.w > width) {
            char buf[256];
            snprintf(buf, sizeof buf,
                     "tile (%d, %d) is %u texels wide, the atlas %u: lower texelsPerMeter or "
                     "tileSize, or widen the atlas",
                     t.kx, t.kz, t.w, width);
            *err = buf;
            return false;
        }
        if (x + t.w > width) {
            x = 0;
            y += row;
            row = 0;
        }
        t.x = x;
        t.y = y;
        x += t.w;
        row = std::max(row, t.h);
    }
    L.width = width;
    L.height = round_up(std::max(y + row, ALIGN), ALIGN);
    L.chart_tile.assign(L.charts.size(), 0);
    for (size_t ti = 0; ti < L.tiles.size(); ti++)
        for (uint32_t c : L.tiles[ti].charts) L.chart_tile[c] = (uint32_t)ti;
    return true;
}

}  // namespace lbk
