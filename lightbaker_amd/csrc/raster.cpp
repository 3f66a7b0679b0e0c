// A tile's texels (crates/lightbaker/src/raster.rs), host-kept: every texel whose centre a chart's
// triangle covers, then, conservatively, every texel a triangle only touches, sampled at the
// triangle's point nearest the texel's centre.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lb_bake.h"
#include "lb_raster.h"

namespace lbk {

namespace {

inline float perp_dot(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline V2 sub2(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline float signum(float v) { return std::signbit(v) ? -1.0f : 1.0f; }

// raster.rs:103-114
V2 nearest(V2 p, V2 a, V2 b, V2 c) {
    auto side = [](V2 o, V2 e, V2 q) { return perp_dot(sub2(e, o), sub2(q, o)); };
    float s = signum(side(a, b, c));
    if (side(a, b, p) * s >= 0.0f && side(b, c, p) * s >= 0.0f && side(c, a, p) * s >= 0.0f)
        return p;
    auto seg = [&](V2 o, V2 e) {
        V2 d = sub2(e, o);
        V2 po = sub2(p, o);
        float len2 = d.x * d.x + d.y * d.y;
        float t = (po.x * d.x + po.y * d.y) / std::fmax(len2, 1e-12f);
        t = std::fmin(std::fmax(t, 0.0f), 1.0f);
        return V2{o.x + d.x * t, o.y + d.y * t};
    };
    V2 cand[3] = {seg(a, b), seg(b, c), seg(c, a)};
    int best = 0;
    float bd = 0.0f;
    for (int i = 0; i < 3; i++) {
        V2 q = sub2(cand[i], p);
        float d = q.x * q.x + q.y * q.y;
        if (i == 0 || d < bd) {
            best = i;
            bd = d;
        }
    }
    return cand[best];
}

inline float bits_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

void rasterize(const lb_scene_desc &s, const Layout &L, const Tile &tile, std::vector<lb_texel> &out) {
    int32_t w = (int32_t)tile.w, h = (int32_t)tile.h;
    out.assign((size_t)w * h, lb_texel{});
    for (int pass = 0; pass < 2; pass++) {
        bool conservative = pass == 1;
        for (uint32_t ci : tile.charts) {
            const Chart &c = L.charts[ci];
            for (uint32_t t : c.tris) {
                const float *pp = s.meshes[c.mesh].positions + (size_t)t * 9;
                V3 pa{pp[0], pp[1], pp[2]}, pb{pp[3], pp[4], pp[5]}, pc{pp[6], pp[7], pp[8]};
                V2 ua = chart_local(c, pa, L.padding);
                V2 ub = chart_local(c, pb, L.padding);
                V2 uc = chart_local(c, pc, L.padding);
                float det = (ub.y - uc.y) * (ua.x - uc.x) + (uc.x - ub.x) * (ua.y - uc.y);
                if (std::fabs(det) < 1e-12f) continue;
                auto bary = [&](V2 q, float *l1, float *l2, float *l3) {
                    *l1 = ((ub.y - uc.y) * (q.x - uc.x) + (uc.x - ub.x) * (q.y - uc.y)) / det;
                    *l2 = ((uc.y - ua.y) * (q.x - uc.x) + (ua.x - uc.x) * (q.y - uc.y)) / det;
                    *l3 = 1.0f - *l1 - *l2;
                };
                auto at = [&](V2 q) {
                    float l1, l2, l3;
                    bary(q, &l1, &l2, &l3);
                    return pa * l1 + pb * l2 + pc * l3;
                };
                V3 n = face_normal(s, c.mesh, t);
                V3 o = at(V2{0, 0});
                V3 dx = at(V2{1, 0}) - o, dy = at(V2{0, 1}) - o;
                V2 lo{std::fmax(std::floor(std::fmin(std::fmin(ua.x, ub.x), uc.x)), 0.0f),
                      std::fmax(std::floor(std::fmin(std::fmin(ua.y, ub.y), uc.y)), 0.0f)};
                V2 hi{std::fmin(std::ceil(std::fmax(std::fmax(ua.x, ub.x), uc.x)), (float)(w - 1)),
                      std::fmin(std::ceil(std::fmax(std::fmax(ua.y, ub.y), uc.y)), (float)(h - 1))};
                for (int32_t y = (int32_t)lo.y; y <= (int32_t)hi.y; y++) {
                    for (int32_t x = (int32_t)lo.x; x <= (int32_t)hi.x; x++) {
                        V2 centre{(float)x + 0.5f, (float)y + 0.5f};
                        lb_texel &tx = out[(size_t)y * w + x];
                        if (tx.pos[3] > 0.5f) continue;
                        V2 q = conservative ? nearest(centre, ua, ub, uc) : centre;
                        if (conservative &&
                            (std::fabs(q.x - centre.x) > 0.5f || std::fabs(q.y - centre.y) > 0.5f))
                            continue;
                        float l1, l2, l3;
                        bary(q, &l1, &l2, &l3);
                        if (!conservative && (l1 < -1e-4f || l2 < -1e-4f || l3 < -1e-4f)) continue;
                        V3 p = at(q);
                        tx = lb_texel{{p.x, p.y, p.z, 1.0f},
                                      {n.x, n.y, n.z, bits_f(ci)},
                                      {dx.x, dx.y, dx.z, 0.0f},
                                      {dy.x, dy.y, dy.z, 0.0f}};
                    }
                }
            }
        }
    }
}

void raster_records(const lb_scene_desc &s, const Layout &L, const std::vector<uint32_t> &tiles,
                    std::vector<RasterTri> &out) {
    out.clear();
    for (size_t k = 0; k < tiles.size(); k++) {
        const Tile &tile = L.tiles[tiles[k]];
        int32_t w = (int32_t)tile.w, h = (int32_t)tile.h;
        for (uint32_t ci : tile.charts) {
            const Chart &c = L.charts[ci];
            for (uint32_t t : c.tris) {
                const float *pp = s.meshes[c.mesh].positions + (size_t)t * 9;
                V3 pa{pp[0], pp[1], pp[2]}, pb{pp[3], pp[4], pp[5]}, pc{pp[6], pp[7], pp[8]};
                V2 ua = chart_local(c, pa, L.padding);
                V2 ub = chart_local(c, pb, L.padding);
                V2 uc = chart_local(c, pc, L.padding);
                float det = (ub.y - uc.y) * (ua.x - uc.x) + (uc.x - ub.x) * (ua.y - uc.y);
                if (std::fabs(det) < 1e-12f) continue;
                auto at = [&](V2 q) {
                    float l1 = ((ub.y - uc.y) * (q.x - uc.x) + (uc.x - ub.x) * (q.y - uc.y)) / det;
                    float l2 = ((uc.y - ua.y) * (q.x - uc.x) + (ua.x - uc.x) * (q.y - uc.y)) / det;
                    float l3 = 1.0f - l1 - l2;
                    return pa * l1 + pb * l2 + pc * l3;
                };
                V3 n = face_normal(s, c.mesh, t);
                V3 o = at(V2{0, 0});
                V3 dx = at(V2{1, 0}) - o, dy = at(V2{0, 1}) - o;
                V2 lo{std::fmax(std::floor(std::fmin(std::fmin(ua.x, ub.x), uc.x)), 0.0f),
                      std::fmax(std::floor(std::fmin(std::fmin(ua.y, ub.y), uc.y)), 0.0f)};
                V2 hi{std::fmin(std::ceil(std::fmax(std::fmax(ua.x, ub.x), uc.x)), (float)(w - 1)),
                      std::fmin(std::ceil(std::fmax(std::fmax(ua.y, ub.y), uc.y)), (float)(h - 1))};
                RasterTri r{};
                r.ua[0] = ua.x; r.ua[1] = ua.y;
                r.ub[0] = ub.x; r.ub[1] = ub.y;
                r.uc[0] = uc.x; r.uc[1] = uc.y;
                r.det = det;
                const V3 *src[6] = {&pa, &pb, &pc, &n, &dx, &dy};
                float *dst[6] = {r.pa, r.pb, r.pc, r.n, r.dx, r.dy};
                for (int q = 0; q < 6; q++) {
                    dst[q][0] = src[q]->x;
                    dst[q][1] = src[q]->y;
                    dst[q][2] = src[q]->z;
                }
                r.chart = ci;
                r.tile = (uint32_t)k;
                r.lo[0] = (int32_t)lo.x;
                r.lo[1] = (int32_t)lo.y;
                r.hi[0] = (int32_t)hi.x;
                r.hi[1] = (int32_t)hi.y;
                out.push_back(r);
            }
        }
    }
}

}  // namespace lbk
