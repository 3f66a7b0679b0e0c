// The CPU BVH build (crates/lightbaker/src/bvh.rs), host-kept as the reference keeps it: a
// 16-bin SAH over the casting, non-degenerate triangles, leaves of at most 4 (more where no split
// beats a leaf), children adjacent. Scalar FP32 in glam's operation order (Vec3 ops are per
// component; no FMA: built with -ffp-contract=off) so the node array matches the reference's.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lightbaker.h"
#include "lb_bake.h"

namespace lbk {

namespace {

constexpr int BINS = 16;   // bvh.rs:9
constexpr size_t LEAF = 4;  // bvh.rs:10

inline V3 vmin3(V3 a, V3 b) { return vmin(a, b); }
inline V3 vmax3(V3 a, V3 b) { return vmax(a, b); }

struct Prim {
    V3 min, max, centre;
    lb_tri tri;
};

inline uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float fbits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// bvh.rs:113-121
float area(V3 mn, V3 mx) {
    V3 d{std::fmax(mx.x - mn.x, 0.0f), std::fmax(mx.y - mn.y, 0.0f), std::fmax(mx.z - mn.z, 0.0f)};
    return d.x * d.y + d.y * d.z + d.z * d.x;
}

const float INF = std::numeric_limits<float>::infinity();

// bvh.rs:119-151: the axis and plane of the cheapest binned split, if one beats a leaf
bool best_split(const Prim *p, size_t n, int *axis_out, float *at_out) {
    V3 cmin{INF, INF, INF}, cmax{-INF, -INF, -INF}, pmin = cmin, pmax = cmax;
    for (size_t i = 0; i < n; i++) {
        cmin = vmin3(cmin, p[i].centre);
        cmax = vmax3(cmax, p[i].centre);
    }
    for (size_t i = 0; i < n; i++) {
        pmin = vmin3(pmin, p[i].min);
        pmax = vmax3(pmax, p[i].max);
    }
    float leaf_cost = (float)n * area(pmin, pmax);
    bool have = false;
    float best_cost = 0.0f, best_at = 0.0f;
    int best_axis = 0;
    for (int axis = 0; axis < 3; axis++) {
        float extent = cmax[axis] - cmin[axis];
        if (extent < 1e-6f) continue;
        V3 bmin[BINS], bmax[BINS];
        uint32_t bn[BINS];
        for (int b = 0; b < BINS; b++) {
            bmin[b] = {INF, INF, INF};
            bmax[b] = {-INF, -INF, -INF};
            bn[b] = 0;
        }
        for (size_t i = 0; i < n; i++) {
            float f = (p[i].centre[axis] - cmin[axis]) / extent * (float)BINS;
            // `as usize` saturates: negative and NaN become 0
            size_t b = !(f > 0.0f) ? 0 : (f >= 4294967296.0f ? (size_t)BINS - 1 : (size_t)f);
            if (b > BINS - 1) b = BINS - 1;
            bmin[b] = vmin3(bmin[b], p[i].min);
            bmax[b] = vmax3(bmax[b], p[i].max);
            bn[b]++;
        }
        for (int s = 1; s < BINS; s++) {
            uint32_t ln = 0, rn = 0;
            for (int b = 0; b < s; b++) ln += bn[b];
            for (int b = s; b < BINS; b++) rn += bn[b];
            if (ln == 0 || rn == 0) continue;
            V3 lmin{INF, INF, INF}, lmax{-INF, -INF, -INF}, rmin = lmin, rmax = lmax;
            for (int b = 0; b < s; b++) {
                lmin = vmin3(lmin, bmin[b]);
                lmax = vmax3(lmax, bmax[b]);
            }
            for (int b = s; b < BINS; b++) {
                rmin = vmin3(rmin, bmin[b]);
                rmax = vmax3(rmax, bmax[b]);
            }
            float cost = (float)ln * area(lmin, lmax) + (float)rn * area(rmin, rmax);
            if (!have || cost < best_cost) {
                have = true;
                best_cost = cost;
                best_axis = axis;
                best_at = cmin[axis] + extent * (float)s / (float)BINS;
            }
        }
    }
    if (!have || !(best_cost < leaf_cost)) return false;
    *axis_out = best_axis;
    *at_out = best_at;
    return true;
}

}  // namespace

// bvh.rs:50-106
void build_bvh(const float *positions, size_t n_tris, const uint32_t *material,
               const uint32_t *chart, const uint8_t *casts, std::vector<lb_node> &nodes,
               std::vector<lb_tri> &tris) {
    std::vector<Prim> prims;
    prims.reserve(n_tris);
    for (size_t t = 0; t < n_tris; t++) {
        if (!casts[material[t]]) continue;
        const float *p = positions + t * 9;
        V3 a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, c{p[6], p[7], p[8]};
        V3 e1{b.x - a.x, b.y - a.y, b.z - a.z}, e2{c.x - a.x, c.y - a.y, c.z - a.z};
        V3 cr{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
        if (cr.x * cr.x + cr.y * cr.y + cr.z * cr.z < 1e-14f) continue;
        Prim pr;
        pr.min = vmin3(vmin3(a, b), c);
        pr.max = vmax3(vmax3(a, b), c);
        pr.centre = {(a.x + b.x + c.x) / 3.0f, (a.y + b.y + c.y) / 3.0f, (a.z + b.z + c.z) / 3.0f};
        pr.tri = lb_tri{{a.x, a.y, a.z, fbits(material[t])},
                        {e1.x, e1.y, e1.z, fbits(chart[t])},
                        {e2.x, e2.y, e2.z, 0.0f}};
        prims.push_back(pr);
    }
    nodes.assign(1, lb_node{});
    tris.clear();
    if (prims.empty()) {
        // bvh.rs:76-80: one empty leaf, a box nothing reaches
        float mx = std::numeric_limits<float>::max();
        nodes[0] = lb_node{{mx, mx, mx}, 0, {-mx, -mx, -mx}, 0};
        tris.assign(1, lb_tri{});
        return;
    }
    std::vector<std::tuple<size_t, size_t, size_t>> stack{{0, 0, prims.size()}};
    auto set = [&](size_t ni, V3 bmin, V3 bmax, uint32_t first, uint32_t count) {
        nodes[ni] = lb_node{{bmin.x, bmin.y, bmin.z}, first, {bmax.x, bmax.y, bmax.z}, count};
    };
    while (!stack.empty()) {
        auto [ni, start, end] = stack.back();
        stack.pop_back();
        Prim *slice = prims.data() + start;
        size_t len = end - start;
        V3 bmin{INF, INF, INF}, bmax{-INF, -INF, -INF};
        for (size_t i = 0; i < len; i++) {
            bmin = vmin3(bmin, slice[i].min);
            bmax = vmax3(bmax, slice[i].max);
        }
        int axis = 0;
        float at = 0.0f;
        bool split = len > LEAF && best_split(slice, len, &axis, &at);
        if (!split) {
            set(ni, bmin, bmax, (uint32_t)start, (uint32_t)len);
            continue;
        }
        // bvh.rs:153-162: swap the prims whose centre lies below `at` to the front
        size_t k = 0;
        for (size_t j = 0; j < len; j++) {
            if (slice[j].centre[axis] < at) {
                std::swap(slice[k], slice[j]);
                k++;
            }
        }
        size_t mid = start + k;
        if (mid == start || mid == end) mid = start + len / 2;
        if (mid - start == 0 || end - mid == 0) {
            set(ni, bmin, bmax, (uint32_t)start, (uint32_t)len);
            continue;
        }
        size_t left = nodes.size();
        nodes.push_back(lb_node{});
        nodes.push_back(lb_node{});
        set(ni, bmin, bmax, (uint32_t)left, 0);
        stack.emplace_back(left + 1, mid, end);
        stack.emplace_back(left, start, mid);
    }
    tris.reserve(prims.size());
    for (auto &p : prims) tris.push_back(p.tri);
}

}  // namespace lbk
