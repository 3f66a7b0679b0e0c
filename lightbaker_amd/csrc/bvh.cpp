// The CPU BVH build (crates/lightbaker/src/bvh.rs), host-kept as the reference keeps it: a
// 16-bin SAH over the casting, non-degenerate triangles, leaves of at most 4 (more where no split
// beats a leaf), children adjacent. Scalar FP32 in glam's operation order (Vec3 ops are per
// component; no FMA: built with -ffp-contract=off) so the node array matches the reference's.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
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

// A slice's reductions, on up to `threads` threads for large slices: minima, maxima and counts
// are exact, so any grouping gives the single loop's values.
struct Bins {
    V3 cmin{INF, INF, INF}, cmax{-INF, -INF, -INF}, pmin{INF, INF, INF}, pmax{-INF, -INF, -INF};
    V3 bmin[3][BINS], bmax[3][BINS];
    uint32_t bn[3][BINS];
};

template <typename F>
void parallel_chunks(size_t n, size_t threads, F f) {
    size_t parts = n >= 65536 && threads > 1 ? threads : 1;
    std::vector<std::thread> pool;
    for (size_t k = 1; k < parts; k++) pool.emplace_back(f, k, n * k / parts, n * (k + 1) / parts);
    f(0, 0, n / parts);
    for (auto &t : pool) t.join();
}

// bvh.rs:119-151: the axis and plane of the cheapest binned split, if one beats a leaf
bool best_split(const Prim *p, size_t n, int *axis_out, float *at_out, size_t threads) {
    size_t parts = n >= 65536 && threads > 1 ? threads : 1;
    std::vector<Bins> part(parts);
    parallel_chunks(n, threads, [&](size_t k, size_t a, size_t b) {
        Bins &q = part[k];
        for (size_t i = a; i < b; i++) {
            q.cmin = vmin3(q.cmin, p[i].centre);
            q.cmax = vmax3(q.cmax, p[i].centre);
            q.pmin = vmin3(q.pmin, p[i].min);
            q.pmax = vmax3(q.pmax, p[i].max);
        }
    });
    V3 cmin{INF, INF, INF}, cmax{-INF, -INF, -INF}, pmin = cmin, pmax = cmax;
    for (const Bins &q : part) {
        cmin = vmin3(cmin, q.cmin);
        cmax = vmax3(cmax, q.cmax);
        pmin = vmin3(pmin, q.pmin);
        pmax = vmax3(pmax, q.pmax);
    }
    float extent3[3] = {cmax.x - cmin.x, cmax.y - cmin.y, cmax.z - cmin.z};
    parallel_chunks(n, threads, [&](size_t k, size_t a, size_t b) {
        Bins &q = part[k];
        for (int axis = 0; axis < 3; axis++)
            for (int j = 0; j < BINS; j++) {
                q.bmin[axis][j] = {INF, INF, INF};
                q.bmax[axis][j] = {-INF, -INF, -INF};
                q.bn[axis][j] = 0;
            }
        for (int axis = 0; axis < 3; axis++) {
            float extent = extent3[axis];
            if (extent < 1e-6f) continue;
            for (size_t i = a; i < b; i++) {
                float f = (p[i].centre[axis] - cmin[axis]) / extent * (float)BINS;
                // `as usize` saturates: negative and NaN become 0
                size_t j = !(f > 0.0f) ? 0 : (f >= 4294967296.0f ? (size_t)BINS - 1 : (size_t)f);
                if (j > BINS - 1) j = BINS - 1;
                q.bmin[axis][j] = vmin3(q.bmin[axis][j], p[i].min);
                q.bmax[axis][j] = vmax3(q.bmax[axis][j], p[i].max);
                q.bn[axis][j]++;
            }
        }
    });
    float leaf_cost = (float)n * area(pmin, pmax);
    bool have = false;
    float best_cost = 0.0f, best_at = 0.0f;
    int best_axis = 0;
    for (int axis = 0; axis < 3; axis++) {
        float extent = extent3[axis];
        if (extent < 1e-6f) continue;
        V3 bmin[BINS], bmax[BINS];
        uint32_t bn[BINS];
        for (int b = 0; b < BINS; b++) {
            bmin[b] = {INF, INF, INF};
            bmax[b] = {-INF, -INF, -INF};
            bn[b] = 0;
            for (const Bins &q : part) {
                bmin[b] = vmin3(bmin[b], q.bmin[axis][b]);
                bmax[b] = vmax3(bmax[b], q.bmax[axis][b]);
                bn[b] += q.bn[axis][b];
            }
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
               std::vector<lb_tri> &tris, int max_threads) {
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
    // The reference's build is one depth-first loop (bvh.rs:82-105): pop a node, split it or make
    // it a leaf, and on a split append its two children to the node array and push them, the
    // left one on top. Node numbers follow the order of those appends. Here the loop runs the
    // same way down to subtrees of at most `deferred` primitives; those are built on their own
    // threads with the same loop over their own slices (disjoint, so the partitions do not meet)
    // and numbered locally, and the numbers are put together afterwards in the order the single
    // loop would have appended them: a subtree's appends all come at the point it is popped.
    struct Task {
        size_t node, start, end;        // provisional node id, slice
        std::vector<lb_node> local;     // its nodes, local 0 its root, children as local ids
    };
    struct Event {
        bool subtree;
        size_t id;  // split: the provisional id of the left child (the right is id + 1); subtree: task
    };
    std::vector<lb_node> prov(1);  // the loop's own nodes, by provisional id
    std::vector<Task> tasks;
    std::vector<Event> events;
    size_t threads = max_threads > 0 ? (size_t)max_threads : std::thread::hardware_concurrency();
    if (threads < 1) threads = 1;
    if (threads > 16) threads = 16;
    size_t deferred = threads > 1 ? std::max<size_t>(4096, prims.size() / (threads * 8)) : 0;

    // one node of the loop (bvh.rs:84-104) over slice [start, end): returns true on a split,
    // with the slice partitioned at *mid
    auto visit = [&](Prim *all, size_t start, size_t end, lb_node &out, size_t *mid,
                     size_t nthreads) {
        Prim *slice = all + start;
        size_t len = end - start;
        V3 bmin{INF, INF, INF}, bmax{-INF, -INF, -INF};
        for (size_t i = 0; i < len; i++) {
            bmin = vmin3(bmin, slice[i].min);
            bmax = vmax3(bmax, slice[i].max);
        }
        out = lb_node{{bmin.x, bmin.y, bmin.z}, (uint32_t)start, {bmax.x, bmax.y, bmax.z},
                      (uint32_t)len};
        int axis = 0;
        float at = 0.0f;
        if (!(len > LEAF && best_split(slice, len, &axis, &at, nthreads))) return false;
        // bvh.rs:153-162: swap the prims whose centre lies below `at` to the front
        size_t k = 0;
        for (size_t j = 0; j < len; j++) {
            if (slice[j].centre[axis] < at) {
                std::swap(slice[k], slice[j]);
                k++;
            }
        }
        size_t m = start + k;
        if (m == start || m == end) m = start + len / 2;
        if (m - start == 0 || end - m == 0) return false;
        *mid = m;
        out.count = 0;
        return true;
    };
    // the loop over one subtree, numbering its nodes from 0
    auto build_local = [&](Task &t) {
        t.local.assign(1, lb_node{});
        std::vector<std::tuple<size_t, size_t, size_t>> st{{0, t.start, t.end}};
        while (!st.empty()) {
            auto [ni, start, end] = st.back();
            st.pop_back();
            size_t mid = 0;
            lb_node nd;
            if (!visit(prims.data(), start, end, nd, &mid, 1)) {
                t.local[ni] = nd;
                continue;
            }
            size_t left = t.local.size();
            t.local.push_back(lb_node{});
            t.local.push_back(lb_node{});
            nd.first = (uint32_t)left;
            t.local[ni] = nd;
            st.emplace_back(left + 1, mid, end);
            st.emplace_back(left, start, mid);
        }
    };

    std::vector<std::tuple<size_t, size_t, size_t>> stack{{0, 0, prims.size()}};
    while (!stack.empty()) {
        auto [pn, start, end] = stack.back();
        stack.pop_back();
        if (end - start <= deferred && end - start > LEAF) {
            events.push_back(Event{true, tasks.size()});
            tasks.push_back(Task{pn, start, end, {}});
            continue;
        }
        size_t mid = 0;
        lb_node nd;
        if (!visit(prims.data(), start, end, nd, &mid, threads)) {
            prov[pn] = nd;
            continue;
        }
        size_t left = prov.size();
        prov.push_back(lb_node{});
        prov.push_back(lb_node{});
        nd.first = (uint32_t)left;  // provisional; renumbered below
        prov[pn] = nd;
        events.push_back(Event{false, left});
        stack.emplace_back(left + 1, mid, end);
        stack.emplace_back(left, start, mid);
    }
    // the subtrees, largest first, on the pool
    if (!tasks.empty()) {
        std::vector<size_t> order(tasks.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = i;
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return tasks[a].end - tasks[a].start > tasks[b].end - tasks[b].start;
        });
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next.fetch_add(1)) < order.size();) build_local(tasks[order[i]]);
        };
        std::vector<std::thread> pool;
        for (size_t i = 1; i < std::min(threads, tasks.size()); i++) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
    }
    // the single loop's numbers: provisional ids and each subtree's local ids, in append order
    std::vector<size_t> pmap(prov.size(), 0);
    std::vector<size_t> tbase(tasks.size(), 0);  // a subtree's local id k > 0 is tbase + k - 1
    size_t g = 1;
    for (const Event &e : events) {
        if (!e.subtree) {
            pmap[e.id] = g;
            pmap[e.id + 1] = g + 1;
            g += 2;
        } else {
            tbase[e.id] = g;
            g += tasks[e.id].local.size() - 1;
        }
    }
    nodes.assign(g, lb_node{});
    for (size_t pn = 0; pn < prov.size(); pn++) {
        lb_node nd = prov[pn];
        if (nd.count == 0) nd.first = (uint32_t)pmap[nd.first];
        nodes[pmap[pn]] = nd;
    }
    for (size_t ti = 0; ti < tasks.size(); ti++) {
        const Task &t = tasks[ti];
        auto gid = [&](size_t k) { return k == 0 ? pmap[t.node] : tbase[ti] + k - 1; };
        for (size_t k = 0; k < t.local.size(); k++) {
            lb_node nd = t.local[k];
            if (nd.count == 0) nd.first = (uint32_t)gid(nd.first);
            nodes[gid(k)] = nd;
        }
    }
    tris.reserve(prims.size());
    for (auto &p : prims) tris.push_back(p.tri);
}

}  // namespace lbk
