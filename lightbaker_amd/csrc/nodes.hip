// Checks a BVH before any kernel walks it, and derives the TUNED form's node layouts from it.
//
// Both layouts are indexed by the reference node index and hold, for an inner node, the boxes of
// its two children (first, first + 1) so one fetch serves the pair the reference reads as three
// 32 B records (trace.wgsl:90,111-112):
//   WIDE   4 x 16 B: a.min xyz a.max.x | a.max.yz b.min.xy | b.min.z b.max xyz | a ref, b ref
//   QUANT  2 x 16 B: an origin word per axis, a ref | 12 bytes of children's boxes, b ref
// QUANT boxes are 8-bit multiples of a power of two per axis above an origin at or below the
// pair's min corner: an FP32 value whose low 9 bits are zero (a box face at a round coordinate
// keeps it exactly), so the axis's word holds e + 151 in its low 8 bits beside a clear bit 8, and
// the word shifted left by 23 reads as 2^24 times the step 2^e (the kernels read the bytes as
// FP16 subnormals, q * 2^-24). Corners are rounded
// outward and checked against the device's own decode (origin + q * 2^e in FP32: the product is
// exact, so one rounding) and in exact arithmetic, so every decoded box contains its FP32 box and
// culling stays conservative. Both layouts carry each child's reference (a leaf's triangles, or
// the inner child's index), so no leaf visit reads another record.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lb_host.h"

namespace lbk {

namespace {
inline uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float fl(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
}  // namespace

bool is_empty_bvh(const lb_node *nodes, size_t n_nodes) {
    return n_nodes == 1 && nodes[0].count == 0;
}

bool validate_bvh(const lb_node *nodes, size_t n_nodes, const lb_tri *tris, size_t n_tris,
                  size_t n_materials, size_t n_charts, std::string *err) {
    char buf[256];
    size_t charts = n_charts ? n_charts : 1;
    for (size_t t = 0; t < n_tris; t++) {
        uint32_t m = bits(tris[t].v0[3]), c = bits(tris[t].e1[3]);
        if (m >= n_materials) {
            snprintf(buf, sizeof buf, "triangle %zu: material %u of %zu", t, m, n_materials);
            *err = buf;
            return false;
        }
        if (c != 0xffffffffu && c >= charts) {
            snprintf(buf, sizeof buf, "triangle %zu: chart %u of %zu", t, c, charts);
            *err = buf;
            return false;
        }
    }
    if (is_empty_bvh(nodes, n_nodes)) return true;
    // a tree: every inner node's children lie after it, every node but the root has one parent
    std::vector<uint8_t> parents(n_nodes, 0);
    for (size_t i = 0; i < n_nodes; i++) {
        const lb_node &n = nodes[i];
        if (n.count > 0) {
            if ((uint64_t)n.first + n.count > n_tris) {
                snprintf(buf, sizeof buf, "leaf %zu: triangles %u+%u of %zu", i, n.first, n.count,
                         n_tris);
                *err = buf;
                return false;
            }
        } else {
            if (n.first <= i || (uint64_t)n.first + 1 >= n_nodes) {
                snprintf(buf, sizeof buf, "node %zu: children %u,%u of %zu", i, n.first,
                         n.first + 1, n_nodes);
                *err = buf;
                return false;
            }
            if (parents[n.first]++ || parents[n.first + 1]++) {
                snprintf(buf, sizeof buf, "node %zu: a child with two parents", i);
                *err = buf;
                return false;
            }
        }
    }
    for (size_t i = 1; i < n_nodes; i++)
        if (parents[i] != 1) {
            snprintf(buf, sizeof buf, "node %zu: not reached from the root", i);
            *err = buf;
            return false;
        }
    return true;
}

// The largest FP32 value with its low 9 bits zero that is at most x (x finite).
static float origin_below(float x) {
    uint32_t u = bits(x);
    if (!(u & 0x1ffu)) return x;
    if (x >= 0.0f) return fl(u & ~0x1ffu);
    return fl((u & ~0x1ffu) + 0x200u);  // a larger magnitude: further below
}

// The outward-rounded 8-bit corners of a child on one axis, for the step 2^(ebits - 127).
// Returns false if 255 steps do not reach (the caller widens the step).
static bool quantize_axis(float origin, uint32_t ebits, float lo, float hi, uint8_t *qlo,
                          uint8_t *qhi) {
    float step = fl(ebits << 23);
    // the corner after rounding (origin + q * step, one rounding) and in exact arithmetic (the
    // traversal's fused form skips the rounding); both must lie outside [lo, hi]
    auto dec = [&](int q) { return origin + (float)q * step; };
    auto ex = [&](int q) { return (long double)origin + (long double)q * (long double)step; };
    auto below = [&](int q) { return dec(q) <= lo && ex(q) <= (long double)lo; };
    auto above = [&](int q) { return dec(q) >= hi && ex(q) >= (long double)hi; };
    int a = (int)std::floor((lo - origin) / step);
    if (a < 0) a = 0;
    while (a > 0 && !below(a)) a--;
    if (!below(a)) return false;
    int b = (int)std::ceil((hi - origin) / step);
    if (b < 0) b = 0;
    while (b <= 255 && !above(b)) b++;
    if (b > 255) return false;
    *qlo = (uint8_t)a;
    *qhi = (uint8_t)b;
    return true;
}

// A TUNED leaf reference: bit 31 set; bits 26..30 the triangle count and bits 0..25 the first
// triangle, or count field 31 and the node index when those do not fit (the kernel then reads
// the node's (first, count) from `leaves`).
static uint32_t leaf_ref(const lb_node &leaf, size_t index) {
    if (leaf.count <= 30 && leaf.first < (1u << 26)) return 0x80000000u | (leaf.count << 26) | leaf.first;
    return 0x80000000u | (31u << 26) | (uint32_t)index;
}

// Both layouts' records of every inner node (zero for leaves), on the host's threads.
static int tuned_records(const lb_node *nodes, size_t n_nodes, std::vector<float> &wide,
                         std::vector<uint32_t> &quant) {
    if (n_nodes >= (1u << 26)) return fail("scene: %zu nodes exceed 26-bit node indices", n_nodes);
    wide.assign(n_nodes * 16, 0.0f);
    quant.assign(n_nodes * 8, 0u);
    bool empty = is_empty_bvh(nodes, n_nodes);
    // every inner node's records on its own: the host's threads take ranges of nodes
    std::atomic<size_t> too_large{SIZE_MAX};
    auto one = [&](size_t i) -> bool {
        const lb_node &n = nodes[i];
        if (n.count > 0) return true;
        const lb_node &a = nodes[n.first], &b = nodes[n.first + 1];
        float *w = &wide[i * 16];
        const float vals[12] = {a.min[0], a.min[1], a.min[2], a.max[0], a.max[1], a.max[2],
                                b.min[0], b.min[1], b.min[2], b.max[0], b.max[1], b.max[2]};
        std::memcpy(w, vals, sizeof vals);
        // each child's reference, a leaf's carrying its (first, count): a leaf visit then reads
        // no node record
        uint32_t ra = a.count > 0 ? leaf_ref(a, n.first) : n.first;
        uint32_t rb = b.count > 0 ? leaf_ref(b, n.first + 1) : n.first + 1;
        w[12] = fl(ra);
        w[13] = fl(rb);
        // quantized: origin at or below the union's min corner, one exponent an axis
        uint32_t *q = &quant[i * 8];
        uint8_t qb[12];
        uint32_t ex[3];
        for (int ax = 0; ax < 3; ax++) {
            float lo = std::fmin(a.min[ax], b.min[ax]);
            float hi = std::fmax(a.max[ax], b.max[ax]);
            float ext = hi - lo;
            int e = ext > 0.0f ? (int)std::ceil(std::log2((double)ext / 250.0)) : -100;
            if (e < -100) e = -100;
            float origin = origin_below(lo);
            if (!std::isfinite(origin)) return false;
            for (;; e++) {
                uint32_t eb = (uint32_t)(e + 127);
                if (quantize_axis(origin, eb, a.min[ax], a.max[ax], &qb[ax], &qb[3 + ax]) &&
                    quantize_axis(origin, eb, b.min[ax], b.max[ax], &qb[6 + ax], &qb[9 + ax])) {
                    ex[ax] = eb;
                    break;
                }
                if (e >= 103) return false;  // e + 151 must fit the byte
            }
            q[ax] = bits(origin) | (ex[ax] + 24u);
        }
        q[3] = ra;
        std::memcpy(&q[4], qb, 12);
        q[7] = rb;
        return true;
    };
    if (!empty) {
        size_t threads = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
        if (n_nodes < 65536) threads = 1;
        auto range = [&](size_t k) {
            for (size_t i = n_nodes * k / threads; i < n_nodes * (k + 1) / threads; i++)
                if (!one(i)) {
                    size_t cur = too_large.load();
                    while (i < cur && !too_large.compare_exchange_weak(cur, i)) {
                    }
                }
        };
        std::vector<std::thread> pool;
        for (size_t k = 1; k < threads; k++) pool.emplace_back(range, k);
        range(0);
        for (auto &t : pool) t.join();
        if (too_large.load() != SIZE_MAX)
            return fail("scene: node %zu is too large to quantize", too_large.load());
    }
    return 0;
}

int build_tuned_nodes(lb_scene *s, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                      size_t n_tris, const float *materials) {
    std::vector<float> wide;
    std::vector<uint32_t> quant;
    if (tuned_records(nodes, n_nodes, wide, quant)) return 1;
    std::vector<uint32_t> leafinfo(n_nodes * 2, 0u);
    bool empty = is_empty_bvh(nodes, n_nodes);
    // leaves: (first triangle, count) at the leaf's own index
    for (size_t i = 0; i < n_nodes; i++) {
        leafinfo[i * 2] = nodes[i].first;
        leafinfo[i * 2 + 1] = nodes[i].count;
    }
    if (empty) {
        // the reference's empty marker (bvh.rs:76-80): one leaf over the zero triangle, which no
        // ray meets (its determinant is 0)
        leafinfo[0] = 0;
        leafinfo[1] = 1;
    }
    s->root_ref = (empty || nodes[0].count > 0) ? (0x80000000u | (31u << 26)) : 0u;
    // the tuned form's triangles carry their material's opacity in e2.w (0 in the reference's
    // record), so stops() reads no material
    std::vector<lb_tri> tt(tris, tris + n_tris);
    for (size_t t = 0; t < n_tris; t++) {
        uint32_t m = bits(tt[t].v0[3]);
        tt[t].e2[3] = materials[(size_t)m * 4 + 3];
    }
    hipStream_t st = s->gpu->stream;
    if (s->tris_t.upload(tt.data(), tt.size() * sizeof(lb_tri), st)) return 1;
    if (s->wide.upload(wide.data(), wide.size() * 4, st)) return 1;
    if (s->quant.upload(quant.data(), quant.size() * 4, st)) return 1;
    if (s->leaves.upload(leafinfo.data(), leafinfo.size() * 4, st)) return 1;
    return 0;
}

}  // namespace lbk

extern "C" int lb_tuned_nodes(const lb_node *nodes, size_t n_nodes, uint32_t *quant,
                              float *wide) {
    if (!nodes || n_nodes == 0) return lbk::fail("lb_tuned_nodes: a BVH has at least one node");
    std::vector<float> w;
    std::vector<uint32_t> q;
    if (lbk::tuned_records(nodes, n_nodes, w, q)) return 1;
    if (quant) std::memcpy(quant, q.data(), q.size() * 4);
    if (wide) std::memcpy(wide, w.data(), w.size() * 4);
    return 0;
}
