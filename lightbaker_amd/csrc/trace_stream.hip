// The POOLED schedule of the TUNED form (schedule 1): a pass's rays form one pool that every lane
// of a persistent grid draws from, and the hits are folded per item by a separate kernel in the
// reference's ray order.
//
//   trace  a wave claims ray ids in batches of 64 from one global counter and hands them to its
//          idle lanes once `refill` lanes are idle; a lane generates its ray from the id (item =
//          id / slots, slot = id % slots, slots in azimuth order so a batch is one wedge of one
//          item's directions) and traverses it in the reference's order (lb_traverse.h). Each
//          iteration the wave takes one step: the leaves of the lanes standing at one once
//          `leaf_batch` of them do (or no lane is at an inner node), else one inner node for the
//          lanes at one. A finished ray writes its hit (8 B) to the item's slot in HBM.
//   shade  one thread an item regenerates its rays and folds their hits in the reference's order
//          with the reference's arithmetic: every output equals the PLAIN form's bit for bit.
//
// Hits: (bits of t, negated for a back face; triangle, NONE for a miss), stored slot-major (hit
// of ray r of item i at r * items + i) so the shading reads are coalesced over items.
//
// The direct pass runs its inside test (trace.wgsl:159-168, 16 rays a point) as rounds: round 0
// tests every item's texel centre, round q (1..9) tests point q of the items still inside after
// round q-1 (exactly the points the reference tests), then one main round traces the sun and
// occlusion rays from each item's chosen point.
#include "lb_host.h"
#include "lb_traverse.h"

namespace lbk {

constexpr size_t MAX_POOL_RAYS = size_t(1) << 28;  // hits a chunk holds (2 GiB)
constexpr uint32_t BATCH = 64;                      // ray ids a wave claims at once

enum RaySource { SRC_GATHER = 0, SRC_VOLUME = 1, SRC_DIRECT = 2, SRC_INSIDE = 3 };

struct PoolArgs {
    SceneView S;
    lb_params P;
    const lb_texel *items;   // the chunk's items
    uint32_t n;              // items of the chunk
    const uint32_t *n_dev;   // SRC_INSIDE rounds 1..9: the listed items' count (else NULL)
    const uint32_t *list;    // SRC_INSIDE rounds 1..9: the listed items (else NULL: every item)
    const uint32_t *pick;    // SRC_DIRECT: the point each item settled on
    uint32_t key0;           // the key index of item 0
    uint32_t q;              // SRC_INSIDE: the round
    uint32_t slots;          // ray slots an item
    uint2 *hits;
    uint32_t *queue;
    uint32_t *spill;
    uint32_t depth, refill, leaf_batch;
    unsigned long long *stats;
};

struct Item {
    f3 pos, n, dx, dy;
    uint32_t w;  // normal.w bits
};
__device__ __forceinline__ Item item_at(const lb_texel *items, uint32_t i) {
    const float4 *tp = reinterpret_cast<const float4 *>(items + i);
    float4 a = tp[0], b = tp[1], c = tp[2], d = tp[3];
    return Item{xyz(a), xyz(b), xyz(c), xyz(d), fbits(b.w)};
}

// Slot s of `total` in azimuth order: for a power of two, the ray whose radical inverse is
// s / total (s bit-reversed), so consecutive slots sweep the azimuth parameter; else s itself.
__device__ __forceinline__ uint32_t az_order(uint32_t s, uint32_t total) {
    if (total < 2u || (total & (total - 1u)) != 0u) return s;
    uint32_t b = 31u - (uint32_t)__builtin_clz(total);
    return __builtin_bitreverse32(s) >> (32u - b);
}

// ---- the rays of each pass, as the reference forms them -----------------------------------------

// trace.wgsl:250-257: gather ray r of an item
__device__ __forceinline__ void gather_ray(const lb_params &P, const Item &tx, uint32_t key,
                                           uint32_t r, f3 *o, f3 *d) {
    uint32_t seed = fbits(P.reach[3]);
    uint32_t rays = P.counts[2];
    frame_t f = frame(tx.n);
    float s1 = rnd(key, 11u, seed), s2 = rnd(key, 13u, seed);
    float jx = rnd(key + r, 19u, seed) - 0.5f;
    float jy = rnd(key + r, 23u, seed) - 0.5f;
    *o = add(add(add(tx.pos, scl(tx.dx, jx)), scl(tx.dy, jy)), scl(tx.n, P.reach[2]));
    *d = cosine(fract(((float)r + 0.5f) / (float)rays + s1), fract(radical(r) + s2), f);
}

// trace.wgsl:278-285: volume ray r of a probe
__device__ __forceinline__ f3 volume_ray(const lb_params &P, uint32_t key, uint32_t r) {
    uint32_t seed = fbits(P.reach[3]);
    uint32_t rays = P.counts[2];
    float s1 = rnd(key, 29u, seed), s2 = rnd(key, 31u, seed);
    float u = fract(((float)r + 0.5f) / (float)rays + s1);
    float v = fract(radical(r) + s2);
    float z = 1.0f - 2.0f * u;
    float rr = sqrtf(vmax(0.0f, 1.0f - z * z));
    float a = v * 2.0f * PI;
    float sa, ca;
    sincos_lb(a, &sa, &ca);
    return mk(rr * ca, z, rr * sa);
}

// the sun-disc direction of sample k of gg (trace.wgsl:213-217, 303-307)
__device__ __forceinline__ f3 sun_ray(const lb_params &P, uint32_t k, uint32_t gg) {
    f3 sun = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    frame_t sf = frame(sun);
    float a = (float)k * 2.399963f;
    float r = P.sun_dir[3] * sqrtf(((float)k + 0.5f) / (float)gg);
    float sa, ca;
    sincos_lb(a, &sa, &ca);
    return normalize3(add(add(sun, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
}

// Point q of the inside test (trace.wgsl:174-193): 0 the texel centre, 1..8 the offsets, 9 the
// centre pushed along the normal; and the key its rays derive from.
__device__ __forceinline__ f3 inside_point(const Item &tx, uint32_t q, uint32_t key,
                                           uint32_t *kq) {
    if (q == 0u) {
        *kq = key;
        return tx.pos;
    }
    if (q < 9u) {
        float cx, cy;
        offset(q, &cx, &cy);
        *kq = key + q * 31u;
        return add(add(tx.pos, scl(tx.dx, cx)), scl(tx.dy, cy));
    }
    *kq = key + 331u;
    return add(tx.pos, scl(tx.n, PUSH));
}

struct RaySpec {
    f3 o, d;
    float tmax;    // 0: an empty slot (nothing to trace)
    bool any;
    uint32_t key;  // stops()'s ray key
    uint32_t hit;  // where the hit goes
};

// Ray `id` of a launch over source SRC (id = item * slots + slot).
template <int SRC>
__device__ __forceinline__ RaySpec make_ray(const PoolArgs &A, uint32_t id) {
    const lb_params &P = A.P;
    uint32_t k = id / A.slots, m = id - k * A.slots;
    RaySpec R;
    R.any = false;
    if (SRC == SRC_GATHER) {
        Item tx = item_at(A.items, k);
        uint32_t key = (A.key0 + k) * 7919u;
        uint32_t r = az_order(m, A.slots);
        gather_ray(P, tx, key, r, &R.o, &R.d);
        R.tmax = P.reach[1];
        R.key = key + r + 4000u;
        R.hit = r * A.n + k;
    } else if (SRC == SRC_VOLUME) {
        Item pr = item_at(A.items, k);
        uint32_t key = pr.w * 7919u + 77u;
        uint32_t nr = P.counts[2];
        R.o = pr.pos;
        if (m < nr) {
            uint32_t r = az_order(m, nr);
            R.d = volume_ray(P, key, r);
            R.tmax = P.reach[1];
            R.key = key + r;
            R.hit = r * A.n + k;
        } else {
            uint32_t s = m - nr, gg = P.counts[0] * P.counts[0];
            R.d = sun_ray(P, s, gg);
            R.tmax = FAR;
            R.any = true;
            R.key = key + s + 9000u;
            R.hit = m * A.n + k;
        }
    } else if (SRC == SRC_DIRECT) {
        Item tx = item_at(A.items, k);
        uint32_t key = (A.key0 + k) * 7919u;
        uint32_t g = P.counts[0], gg = g * g, nr = P.counts[1];
        uint32_t q = A.pick[k], kq;
        f3 at = inside_point(tx, q, key, &kq);
        f3 lift = scl(tx.n, P.reach[2]);
        if (m < 2u * gg) {
            R.hit = m * A.n + k;
            f3 sun = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
            if (!(dot3(sun, tx.n) > 0.0f)) {
                R.tmax = 0.0f;
                return R;
            }
            float span = q == 0u ? 1.0f : 0.4f;
            uint32_t s = m < gg ? m : m - gg;
            float cx = (((float)(s % g) + 0.5f) / (float)g - 0.5f) * span;
            float cy = (((float)(s / g) + 0.5f) / (float)g - 0.5f) * span;
            R.o = add(add(add(at, scl(tx.dx, cx)), scl(tx.dy, cy)), lift);
            R.tmax = FAR;
            R.any = true;
            if (m < gg) {
                R.d = sun;
                R.key = key + s;
            } else {
                float kr = rnd(key, 17u, fbits(P.reach[3]));
                R.d = sun_ray(P, (s + (uint32_t)(kr * (float)gg)) % gg, gg);
                R.key = key + s + 1000u;
            }
        } else {
            uint32_t seed = fbits(P.reach[3]);
            uint32_t r = az_order(m - 2u * gg, nr);
            float s1 = rnd(key, 3u, seed), s2 = rnd(key, 5u, seed);
            frame_t f = frame(tx.n);
            R.o = add(at, lift);
            R.d = cosine(fract(((float)r + 0.5f) / (float)nr + s1), fract(radical(r) + s2), f);
            R.tmax = P.reach[0];
            R.key = key + r + 2000u;
            R.hit = (2u * gg + r) * A.n + k;
        }
    } else {
        uint32_t li = A.list ? A.list[k] : k;
        Item tx = item_at(A.items, li);
        uint32_t key = (A.key0 + li) * 7919u, kq;
        f3 p = inside_point(tx, A.q, key, &kq);
        R.o = add(p, scl(tx.n, P.reach[2]));
        R.d = cosine(((float)m + 0.5f) / 16.0f, radical(m), frame(tx.n));
        R.tmax = P.ground[3];
        R.key = kq + m + 7000u;
        R.hit = k * 16u + m;
    }
    return R;
}

// ---- the trace kernel -------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <int SRC, int LAYOUT, bool STATS>
__global__ void k_pool(PoolArgs A) {
    StackView st = make_stack(A.spill, A.depth);
    uint32_t lane = threadIdx.x & 63u;
    uint32_t total = (A.n_dev ? *A.n_dev : A.n) * A.slots;
    Counts cnt{0, 0, 0, 0, 0, 0};
    Ray R;
    uint32_t hit = 0;
    bool active = false;
    // wave-uniform: the wave's claimed ids [next, end); no ray is left for it once drained
    uint32_t next = 0, end = 0;
    bool drained = false;
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (!drained && (__popcll(idle) >= (int)A.refill || idle == ~0ull)) {
            uint32_t need = (uint32_t)__popcll(idle);
            uint32_t have = end - next;
            uint32_t fresh = end;
            if (have < need) {
                // the ids beyond this batch come from a new one; claims only grow, so once a
                // wave's ids reach `total` every later claim is past it too
                uint32_t b = 0;
                if (lane == 0u) b = atomicAdd(A.queue, BATCH);
                fresh = __builtin_amdgcn_readfirstlane(b);
            }
            uint32_t rank = lanes_below(idle);
            uint32_t my = rank < have ? next + rank : fresh + (rank - have);
            if (have < need) {
                next = fresh + (need - have);
                end = fresh + BATCH;
            } else {
                next += need;
            }
            drained = next >= total;
            if (!active && my < total) {
                RaySpec s = make_ray<SRC>(A, my);
                if (s.tmax == 0.0f) {
                    A.hits[s.hit] = make_uint2(0u, NONE);
                } else {
                    ray_start(R, A.S.root, s.o, s.d, s.tmax, s.any, s.key);
                    hit = s.hit;
                    active = true;
                    if (STATS) cnt.rays++;
                }
            }
        }
        uint64_t busy = __ballot(active);
        if (busy == 0ull) {
            if (drained) break;
            continue;
        }
        bool at_leaf = active && (R.n & LEAF_BIT) != 0u;
        uint64_t leaves = __ballot(at_leaf);
        bool leaf_step = leaves == busy || __popcll(leaves) >= (int)A.leaf_batch;
        bool done = false;
        if (leaf_step) {
            if (at_leaf) done = leaf_tests<STATS>(A.S, A.P, R, cnt) || pop(R, st);
        } else if (active && !at_leaf) {
            done = !inner_visit<LAYOUT, STATS>(A.S, R, st, cnt) && pop(R, st);
        }
        if (done) {
            A.hits[hit] = make_uint2(fbits(R.back ? -R.t : R.t), R.tri);
            active = false;
        }
    }
    if (STATS) flush_counts(cnt, A.stats);
}

__device__ __forceinline__ Hit hit_of(uint2 h) {
    return Hit{fabsf(bitsf(h.x)), h.y, (h.x >> 31) != 0u};
}

// ---- shading in the reference's order ---------------------------------------------------------

// trace.wgsl:245-270
template <bool STATS>
__global__ void k_shade_gather(SceneView S, lb_params P, const lb_texel *items, uint32_t nc,
                               uint32_t key0, const uint2 *hits, float4 *out,
                               unsigned long long *stats) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    Counts cnt{0, 0, 0, 0, 0, 0};
    uint32_t total = P.counts[2];
    Item tx = item_at(items, li);
    uint32_t key = (key0 + li) * 7919u;
    f3 sum = mk(0.0f, 0.0f, 0.0f);
    float back = 0.0f;
    for (uint32_t r = 0u; r < total; r++) {
        Hit h = hit_of(hits[(size_t)r * nc + li]);
        if (h.back && h.tri != NONE) {
            back += 1.0f;
            continue;
        }
        f3 o, d;
        gather_ray(P, tx, key, r, &o, &d);
        if (h.tri == NONE) {
            sum = add(sum, sky(S.sh, P, d));
            continue;
        }
        if (STATS) {
            cnt.mats++;
            cnt.gathers++;
        }
        sum = add(sum, bounced(S, P, o, d, h));
    }
    out[li] = make_float4(sum.x * PI / (float)total, sum.y * PI / (float)total,
                          sum.z * PI / (float)total, back / (float)total);
    if (STATS) flush_counts(cnt, stats);
}

// trace.wgsl:272-324
template <bool STATS>
__global__ void k_shade_volume(SceneView S, lb_params P, const lb_texel *items, uint32_t nc,
                               const uint2 *hits, float4 *out, unsigned long long *stats) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    Counts cnt{0, 0, 0, 0, 0, 0};
    uint32_t nr = P.counts[2];
    uint32_t g = P.counts[0], gg = g * g;
    Item pr = item_at(items, li);
    uint32_t key = pr.w * 7919u + 77u;
    f3 l0 = mk(0, 0, 0), l1x = mk(0, 0, 0), l1y = mk(0, 0, 0), l1z = mk(0, 0, 0);
    float back = 0.0f;
    for (uint32_t r = 0u; r < nr; r++) {
        Hit h = hit_of(hits[(size_t)r * nc + li]);
        f3 d = volume_ray(P, key, r);
        f3 l = sky(S.sh, P, d);
        if (h.tri != NONE) {
            if (h.back) back += 1.0f;
            if (STATS && !h.back) {
                cnt.mats++;
                cnt.gathers++;
            }
            l = h.back ? mk(0.0f, 0.0f, 0.0f) : bounced(S, P, pr.pos, d, h);
        }
        l0 = add(l0, scl(l, 0.282095f));
        l1x = add(l1x, scl(scl(l, 0.488603f), d.x));
        l1y = add(l1y, scl(scl(l, 0.488603f), d.y));
        l1z = add(l1z, scl(scl(l, 0.488603f), d.z));
    }
    float w = 4.0f * PI / (float)nr;
    float sun = 0.0f;
    for (uint32_t s = 0u; s < gg; s++)
        if (hits[(size_t)(nr + s) * nc + li].y == NONE) sun += 1.0f;
    out[li * 4u] = make_float4(l0.x * w, l0.y * w, l0.z * w, sun / (float)gg);
    out[li * 4u + 1u] = make_float4(l1x.x * w, l1x.y * w, l1x.z * w, back / (float)nr);
    out[li * 4u + 2u] = make_float4(l1y.x * w, l1y.y * w, l1y.z * w, 0.0f);
    out[li * 4u + 3u] = make_float4(l1z.x * w, l1z.y * w, l1z.z * w, 0.0f);
    if (STATS) flush_counts(cnt, stats);
}

// trace.wgsl:166-167 on round q's hits: an item outside at point q settles on it (round 0: on
// the centre); one still inside goes on to point q + 1, and after point 9 settles on the centre.
__global__ void k_settle_inside(lb_params P, uint32_t nc, uint32_t q, const uint32_t *list,
                                const uint32_t *count, const uint2 *hits, uint32_t *next,
                                uint32_t *next_count, uint32_t *pick) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = count ? *count : nc;
    if (k >= m) return;
    uint32_t li = list ? list[k] : k;
    float back = 0.0f;
    for (uint32_t r = 0u; r < 16u; r++) {
        uint2 h = hits[(size_t)k * 16u + r];
        if (h.y != NONE && (h.x >> 31) != 0u) back += 1.0f;
    }
    bool inside = back / 16.0f > bitsf(P.atlas[3]);
    if (!inside) {
        pick[li] = q;
    } else if (q < 9u) {
        next[atomicAdd(next_count, 1u)] = li;
    } else {
        pick[li] = 0u;
    }
}

// trace.wgsl:194-243 on the main round's hits
__global__ void k_shade_direct(lb_params P, const lb_texel *items, uint32_t nc,
                               const uint32_t *pick, const uint2 *hits, float4 *out) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    uint32_t g = P.counts[0], gg = g * g;
    uint32_t nr = P.counts[1];
    f3 n = xyz(reinterpret_cast<const float4 *>(items + li)[1]);
    f3 sun = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    uint32_t hard = 0u;
    float soft = 0.0f;
    if (dot3(sun, n) > 0.0f) {
        for (uint32_t s = 0u; s < gg; s++) {
            if (hits[(size_t)s * nc + li].y == NONE) hard |= 1u << s;
            if (hits[(size_t)(gg + s) * nc + li].y == NONE) soft += 1.0f;
        }
        soft /= (float)gg;
    }
    float occ = 0.0f, back = 0.0f;
    for (uint32_t r = 0u; r < nr; r++) {
        Hit h = hit_of(hits[(size_t)(2u * gg + r) * nc + li]);
        if (h.tri == NONE) continue;
        float k = h.t / P.reach[0];
        occ += 1.0f - k * k;
        if (h.back && h.t < P.ground[3]) back += 1.0f;
    }
    out[li] = make_float4(soft, bitsf(hard | (pick[li] << 24u)), occ / (float)nr,
                          back / (float)nr);
}

__global__ void k_zero_words(uint32_t *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- host side ------------------------------------------------------------------------------

using PoolFn = void (*)(PoolArgs);

template <int SRC>
static PoolFn pick_pool_src(int layout, bool stats) {
    if (layout == LB_NODES_WIDE) return stats ? k_pool<SRC, 0, true> : k_pool<SRC, 0, false>;
    return stats ? k_pool<SRC, 1, true> : k_pool<SRC, 1, false>;
}

static PoolFn pick_pool(int src, int layout, bool stats) {
    switch (src) {
    case SRC_GATHER: return pick_pool_src<SRC_GATHER>(layout, stats);
    case SRC_VOLUME: return pick_pool_src<SRC_VOLUME>(layout, stats);
    case SRC_DIRECT: return pick_pool_src<SRC_DIRECT>(layout, stats);
    default: return pick_pool_src<SRC_INSIDE>(layout, stats);
    }
}

// Ray slots an item takes in its largest round.
static uint32_t slots_per_item(int pass, const lb_params &P) {
    uint32_t gg = P.counts[0] * P.counts[0];
    if (pass == LB_PASS_GATHER) return P.counts[2];
    if (pass == LB_PASS_VOLUME) return P.counts[2] + gg;
    uint32_t main = 2u * gg + P.counts[1];
    return main > 16u ? main : 16u;
}

struct PoolPlan {
    uint32_t block, grid, chunk;  // grid: the largest of the sources' grids
    size_t lds;
};

// A persistent grid: as many workgroups as the device keeps resident at once.
static hipError_t pool_grid(lb_gpu *gpu, PoolFn fn, uint32_t block, size_t lds, uint32_t *grid) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(fn), (int)block, lds);
    if (e != hipSuccess) return e;
    *grid = (uint32_t)gpu->cus * (uint32_t)(per_cu < 1 ? 1 : per_cu);
    return hipSuccess;
}

// Grid and chunk of a launch over n items; reserves the hits, item lists and spill area.
static hipError_t pool_plan(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n,
                            PoolPlan *out) {
    const Tuning &T = gpu->tuning[pass];
    out->block = (uint32_t)T.workgroup;
    out->lds = stack_lds_bytes(T.lds_stack, (int)out->block);
    // the spill area covers the largest grid any source's kernel launches with
    out->grid = 1;
    for (int src = 0; src < 4; src++)
        for (int stats = 0; stats < 2; stats++) {
            uint32_t g = 0;
            hipError_t e = pool_grid(gpu, pick_pool(src, T.nodes, stats != 0), out->block,
                                     out->lds, &g);
            if (e != hipSuccess) return e;
            if (g > out->grid) out->grid = g;
        }
    uint32_t per = slots_per_item(pass, P);
    size_t cap = T.chunk > 0 ? (size_t)T.chunk : MAX_POOL_RAYS;
    size_t chunk = cap / per;
    if (chunk < 1) chunk = 1;
    out->chunk = (uint32_t)(chunk < n ? chunk : (n ? n : 1u));
    size_t rays = (size_t)out->chunk * per;
    size_t spill = (size_t)out->grid * out->block * (size_t)(STACK - T.lds_stack) * 4u;
    // aux words: picks and two item lists (direct pass), then one count a round
    size_t aux = ((size_t)out->chunk * 3u + 64u) * 4u;
    if (gpu->spill.reserve(spill < 16 ? 16 : spill) || gpu->stream_hits.reserve(rays * 8u) ||
        gpu->stream_aux.reserve(aux))
        return hipErrorOutOfMemory;
    return hipSuccess;
}

hipError_t stream_prepare(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n) {
    PoolPlan pp;
    return pool_plan(gpu, pass, P, n, &pp);
}

static hipError_t pool_launch(lb_gpu *gpu, int pass, const PoolPlan &pp, int src, bool stats,
                              PoolArgs A) {
    const Tuning &T = gpu->tuning[pass];
    PoolFn fn = pick_pool(src, T.nodes, stats);
    uint32_t grid = 0;
    hipError_t e = pool_grid(gpu, fn, pp.block, pp.lds, &grid);
    if (e != hipSuccess) return e;
    if (grid > pp.grid) return hipErrorInvalidConfiguration;  // beyond the spill area
    A.hits = static_cast<uint2 *>(gpu->stream_hits.ptr);
    A.queue = gpu->queue;
    A.spill = static_cast<uint32_t *>(gpu->spill.ptr);
    A.depth = (uint32_t)T.lds_stack;
    A.refill = (uint32_t)T.refill;
    A.leaf_batch = (uint32_t)T.leaf_batch;
    A.stats = stats ? gpu->counters : nullptr;
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, gpu->stream, gpu->queue, 1u);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(pp.block), pp.lds, gpu->stream, A);
    return hipSuccess;
}

hipError_t launch_stream(lb_gpu *gpu, int pass, bool stats, const SceneView &S,
                         const lb_params &P, const lb_texel *items, uint32_t n, uint32_t key_base,
                         float4 *out) {
    PoolPlan pp;
    hipError_t e = pool_plan(gpu, pass, P, n, &pp);
    if (e != hipSuccess) return e;
    const uint2 *hits = static_cast<const uint2 *>(gpu->stream_hits.ptr);
    uint32_t *aux = static_cast<uint32_t *>(gpu->stream_aux.ptr);
    unsigned long long *ctr = stats ? gpu->counters : nullptr;
    hipStream_t st = gpu->stream;
    const dim3 blk(256);
    for (uint32_t c0 = 0; c0 < n; c0 += pp.chunk) {
        uint32_t nc = n - c0 < pp.chunk ? n - c0 : pp.chunk;
        dim3 items_grid((nc + 255u) / 256u);
        PoolArgs A{};
        A.S = S;
        A.P = P;
        A.items = items + c0;
        A.n = nc;
        A.key0 = key_base + c0;
        if (pass == LB_PASS_GATHER) {
            A.slots = P.counts[2];
            if ((e = pool_launch(gpu, pass, pp, SRC_GATHER, stats, A)) != hipSuccess) return e;
            if (stats)
                hipLaunchKernelGGL(k_shade_gather<true>, items_grid, blk, 0, st, S, P, A.items,
                                   nc, A.key0, hits, out + c0, ctr);
            else
                hipLaunchKernelGGL(k_shade_gather<false>, items_grid, blk, 0, st, S, P, A.items,
                                   nc, A.key0, hits, out + c0, ctr);
        } else if (pass == LB_PASS_VOLUME) {
            A.slots = P.counts[2] + P.counts[0] * P.counts[0];
            if ((e = pool_launch(gpu, pass, pp, SRC_VOLUME, stats, A)) != hipSuccess) return e;
            if (stats)
                hipLaunchKernelGGL(k_shade_volume<true>, items_grid, blk, 0, st, S, P, A.items,
                                   nc, hits, out + (size_t)c0 * 4u, ctr);
            else
                hipLaunchKernelGGL(k_shade_volume<false>, items_grid, blk, 0, st, S, P, A.items,
                                   nc, hits, out + (size_t)c0 * 4u, ctr);
        } else {
            uint32_t *pick = aux;
            uint32_t *list[2] = {aux + pp.chunk, aux + 2u * (size_t)pp.chunk};
            uint32_t *counts = aux + 3u * (size_t)pp.chunk;  // round q's item count at [q]
            hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, counts, 16u);
            for (uint32_t q = 0; q < 10u; q++) {
                const uint32_t *cur = q == 0u ? nullptr : list[q & 1u];
                const uint32_t *count = q == 0u ? nullptr : counts + q;
                PoolArgs B = A;
                B.slots = 16u;
                B.q = q;
                B.list = cur;
                B.n_dev = count;
                if ((e = pool_launch(gpu, pass, pp, SRC_INSIDE, stats, B)) != hipSuccess) return e;
                hipLaunchKernelGGL(k_settle_inside, items_grid, blk, 0, st, P, nc, q, cur, count,
                                   hits, list[(q + 1u) & 1u], counts + q + 1u, pick);
            }
            A.slots = 2u * P.counts[0] * P.counts[0] + P.counts[1];
            A.pick = pick;
            if ((e = pool_launch(gpu, pass, pp, SRC_DIRECT, stats, A)) != hipSuccess) return e;
            hipLaunchKernelGGL(k_shade_direct, items_grid, blk, 0, st, P, A.items, nc, pick, hits,
                               out + c0);
        }
    }
    return hipGetLastError();
}

}  // namespace lbk
