// The STREAMED schedule of the TUNED form (schedule 2): a pass becomes ray generation, one trace
// kernel that knows nothing of the pass, and a shading kernel, with the rays and their hits in
// HBM between them.
//
//   generate  one thread an item and a run of its ray slots writes each ray's record;
//   trace     a persistent grid; every lane takes the next ray from one global queue as soon as
//             its previous ray ended (a wave refills once `refill` of its lanes are idle), and
//             runs the reference's traversal (lb_traverse.h) as while-while: inner nodes until
//             every busy lane stands at a leaf or is done, then the leaves;
//   shade     one thread an item folds its hits in the reference's ray order with the
//             reference's arithmetic, so every output matches the PLAIN form bit for bit.
//
// A lane's work no longer waits for the longest ray of its item or of its wave: the trace kernel
// keeps lanes busy across items. The price is 32 B of ray record written and read twice and 8 B
// of hit written and read once a ray.
//
// Records: a ray is (o.xyz, tmax) then (d.xyz, stops() key); tmax < 0 marks an any-hit ray
// (|tmax| is its range) and tmax == 0 an empty slot that traces nothing. A hit is (bits of t,
// negated for a back face; triangle, NONE for a miss). Ray id = slot * items + item (slot-major),
// so generation writes and shading reads are coalesced over items.
//
// The direct pass runs its inside test (trace.wgsl:159-168, 16 rays a point) as rounds: round 0
// tests every item's texel centre, round q (1..9) tests point q of the items still inside after
// round q-1, exactly the points the reference tests, then one main round traces the sun and
// occlusion rays from each item's chosen point.
#include "lb_host.h"
#include "lb_traverse.h"

namespace lbk {

constexpr uint32_t GEN_RUN = 16;  // ray slots one generating thread writes
constexpr size_t MAX_STREAM_RAYS = size_t(1) << 26;  // rays a chunk holds (2 GiB of records)

struct TraceArgs {
    SceneView S;
    lb_params P;
    const float4 *rays;
    uint2 *hits;
    uint32_t n;              // rays, unless n_dev is set
    const uint32_t *n_dev;   // device-side item count; rays = *n_dev * n_mul
    uint32_t n_mul;
    uint32_t *queue;
    uint32_t *spill;
    uint32_t depth, refill;
    unsigned long long *stats;
};

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <int LAYOUT, bool STATS>
__global__ void k_trace_stream(TraceArgs A) {
    StackView st = make_stack(A.spill, A.depth);
    uint32_t lane = threadIdx.x & 63u;
    uint32_t n = A.n_dev ? *A.n_dev * A.n_mul : A.n;
    Counts cnt{0, 0, 0, 0, 0, 0};
    Ray R;
    uint32_t id = 0;
    bool active = false;
    bool drained = false;  // wave-uniform: the queue holds no more rays
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (!drained && (__popcll(idle) >= (int)A.refill || idle == ~0ull)) {
            uint32_t want = (uint32_t)__popcll(idle);
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(A.queue, want);
            base = __builtin_amdgcn_readfirstlane(base);
            // base stays below n + 64 * (waves of the grid): no wrap for n < 2^31
            if (base + want >= n) drained = true;
            if (!active) {
                uint32_t my = base + lanes_below(idle);
                if (my < n) {
                    float4 a = A.rays[2u * (size_t)my], b = A.rays[2u * (size_t)my + 1u];
                    if (a.w == 0.0f) {
                        A.hits[my] = make_uint2(0u, NONE);
                    } else {
                        ray_start(R, A.S.root, xyz(a), xyz(b), fabsf(a.w), a.w < 0.0f, fbits(b.w));
                        if (STATS) cnt.rays++;
                        id = my;
                        active = true;
                    }
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (drained) break;
            continue;
        }
        if (active) {
            bool done = false;
            while (!(R.n & LEAF_BIT)) {
                if (!inner_visit<LAYOUT, STATS>(A.S, R, st, cnt) && pop(R, st)) {
                    done = true;
                    break;
                }
            }
            if (!done) done = leaf_tests<STATS>(A.S, A.P, R, cnt) || pop(R, st);
            if (done) {
                A.hits[id] = make_uint2(fbits(R.back ? -R.t : R.t), R.tri);
                active = false;
            }
        }
    }
    if (STATS) flush_counts(cnt, A.stats);
}

__device__ __forceinline__ Hit hit_of(uint2 h) {
    return Hit{fabsf(bitsf(h.x)), h.y, (h.x >> 31) != 0u};
}

struct Item {
    f3 pos, n, dx, dy;
    uint32_t w;  // normal.w bits
};
__device__ __forceinline__ Item item_at(const lb_texel *items, uint32_t i) {
    const float4 *tp = reinterpret_cast<const float4 *>(items + i);
    float4 a = tp[0], b = tp[1], c = tp[2], d = tp[3];
    return Item{xyz(a), xyz(b), xyz(c), xyz(d), fbits(b.w)};
}

__device__ __forceinline__ void put_ray(float4 *rays, size_t id, f3 o, float tmax, f3 d,
                                        uint32_t key) {
    rays[2u * id] = make_float4(o.x, o.y, o.z, tmax);
    rays[2u * id + 1u] = make_float4(d.x, d.y, d.z, bitsf(key));
}

// ---- gather (trace.wgsl:245-270) ------------------------------------------------------------

__global__ void k_gen_gather(lb_params P, const lb_texel *items, uint32_t nc, uint32_t key0,
                             float4 *rays) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    uint32_t total = P.counts[2];
    uint32_t r0 = blockIdx.y * GEN_RUN;
    Item tx = item_at(items, li);
    frame_t f = frame(tx.n);
    uint32_t key = (key0 + li) * 7919u;
    uint32_t seed = fbits(P.reach[3]);
    float s1 = rnd(key, 11u, seed), s2 = rnd(key, 13u, seed);
    for (uint32_t r = r0; r < r0 + GEN_RUN && r < total; r++) {
        float jx = rnd(key + r, 19u, seed) - 0.5f;
        float jy = rnd(key + r, 23u, seed) - 0.5f;
        f3 o = add(add(add(tx.pos, scl(tx.dx, jx)), scl(tx.dy, jy)), scl(tx.n, P.reach[2]));
        f3 d = cosine(fract(((float)r + 0.5f) / (float)total + s1), fract(radical(r) + s2), f);
        put_ray(rays, (size_t)r * nc + li, o, P.reach[1], d, key + r + 4000u);
    }
}

template <bool STATS>
__global__ void k_shade_gather(SceneView S, lb_params P, uint32_t nc, const float4 *rays,
                               const uint2 *hits, float4 *out, unsigned long long *stats) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    Counts cnt{0, 0, 0, 0, 0, 0};
    uint32_t total = P.counts[2];
    f3 sum = mk(0.0f, 0.0f, 0.0f);
    float back = 0.0f;
    for (uint32_t r = 0u; r < total; r++) {
        size_t id = (size_t)r * nc + li;
        Hit h = hit_of(hits[id]);
        f3 d = xyz(rays[2u * id + 1u]);
        if (h.tri == NONE) {
            sum = add(sum, sky(P, d));
            continue;
        }
        if (h.back) {
            back += 1.0f;
            continue;
        }
        if (STATS) {
            cnt.mats++;
            cnt.gathers++;
        }
        f3 o = xyz(rays[2u * id]);
        sum = add(sum, bounced(S, P, o, d, h));
    }
    out[li] = make_float4(sum.x * PI / (float)total, sum.y * PI / (float)total,
                          sum.z * PI / (float)total, back / (float)total);
    if (STATS) flush_counts(cnt, stats);
}

// ---- volume (trace.wgsl:272-324) ------------------------------------------------------------

__global__ void k_gen_volume(lb_params P, const lb_texel *items, uint32_t nc, float4 *rays) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    uint32_t nr = P.counts[2];
    uint32_t g = P.counts[0], gg = g * g;
    uint32_t total = nr + gg;
    uint32_t r0 = blockIdx.y * GEN_RUN;
    Item pr = item_at(items, li);
    uint32_t key = pr.w * 7919u + 77u;
    uint32_t seed = fbits(P.reach[3]);
    float s1 = rnd(key, 29u, seed), s2 = rnd(key, 31u, seed);
    f3 sund = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    frame_t sf = frame(sund);
    for (uint32_t slot = r0; slot < r0 + GEN_RUN && slot < total; slot++) {
        size_t id = (size_t)slot * nc + li;
        if (slot < nr) {
            float u = fract(((float)slot + 0.5f) / (float)nr + s1);
            float v = fract(radical(slot) + s2);
            float z = 1.0f - 2.0f * u;
            float rr = sqrtf(vmax(0.0f, 1.0f - z * z));
            float a = v * 2.0f * PI;
            float sa, ca;
            sincos_lb(a, &sa, &ca);
            put_ray(rays, id, pr.pos, P.reach[1], mk(rr * ca, z, rr * sa), key + slot);
        } else {
            uint32_t s = slot - nr;
            float a = (float)s * 2.399963f;
            float r = P.sun_dir[3] * sqrtf(((float)s + 0.5f) / (float)gg);
            float sa, ca;
            sincos_lb(a, &sa, &ca);
            f3 d = normalize3(add(add(sund, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
            put_ray(rays, id, pr.pos, -FAR, d, key + s + 9000u);
        }
    }
}

template <bool STATS>
__global__ void k_shade_volume(SceneView S, lb_params P, uint32_t nc, const float4 *rays,
                               const uint2 *hits, float4 *out, unsigned long long *stats) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    Counts cnt{0, 0, 0, 0, 0, 0};
    uint32_t nr = P.counts[2];
    uint32_t g = P.counts[0], gg = g * g;
    f3 l0 = mk(0, 0, 0), l1x = mk(0, 0, 0), l1y = mk(0, 0, 0), l1z = mk(0, 0, 0);
    float back = 0.0f;
    for (uint32_t r = 0u; r < nr; r++) {
        size_t id = (size_t)r * nc + li;
        Hit h = hit_of(hits[id]);
        f3 d = xyz(rays[2u * id + 1u]);
        f3 l = sky(P, d);
        if (h.tri != NONE) {
            if (h.back) back += 1.0f;
            if (STATS && !h.back) {
                cnt.mats++;
                cnt.gathers++;
            }
            l = h.back ? mk(0.0f, 0.0f, 0.0f) : bounced(S, P, xyz(rays[2u * id]), d, h);
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

// ---- direct (trace.wgsl:170-243) ------------------------------------------------------------

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

struct InsideArgs {
    lb_params P;
    const lb_texel *items;
    uint32_t nc, key0, q;
    const uint32_t *list;  // round q's items (NULL in round 0: every item)
    const uint32_t *count; // round q's item count (NULL in round 0: nc)
    uint32_t *next;        // the items still inside after round q
    uint32_t *next_count;
    uint32_t *pick;        // per item: the point it settles on
    float4 *rays;
    const uint2 *hits;
};

// 16 rays a listed item, item-major (id = k * 16 + r)
__global__ void k_gen_inside(InsideArgs A) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = A.count ? *A.count : A.nc;
    if (k >= m) return;
    uint32_t li = A.list ? A.list[k] : k;
    const lb_params &P = A.P;
    Item tx = item_at(A.items, li);
    uint32_t key = (A.key0 + li) * 7919u;
    uint32_t kq;
    f3 p = inside_point(tx, A.q, key, &kq);
    f3 o = add(p, scl(tx.n, P.reach[2]));
    frame_t f = frame(tx.n);
    for (uint32_t r = 0u; r < 16u; r++) {
        f3 d = cosine(((float)r + 0.5f) / 16.0f, radical(r), f);
        put_ray(A.rays, (size_t)k * 16u + r, o, P.ground[3], d, kq + r + 7000u);
    }
}

// trace.wgsl:166-167 on round q's hits: an item outside at point q settles on it (round 0: on
// the centre); one still inside goes on to point q + 1, and after point 9 settles on the centre.
__global__ void k_settle_inside(InsideArgs A) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = A.count ? *A.count : A.nc;
    if (k >= m) return;
    uint32_t li = A.list ? A.list[k] : k;
    float back = 0.0f;
    for (uint32_t r = 0u; r < 16u; r++) {
        uint2 h = A.hits[(size_t)k * 16u + r];
        if (h.y != NONE && (h.x >> 31) != 0u) back += 1.0f;
    }
    bool inside = back / 16.0f > bitsf(A.P.atlas[3]);
    if (!inside) {
        A.pick[li] = A.q;
    } else if (A.q < 9u) {
        uint32_t slot = atomicAdd(A.next_count, 1u);
        A.next[slot] = li;
    } else {
        A.pick[li] = 0u;
    }
}

// Sun slots 0..gg-1 (hard), gg..2gg-1 (soft) and occlusion slots after them; an unlit item's sun
// slots stay empty.
__global__ void k_gen_direct(lb_params P, const lb_texel *items, uint32_t nc, uint32_t key0,
                             const uint32_t *pick, float4 *rays) {
    uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nc) return;
    uint32_t g = P.counts[0], gg = g * g;
    uint32_t nr = P.counts[1];
    uint32_t total = 2u * gg + nr;
    uint32_t s0 = blockIdx.y * GEN_RUN;
    Item tx = item_at(items, li);
    uint32_t key = (key0 + li) * 7919u;
    uint32_t seed = fbits(P.reach[3]);
    uint32_t q = pick[li];
    uint32_t kq;
    f3 at = inside_point(tx, q, key, &kq);
    float span = q == 0u ? 1.0f : 0.4f;
    f3 lift = scl(tx.n, P.reach[2]);
    f3 sun = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    bool lit = dot3(sun, tx.n) > 0.0f;
    frame_t sf = frame(sun);
    frame_t f = frame(tx.n);
    float kr = rnd(key, 17u, seed);
    float s1 = rnd(key, 3u, seed), s2 = rnd(key, 5u, seed);
    f3 o_ao = add(at, lift);
    for (uint32_t slot = s0; slot < s0 + GEN_RUN && slot < total; slot++) {
        size_t id = (size_t)slot * nc + li;
        if (slot < 2u * gg) {
            if (!lit) {
                put_ray(rays, id, mk(0.0f, 0.0f, 0.0f), 0.0f, mk(0.0f, 0.0f, 0.0f), 0u);
                continue;
            }
            uint32_t s = slot < gg ? slot : slot - gg;
            float cx = (((float)(s % g) + 0.5f) / (float)g - 0.5f) * span;
            float cy = (((float)(s / g) + 0.5f) / (float)g - 0.5f) * span;
            f3 o = add(add(add(at, scl(tx.dx, cx)), scl(tx.dy, cy)), lift);
            if (slot < gg) {
                put_ray(rays, id, o, -FAR, sun, key + s);
            } else {
                uint32_t k = (s + (uint32_t)(kr * (float)gg)) % gg;
                float a = (float)k * 2.399963f;
                float r = P.sun_dir[3] * sqrtf(((float)k + 0.5f) / (float)gg);
                float sa, ca;
                sincos_lb(a, &sa, &ca);
                f3 d = normalize3(add(add(sun, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
                put_ray(rays, id, o, -FAR, d, key + s + 1000u);
            }
        } else {
            uint32_t r = slot - 2u * gg;
            f3 d = cosine(fract(((float)r + 0.5f) / (float)nr + s1), fract(radical(r) + s2), f);
            put_ray(rays, id, o_ao, P.reach[0], d, key + r + 2000u);
        }
    }
}

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

using TraceFn = void (*)(TraceArgs);

static TraceFn pick_trace(int layout, bool stats) {
    if (layout == LB_NODES_WIDE) return stats ? k_trace_stream<0, true> : k_trace_stream<0, false>;
    return stats ? k_trace_stream<1, true> : k_trace_stream<1, false>;
}

// Ray slots an item takes in its largest round, and items a chunk holds.
static uint32_t slots_per_item(int pass, const lb_params &P) {
    uint32_t gg = P.counts[0] * P.counts[0];
    if (pass == LB_PASS_GATHER) return P.counts[2];
    if (pass == LB_PASS_VOLUME) return P.counts[2] + gg;
    uint32_t main = 2u * gg + P.counts[1];
    return main > 16u ? main : 16u;
}

struct StreamPlan {
    TraceFn fn;
    uint32_t block, grid, chunk;
    size_t lds;
};

static hipError_t stream_plan(lb_gpu *gpu, int pass, bool stats, const lb_params &P, uint32_t n,
                              StreamPlan *out) {
    const Tuning &T = gpu->tuning;
    out->fn = pick_trace(T.nodes, stats);
    out->block = (uint32_t)T.workgroup;
    out->lds = (size_t)T.lds_stack * out->block * 4u;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(out->fn), (int)out->block, out->lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    out->grid = (uint32_t)gpu->cus * (uint32_t)per_cu;
    uint32_t per = slots_per_item(pass, P);
    size_t cap = T.chunk > 0 ? (size_t)T.chunk : MAX_STREAM_RAYS;
    size_t chunk = cap / per;
    if (chunk < 1) chunk = 1;
    out->chunk = (uint32_t)(chunk < n ? chunk : (n ? n : 1u));
    size_t rays = (size_t)out->chunk * per;
    size_t spill = (size_t)out->grid * out->block * (size_t)(STACK - T.lds_stack) * 4u;
    // aux words: pick, two item lists, two counters (direct pass)
    size_t aux = ((size_t)out->chunk * 3u + 64u) * 4u;
    if (gpu->spill.reserve(spill < 16 ? 16 : spill) || gpu->stream_rays.reserve(rays * 32u) ||
        gpu->stream_hits.reserve(rays * 8u) || gpu->stream_aux.reserve(aux))
        return hipErrorOutOfMemory;
    return hipSuccess;
}

hipError_t stream_prepare(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n) {
    StreamPlan sp;
    hipError_t e = stream_plan(gpu, pass, false, P, n, &sp);
    if (e != hipSuccess) return e;
    return stream_plan(gpu, pass, true, P, n, &sp);
}

static void trace_launch(lb_gpu *gpu, const StreamPlan &sp, const SceneView &S,
                         const lb_params &P, uint32_t n, const uint32_t *n_dev, uint32_t n_mul,
                         bool stats) {
    TraceArgs A;
    A.S = S;
    A.P = P;
    A.rays = static_cast<const float4 *>(gpu->stream_rays.ptr);
    A.hits = static_cast<uint2 *>(gpu->stream_hits.ptr);
    A.n = n;
    A.n_dev = n_dev;
    A.n_mul = n_mul;
    A.queue = gpu->queue;
    A.spill = static_cast<uint32_t *>(gpu->spill.ptr);
    A.depth = (uint32_t)gpu->tuning.lds_stack;
    A.refill = (uint32_t)gpu->tuning.refill;
    A.stats = stats ? gpu->counters : nullptr;
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, gpu->stream, gpu->queue, 1u);
    hipLaunchKernelGGL(sp.fn, dim3(sp.grid), dim3(sp.block), sp.lds, gpu->stream, A);
}

hipError_t launch_stream(lb_gpu *gpu, int pass, bool stats, const SceneView &S,
                         const lb_params &P, const lb_texel *items, uint32_t n, uint32_t key_base,
                         float4 *out) {
    StreamPlan sp;
    hipError_t e = stream_plan(gpu, pass, stats, P, n, &sp);
    if (e != hipSuccess) return e;
    float4 *rays = static_cast<float4 *>(gpu->stream_rays.ptr);
    const uint2 *hits = static_cast<const uint2 *>(gpu->stream_hits.ptr);
    uint32_t *aux = static_cast<uint32_t *>(gpu->stream_aux.ptr);
    unsigned long long *ctr = stats ? gpu->counters : nullptr;
    hipStream_t st = gpu->stream;
    const dim3 blk(256);
    for (uint32_t c0 = 0; c0 < n; c0 += sp.chunk) {
        uint32_t nc = n - c0 < sp.chunk ? n - c0 : sp.chunk;
        const lb_texel *it = items + c0;
        dim3 items_grid((nc + 255u) / 256u);
        if (pass == LB_PASS_GATHER) {
            uint32_t total = P.counts[2];
            hipLaunchKernelGGL(k_gen_gather, dim3(items_grid.x, (total + GEN_RUN - 1) / GEN_RUN),
                               blk, 0, st, P, it, nc, key_base + c0, rays);
            trace_launch(gpu, sp, S, P, nc * total, nullptr, 0, stats);
            if (stats)
                hipLaunchKernelGGL(k_shade_gather<true>, items_grid, blk, 0, st, S, P, nc, rays,
                                   hits, out + c0, ctr);
            else
                hipLaunchKernelGGL(k_shade_gather<false>, items_grid, blk, 0, st, S, P, nc, rays,
                                   hits, out + c0, ctr);
        } else if (pass == LB_PASS_VOLUME) {
            uint32_t total = P.counts[2] + P.counts[0] * P.counts[0];
            hipLaunchKernelGGL(k_gen_volume, dim3(items_grid.x, (total + GEN_RUN - 1) / GEN_RUN),
                               blk, 0, st, P, it, nc, rays);
            trace_launch(gpu, sp, S, P, nc * total, nullptr, 0, stats);
            if (stats)
                hipLaunchKernelGGL(k_shade_volume<true>, items_grid, blk, 0, st, S, P, nc, rays,
                                   hits, out + (size_t)c0 * 4u, ctr);
            else
                hipLaunchKernelGGL(k_shade_volume<false>, items_grid, blk, 0, st, S, P, nc, rays,
                                   hits, out + (size_t)c0 * 4u, ctr);
        } else {
            uint32_t *pick = aux;
            uint32_t *list[2] = {aux + sp.chunk, aux + 2u * (size_t)sp.chunk};
            uint32_t *counts = aux + 3u * (size_t)sp.chunk;  // one word a round
            hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, counts, 16u);
            for (uint32_t q = 0; q < 10u; q++) {
                InsideArgs A;
                A.P = P;
                A.items = it;
                A.nc = nc;
                A.key0 = key_base + c0;
                A.q = q;
                A.list = q == 0u ? nullptr : list[q & 1u];
                A.count = q == 0u ? nullptr : counts + q;
                A.next = list[(q + 1u) & 1u];
                A.next_count = counts + q + 1u;
                A.pick = pick;
                A.rays = rays;
                A.hits = hits;
                hipLaunchKernelGGL(k_gen_inside, items_grid, blk, 0, st, A);
                trace_launch(gpu, sp, S, P, nc * 16u, A.count, 16u, stats);
                hipLaunchKernelGGL(k_settle_inside, items_grid, blk, 0, st, A);
            }
            uint32_t total = 2u * P.counts[0] * P.counts[0] + P.counts[1];
            hipLaunchKernelGGL(k_gen_direct, dim3(items_grid.x, (total + GEN_RUN - 1) / GEN_RUN),
                               blk, 0, st, P, it, nc, key_base + c0, pick, rays);
            trace_launch(gpu, sp, S, P, nc * total, nullptr, 0, stats);
            hipLaunchKernelGGL(k_shade_direct, items_grid, blk, 0, st, P, it, nc, pick, hits,
                               out + c0);
        }
    }
    return hipGetLastError();
}

}  // namespace lbk
