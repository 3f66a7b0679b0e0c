// The TUNED form of the three ray passes (crates/lightbaker/src/gpu/trace.wgsl) for gfx950, in
// its lock-step schedule (schedule 0); the pooled schedule (1) is in trace_stream.hip.
//
// Mapping: a wave64 takes work units from one global atomic queue until it is empty (a persistent
// grid sized to the device's residency). A unit is 64 / L items; L lanes of the wave (a group,
// L a power of two in 8..64) share one item and spread its ray slots over themselves, and the
// slots run in azimuth order (consecutive slots sweep the azimuth), so the rays a wave traces
// together point into one wedge. Each lane sums its rays with Kahan-Neumaier compensation and the
// group combines the (sum, compensation) pairs with ds_swizzle butterflies (xor 1..16) and one
// cross-half exchange (xor 32). Counts (shadow bits, back faces, sun hits) are exact integers, so
// they equal the reference's whatever the order.
//
// Traversal: the reference's order exactly (near child first on ta <= tb, far child pushed,
// 48 entries, a push beyond them dropped) over a node record that holds both children's boxes,
// FP32 (WIDE, 64 B a visit) or 8-bit outward-rounded (QUANT, 32 B a visit). The stack's first
// `lds_stack` entries live in LDS; deeper entries spill to a per-lane global area.
#include <type_traits>

#include "lb_host.h"
#include "lb_records.h"
#include "lb_traverse.h"

namespace lbk {

// Words of LDS a group of lanes keeps its item's state in, after the workgroup's stacks.
constexpr uint32_t SLOT_WORDS = 36;

// This lane's group's slot: the dynamic LDS holds the stacks, then one slot a group.
__device__ __forceinline__ lds_u32 *group_slot(uint32_t depth, uint32_t lanes) {
    extern __shared__ uint32_t lds_stack[];
    lds_u32 *slots = (lds_u32 *)lds_stack + lane_stride(depth) * blockDim.x;
    return slots + (threadIdx.x / lanes) * SLOT_WORDS;
}
__device__ __forceinline__ void put_f3(lds_u32 *s, f3 v) {
    s[0] = fbits(v.x);
    s[1] = fbits(v.y);
    s[2] = fbits(v.z);
}
__device__ __forceinline__ f3 get_f3(const lds_u32 *s) {
    return mk(bitsf(s[0]), bitsf(s[1]), bitsf(s[2]));
}

// trace.wgsl:83-130 over the tuned node records, to the end, the lanes of a wave in step: they
// go through inner nodes until leaf_batch of the lanes still tracing stand at a leaf (64: until
// none is at an inner node), then those take their leaves, and again. The count is the exec
// mask's population, a scalar operation, so a step costs no extra vector instruction; it is
// taken every VISITS inner visits (a lane that reaches a leaf before then waits for the count),
// which saves the loop's scalar bookkeeping between them. A ray that ends (a missed pair with an
// empty stack, an any-hit ray's first hit) takes the ENDED reference and leaves the loops through
// their own conditions: a break or return from inside a divergent loop costs the exec-mask
// bookkeeping of another exit on every step (map gather launch 10.31 -> 9.85 ms, the city's
// step 1 993 -> 1 866 ms). Each lane still takes its own ray's steps in the reference's order.
template <int LAYOUT, bool STATS, int VISITS, typename C>
__device__ __forceinline__ void trace_locked(const SceneView &S, const lb_params &P, Ray &R,
                                             const StackView &st, C &cnt, int leaf_batch) {
    static_assert(VISITS >= 1 && VISITS <= 4, "1 to 4 visits between counts");
    // an inner visit, and a pop when it missed both children; with the stack empty the ray ends,
    // marked by a reference no node has (node indices stay below 2^26 - 1), which the leaf test
    // ends the inner loop on as it does for a leaf
    constexpr uint32_t ENDED = 0xffffffffu;
    auto step = [&] {
        if (!inner_visit<LAYOUT, STATS>(S, R, st, cnt) && pop(R, st)) R.n = ENDED;
    };
    while (R.n != ENDED) {
        // lanes that may stay at inner nodes before the wave turns to the leaves
        int keep = __popcll(__ballot(1)) - leaf_batch;
        while (!(R.n & LEAF_BIT)) {
            step();
            // the further visits written out: a loop over them compiles to slower code
            if constexpr (VISITS >= 2) {
                if (R.n & LEAF_BIT) break;
                step();
            }
            if constexpr (VISITS >= 3) {
                if (R.n & LEAF_BIT) break;
                step();
            }
            if constexpr (VISITS >= 4) {
                if (R.n & LEAF_BIT) break;
                step();
            }
            if (__popcll(__ballot(1)) <= keep) break;
        }
        if ((R.n & LEAF_BIT) && R.n != ENDED && (leaf_tests<STATS>(S, P, R, cnt) || pop(R, st)))
            R.n = ENDED;
    }
}

template <int LAYOUT, bool STATS, int VISITS, typename C>
__device__ Hit trace_t(const SceneView &S, const lb_params &P, uint32_t root, f3 o, f3 d,
                       float tmax, bool any, uint32_t ray, const StackView &st, C &cnt,
                       int leaf_batch) {
    Ray R;
    ray_start(R, root, o, d, tmax, any, ray);
    if (STATS) cnt.rays++;
    trace_locked<LAYOUT, STATS, VISITS>(S, P, R, st, cnt, leaf_batch);
    return Hit{R.t, R.tri, R.back};
}

// Inner visits between the lock-step loop's counts of its lanes (trace_locked): 3 took the map's
// ranked gather launch from 10.84 to 10.23 ms and its direct pass from 11.40 to 10.73 ms (2 and
// 4 gave less on the city). The gather's counting launch, whose rays are unranked and part
// sooner, takes COUNTING_VISITS: with a break out of the loops on an ended ray 2 made it slower
// and 3 slower still (15.3 -> 17.3 ms); with the ended ray's sentinel, 2 shortened the map's
// bounces by 0.7 ms and the city's by 11 ms.
constexpr int LOCKED_VISITS = 3;
constexpr int COUNTING_VISITS = 2;

// ---- group reductions -------------------------------------------------------------------------

template <int M>
__device__ __forceinline__ float swz(float v) {
    // ds_swizzle bitmask mode: and 0x1f, or 0, xor M (within each 32-lane half)
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1f | (M << 10)));
}
template <int M>
__device__ __forceinline__ uint32_t swzu(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1f | (M << 10));
}
__device__ __forceinline__ float x32(float v) { return __shfl_xor(v, 32); }
__device__ __forceinline__ uint32_t x32u(uint32_t v) { return (uint32_t)__shfl_xor((int)v, 32); }

// Kahan-Neumaier running sum
struct KSum {
    float s, c;
    __device__ __forceinline__ void add(float x) {
        float t = s + x;
        if (fabsf(s) >= fabsf(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    // TwoSum merge of another compensated sum
    __device__ __forceinline__ void merge(float s2, float c2) {
        float t = s + s2;
        float bp = t - s;
        float err = (s - (t - bp)) + (s2 - bp);
        s = t;
        c = c + c2 + err;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

template <int M>
__device__ __forceinline__ void kstep(KSum &k) {
    float s2 = swz<M>(k.s), c2 = swz<M>(k.c);
    k.merge(s2, c2);
}

// Every lane of an aligned group of L lanes ends with the group's compensated total.
__device__ __forceinline__ void group_ksum(KSum &k, int L) {
    if (L > 1) kstep<1>(k);
    if (L > 2) kstep<2>(k);
    if (L > 4) kstep<4>(k);
    if (L > 8) kstep<8>(k);
    if (L > 16) kstep<16>(k);
    if (L > 32) {
        float s2 = x32(k.s), c2 = x32(k.c);
        k.merge(s2, c2);
    }
}
__device__ __forceinline__ float group_sum(float v, int L) {
    if (L > 1) v += swz<1>(v);
    if (L > 2) v += swz<2>(v);
    if (L > 4) v += swz<4>(v);
    if (L > 8) v += swz<8>(v);
    if (L > 16) v += swz<16>(v);
    if (L > 32) v += x32(v);
    return v;
}
__device__ __forceinline__ uint32_t group_or(uint32_t v, int L) {
    if (L > 1) v |= swzu<1>(v);
    if (L > 2) v |= swzu<2>(v);
    if (L > 4) v |= swzu<4>(v);
    if (L > 8) v |= swzu<8>(v);
    if (L > 16) v |= swzu<16>(v);
    if (L > 32) v |= x32u(v);
    return v;
}

// ---- the persistent frame -----------------------------------------------------------------------

struct TunedArgs {
    SceneView S;
    lb_params P;
    const lb_texel *items;
    uint32_t n_items, key_base;
    float4 *out;
    unsigned long long *stats;
    uint32_t *queue;
    uint32_t *spill;
    uint32_t lanes_per_item, depth;
    int leaf_batch;  // lanes at a leaf that end a wave's inner phase early (64: never)
    uint8_t *cost;        // gather, counting launch: each ray's steps by azimuth slot (or null)
    const uint8_t *perm;  // gather: slot s of item i takes azimuth slot perm[i * rays + s]
    uint2 *rec;           // gather: each ray's record at r * n_items + i (or null)
    uint32_t xcd_split;   // units split over the chiplets' queues (next_unit_split)
    uint32_t el_order;    // gather without perm: slots in elevation bands (el_az_slot)
};


// Takes the next unit for the whole wave; UINT_MAX when the queue is empty.
__device__ __forceinline__ uint32_t next_unit(uint32_t *queue, uint32_t n_units) {
    uint32_t u = 0;
    if ((threadIdx.x & 63u) == 0u) u = atomicAdd(queue, 1u);
    u = __builtin_amdgcn_readfirstlane(u);
    return u < n_units ? u : 0xffffffffu;
}

// The units split into XCDS contiguous ranges with a counter each (128 bytes apart). The
// dispatcher deals workgroups to the chiplets in turn, so workgroup b runs on chiplet b % XCDS:
// a wave takes units from its own chiplet's range, then from the next ranges as they run dry,
// and ends when it comes back round. Neighbouring items trace into neighbouring geometry, so a
// chiplet's L2 holds the part of the scene its range sees. `range` is the wave's current range
// (blockIdx.x % XCDS to start).
constexpr uint32_t XCDS = 8, QUEUE_STRIDE = 32;
__device__ __forceinline__ uint32_t next_unit_split(uint32_t *queue, uint32_t n_units,
                                                    uint32_t &range) {
    uint32_t home = blockIdx.x % XCDS;
    for (;;) {
        uint32_t lo = (uint32_t)((uint64_t)n_units * range / XCDS);
        uint32_t hi = (uint32_t)((uint64_t)n_units * (range + 1u) / XCDS);
        uint32_t u = 0;
        if ((threadIdx.x & 63u) == 0u) u = atomicAdd(queue + range * QUEUE_STRIDE, 1u);
        u = __builtin_amdgcn_readfirstlane(u);
        if (u < hi - lo) return lo + u;
        range = (range + 1u) % XCDS;
        if (range == home) return 0xffffffffu;
    }
}

__device__ __forceinline__ uint32_t take_unit(const TunedArgs &A, uint32_t n_units,
                                              uint32_t &range) {
    return A.xcd_split ? next_unit_split(A.queue, n_units, range) : next_unit(A.queue, n_units);
}

struct ItemIn {
    f3 pos, n, dx, dy;
    uint32_t w;  // normal.w bits
};
__device__ __forceinline__ ItemIn load_item(const lb_texel *items, uint32_t i) {
    const float4 *tp = reinterpret_cast<const float4 *>(items + i);
    float4 a = tp[0], b = tp[1], c = tp[2], d = tp[3];
    return ItemIn{xyz(a), xyz(b), xyz(c), xyz(d), fbits(b.w)};
}

// trace.wgsl:159-168 with the 16 rays over the group, from point q of the item whose state is in
// the group's LDS slot (the texel centre, an offset, or the centre pushed along the normal)
template <int LAYOUT, bool STATS, int VISITS>
__device__ __forceinline__ bool inside_g(const TunedArgs &A, const StackView &st,
                                         const lds_u32 *slot, uint32_t q, uint32_t lig, int L,
                                         Counts &cnt) {
    float back = 0.0f;
    for (uint32_t r = lig; r < 16u; r += (uint32_t)L) {
        // read the slot afresh (the compiler would otherwise keep the stored values live)
        asm volatile("" ::: "memory");
        f3 pos = get_f3(slot), lift = get_f3(slot + 9);
        frame_t f;
        f.t = get_f3(slot + 12);
        f.b = get_f3(slot + 15);
        f.n = get_f3(slot + 18);
        uint32_t key = slot[31], kq = key;
        f3 p = pos;
        if (q > 0u && q < 9u) {
            float cx, cy;
            offset(q, &cx, &cy);
            p = add(add(pos, scl(get_f3(slot + 3), cx)), scl(get_f3(slot + 6), cy));
            kq = key + q * 31u;
        } else if (q == 9u) {
            p = add(pos, scl(f.n, PUSH));
            kq = key + 331u;
        }
        f3 d = cosine(((float)r + 0.5f) / 16.0f, radical(r), f);
        Hit h = trace_t<LAYOUT, STATS, VISITS>(A.S, A.P, slot[35], add(p, lift), d,
                                               A.P.ground[3], false, kq + r + 7000u, st, cnt,
                                               A.leaf_batch);
        if (h.tri != NONE && h.back) back += 1.0f;
    }
    back = group_sum(back, L);
    return back / 16.0f > bitsf(A.P.atlas[3]);
}

// Slot s of an item's `total` rays in azimuth order: for a power of two, slot s is the ray whose
// azimuth parameter v = fract(radical(r) + rot) is the s-th from 0 (r = (s + m0) bit-reversed,
// m0 = -rot * total rounded up), so consecutive slots sweep v from 0 whatever the item's
// rotation `rot`, and the items a wave traces together point into the same wedge of
// directions. Any other count keeps r = s.
__device__ __forceinline__ uint32_t az_slot(uint32_t s, uint32_t total, float rot) {
    if (total < 2u || (total & (total - 1u)) != 0u) return s;
    uint32_t b = 31u - (uint32_t)__builtin_clz(total);
    uint32_t m0 = (total - (uint32_t)(rot * (float)total)) & (total - 1u);
    return __builtin_bitreverse32((s + m0) & (total - 1u)) >> (32u - b);
}

// az_slot's inverse: the slot of ray r
__device__ __forceinline__ uint32_t az_slot_of(uint32_t r, uint32_t total, float rot) {
    if (total < 2u || (total & (total - 1u)) != 0u) return r;
    uint32_t b = 31u - (uint32_t)__builtin_clz(total);
    uint32_t m0 = (total - (uint32_t)(rot * (float)total)) & (total - 1u);
    return ((__builtin_bitreverse32(r) >> (32u - b)) - m0) & (total - 1u);
}

// Slot s of `total` gather rays (a power of two, at least 64) in elevation bands: band s / 32
// holds the 32 rays of the next smallest elevation parameters u = fract((r + 0.5) / total +
// rot_u) (from the normal out to the horizon), and within a band position s % 32 takes the ray
// of the p-th azimuth sector from v = 0, v = fract(radical(r) + rot_v). Any 32 consecutive r hold
// every residue of r mod 32 once, and radical's top 5 bits are those bits reversed, so a band
// holds one ray of each of the 32 sectors. The rays of a round (one band) then end at like
// depths without counting them first.
__device__ __forceinline__ uint32_t el_az_slot(uint32_t s, uint32_t total, float rot_u,
                                               float rot_v) {
    uint32_t mask = total - 1u;
    // the ray of smallest u: where (r + 0.5) / total + rot_u passes 1
    uint32_t r_min = (uint32_t)ceilf((float)total * (1.0f - rot_u) - 0.5f) & mask;
    uint32_t r0 = (r_min + (s & ~31u)) & mask;
    uint32_t m = (32u - (uint32_t)(rot_v * 32.0f)) & 31u;
    uint32_t low = __builtin_bitreverse32((s + m) & 31u) >> 27;
    return (r0 + ((low - r0) & 31u)) & mask;
}

// The slot loop of one item's rays: each lane traces its slot to the end (trace_locked), then
// takes slot + L. start(slot, R) sets up slot's ray; shade(slot, R) folds its result into the
// lane's sums.
template <int LAYOUT, bool STATS, int VISITS, typename C, typename Start, typename Shade>
__device__ __forceinline__ void run_slots(const TunedArgs &A, const StackView &st, uint32_t first,
                                          uint32_t total, uint32_t L, C &cnt, Start start,
                                          Shade shade) {
    Ray R;
    for (uint32_t slot = first; slot < total; slot += L) {
        start(slot, R);
        if (STATS) cnt.rays++;
        trace_locked<LAYOUT, STATS, VISITS>(A.S, A.P, R, st, cnt, A.leaf_batch);
        shade(slot, R);
    }
}

// trace.wgsl:170-243. The inside tests (16 rays, and up to 9 more points) keep the group in
// step; the sun and occlusion rays (2 g^2 + aoRays slots) run in the chosen schedule.
template <int LAYOUT, bool STATS, int WAVES>
__global__ __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1))) void k_direct_tuned(
    TunedArgs A) {
    StackView st = make_stack(A.spill, A.depth);
    const lb_params &P = A.P;
    int L = (int)A.lanes_per_item;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t per_wave = 64u / (uint32_t)L;
    uint32_t grp = lane / (uint32_t)L, lig = lane % (uint32_t)L;
    uint32_t n_units = (A.n_items + per_wave - 1u) / per_wave;
    uint32_t seed = fbits(P.reach[3]);
    uint32_t g = P.counts[0], gg = g * g;
    f3 sun = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    lds_u32 *slot_lds = group_slot(A.depth, (uint32_t)L);
    Counts cnt{0, 0, 0, 0, 0, 0};
    uint32_t range = blockIdx.x % XCDS;
    for (;;) {
        uint32_t unit = take_unit(A, n_units, range);
        if (unit == 0xffffffffu) break;
        uint32_t i = unit * per_wave + grp;
        bool valid = i < A.n_items;
        // the item's state goes to the group's LDS slot and is read back where it is used, so
        // it holds no vector registers across the traversals
        bool lit;
        {
            ItemIn tx = load_item(A.items, valid ? i : 0u);
            lit = dot3(sun, tx.n) > 0.0f;
            // where the occlusion and inside-test rays start (entry_node): their origins lie
            // within 0.45 of the texel's axes from its centre, or PUSH along the normal, all
            // lifted; they reach at most aoReach or the inside reach
            frame_t fi = frame(tx.n);
            f3 c = add(tx.pos, scl(tx.n, P.reach[2]));
            float r = vmax(P.reach[0], P.ground[3]) * 1.001f + 1e-4f;
            f3 e = mk(0.45f * (fabsf(tx.dx.x) + fabsf(tx.dy.x)) + fabsf(fi.n.x) * PUSH + r,
                      0.45f * (fabsf(tx.dx.y) + fabsf(tx.dy.y)) + fabsf(fi.n.y) * PUSH + r,
                      0.45f * (fabsf(tx.dx.z) + fabsf(tx.dy.z)) + fabsf(fi.n.z) * PUSH + r);
            uint32_t entry = entry_node<LAYOUT>(A.S, sub(c, e), add(c, e));
            if (lig == 0u) {
                slot_lds[35] = entry;
                uint32_t key = (A.key_base + i) * 7919u;
                frame_t f = frame(tx.n), sf = frame(sun);
                put_f3(slot_lds, tx.pos);
                put_f3(slot_lds + 3, tx.dx);
                put_f3(slot_lds + 6, tx.dy);
                put_f3(slot_lds + 9, scl(tx.n, P.reach[2]));
                put_f3(slot_lds + 12, f.t);
                put_f3(slot_lds + 15, f.b);
                put_f3(slot_lds + 18, f.n);
                put_f3(slot_lds + 21, sf.t);
                put_f3(slot_lds + 24, sf.b);
                slot_lds[28] = fbits(rnd(key, 17u, seed));
                slot_lds[29] = fbits(rnd(key, 3u, seed));
                slot_lds[30] = fbits(rnd(key, 5u, seed));
                slot_lds[31] = key;
            }
        }
        // trace.wgsl:174-193: the point the rays leave from, 0 (the centre) unless the centre
        // lies inside a solid
        uint32_t pick = 0u;
        if (valid && inside_g<LAYOUT, STATS, LOCKED_VISITS>(A, st, slot_lds, 0u, lig, L, cnt)) {
            for (uint32_t q = 1u; q < 10u; q++)
                if (!inside_g<LAYOUT, STATS, LOCKED_VISITS>(A, st, slot_lds, q, lig, L, cnt)) {
                    pick = q;
                    break;
                }
        }
        if (lig == 0u) {
            asm volatile("" ::: "memory");
            f3 pos = get_f3(slot_lds), at = pos;
            if (pick > 0u && pick < 9u) {
                float cx, cy;
                offset(pick, &cx, &cy);
                at = add(add(pos, scl(get_f3(slot_lds + 3), cx)), scl(get_f3(slot_lds + 6), cy));
            } else if (pick == 9u) {
                at = add(pos, scl(get_f3(slot_lds + 18), PUSH));
            }
            put_f3(slot_lds + 32, at);
            slot_lds[27] = fbits(pick == 0u ? 1.0f : 0.4f);
        }
        uint32_t n_sun = lit ? 2u * gg : 0u;
        uint32_t rays = P.counts[1];
        uint32_t total = valid ? n_sun + rays : 0u;
        uint32_t hard = 0u;
        float soft = 0.0f, back = 0.0f;
        KSum occ{0.0f, 0.0f};
        auto start = [&](uint32_t slot, Ray &R) {
            // read the slot afresh (the compiler would otherwise keep the stored values live)
            asm volatile("" ::: "memory");
            f3 at = get_f3(slot_lds + 32), lift = get_f3(slot_lds + 9);
            uint32_t key = slot_lds[31];
            if (slot < n_sun) {
                f3 dx = get_f3(slot_lds + 3), dy = get_f3(slot_lds + 6);
                float span = bitsf(slot_lds[27]);
                uint32_t s = slot < gg ? slot : slot - gg;
                float cx = (((float)(s % g) + 0.5f) / (float)g - 0.5f) * span;
                float cy = (((float)(s / g) + 0.5f) / (float)g - 0.5f) * span;
                f3 o = add(add(add(at, scl(dx, cx)), scl(dy, cy)), lift);
                if (slot < gg) {
                    ray_start(R, A.S.root, o, sun, FAR, true, key + s);
                } else {
                    frame_t sf;
                    sf.t = get_f3(slot_lds + 21);
                    sf.b = get_f3(slot_lds + 24);
                    float kr = bitsf(slot_lds[28]);
                    uint32_t k = (s + (uint32_t)(kr * (float)gg)) % gg;
                    float a = (float)k * 2.399963f;
                    float r = P.sun_dir[3] * sqrtf(((float)k + 0.5f) / (float)gg);
                    float sa, ca;
                    sincos_lb(a, &sa, &ca);
                    f3 d = normalize3(add(add(sun, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
                    ray_start(R, A.S.root, o, d, FAR, true, key + s + 1000u);
                }
            } else {
                frame_t f;
                f.t = get_f3(slot_lds + 12);
                f.b = get_f3(slot_lds + 15);
                f.n = get_f3(slot_lds + 18);
                float s1 = bitsf(slot_lds[29]), s2 = bitsf(slot_lds[30]);
                uint32_t r = slot - n_sun;
                r = az_slot(r, rays, s2);
                f3 d = cosine(fract(((float)r + 0.5f) / (float)rays + s1), fract(radical(r) + s2), f);
                ray_start(R, slot_lds[35], add(at, lift), d, P.reach[0], false, key + r + 2000u);
            }
        };
        auto shade = [&](uint32_t slot, const Ray &R) {
            if (slot < n_sun) {
                if (R.tri == NONE) {
                    if (slot < gg)
                        hard |= 1u << slot;
                    else
                        soft += 1.0f;
                }
            } else if (R.tri != NONE) {
                float k = R.t / P.reach[0];
                occ.add(1.0f - k * k);
                if (R.back && R.t < P.ground[3]) back += 1.0f;
            }
        };
        run_slots<LAYOUT, STATS, LOCKED_VISITS>(A, st, lig, total, (uint32_t)L, cnt, start,
                                                shade);
        hard = group_or(hard, L);
        soft = group_sum(soft, L);
        back = group_sum(back, L);
        group_ksum(occ, L);
        if (valid && lig == 0u) {
            float sv = lit ? soft / (float)gg : 0.0f;
            A.out[i] = make_float4(sv, bitsf(hard | (pick << 24u)), occ.value() / (float)rays,
                                   back / (float)rays);
        }
    }
    if (STATS) flush_counts(cnt, A.stats);
}

// trace.wgsl:245-270: lane lig of an item's group takes its rays r = lig, lig + L, ...
// WAVES: waves a SIMD the kernel is compiled to keep resident (0: the compiler's choice).
// MODE: 0 plain; 1 the counting launch (each ray's steps into A.cost, nothing else counted);
// 2 the launch's traversal counters as well (lb_stats).
template <int LAYOUT, int MODE, int WAVES>
__global__ __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1))) void k_gather_tuned(
    TunedArgs A) {
    constexpr bool STATS = MODE != 0;
    using C = typename std::conditional<MODE == 2, Counts, Counts32>::type;
    StackView st = make_stack(A.spill, A.depth);
    const lb_params &P = A.P;
    int L = (int)A.lanes_per_item;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t per_wave = 64u / (uint32_t)L;
    uint32_t grp = lane / (uint32_t)L, lig = lane % (uint32_t)L;
    uint32_t n_units = (A.n_items + per_wave - 1u) / per_wave;
    uint32_t seed = fbits(P.reach[3]);
    uint32_t rays = P.counts[2];
    lds_u32 *slot_lds = group_slot(A.depth, (uint32_t)L);
    C cnt{0, 0, 0, 0, 0, 0};
    uint32_t range = blockIdx.x % XCDS;
    for (;;) {
        uint32_t unit = take_unit(A, n_units, range);
        if (unit == 0xffffffffu) break;
        uint32_t i = unit * per_wave + grp;
        bool valid = i < A.n_items;
        {
            // the item's state goes to the group's LDS slot and is read back at each ray start,
            // so it holds no vector registers across the traversal
            ItemIn tx = load_item(A.items, valid ? i : 0u);
            frame_t f = frame(tx.n);
            uint32_t key = (A.key_base + i) * 7919u;
            float s1 = rnd(key, 11u, seed), s2 = rnd(key, 13u, seed);
            if (lig == 0u) {
                put_f3(slot_lds, tx.pos);
                put_f3(slot_lds + 3, tx.dx);
                put_f3(slot_lds + 6, tx.dy);
                put_f3(slot_lds + 9, tx.n);
                put_f3(slot_lds + 12, f.t);
                put_f3(slot_lds + 15, f.b);
                slot_lds[18] = fbits(s1);
                slot_lds[19] = fbits(s2);
                slot_lds[20] = key;
            }
        }
        KSum sr{0, 0}, sg{0, 0}, sb{0, 0};
        float back = 0.0f;
        decltype(cnt.inner) steps0 = 0;
        uint32_t ray_of_slot = 0, cost_slot = 0;
        auto start = [&](uint32_t slot, Ray &R) {
            // read the slot afresh (the compiler would otherwise keep the stored values live)
            asm volatile("" ::: "memory");
            f3 pos = get_f3(slot_lds), dx = get_f3(slot_lds + 3), dy = get_f3(slot_lds + 6),
               n = get_f3(slot_lds + 9);
            frame_t f;
            f.t = get_f3(slot_lds + 12);
            f.b = get_f3(slot_lds + 15);
            f.n = n;
            float s1 = bitsf(slot_lds[18]), s2 = bitsf(slot_lds[19]);
            uint32_t key = slot_lds[20];
            uint32_t s, r;
            if (A.perm) {
                s = (uint32_t)A.perm[(size_t)i * rays + slot];
                r = az_slot(s, rays, s2);
            } else if (A.el_order) {
                r = el_az_slot(slot, rays, s1, s2);
                s = az_slot_of(r, rays, s2);
            } else {
                s = slot;
                r = az_slot(s, rays, s2);
            }
            ray_of_slot = r;
            cost_slot = s;
            if (STATS) steps0 = cnt.inner + cnt.tris;
            float jx = rnd(key + r, 19u, seed) - 0.5f;
            float jy = rnd(key + r, 23u, seed) - 0.5f;
            f3 o = add(add(add(pos, scl(dx, jx)), scl(dy, jy)), scl(n, P.reach[2]));
            f3 d = cosine(fract(((float)r + 0.5f) / (float)rays + s1), fract(radical(r) + s2), f);
            ray_start(R, A.S.root, o, d, P.reach[1], false, key + r + 4000u);
        };
        auto shade = [&](uint32_t slot, const Ray &R) {
            if (A.rec)
                A.rec[(size_t)ray_of_slot * A.n_items + i] =
                    gather_record(A.S, P, R.o, R.d, R.tri, R.back, R.t);
            if (STATS && A.cost) {
                uint32_t steps = (uint32_t)(cnt.inner + cnt.tris - steps0);
                A.cost[(size_t)i * rays + cost_slot] = (uint8_t)(steps < 255u ? steps : 255u);
            }
            if (R.tri == NONE) {
                f3 l = sky(A.S.sh, P, R.d);
                sr.add(l.x);
                sg.add(l.y);
                sb.add(l.z);
            } else if (R.back) {
                back += 1.0f;
            } else {
                if (STATS) {
                    cnt.mats++;
                    cnt.gathers++;
                }
                f3 l = bounced(A.S, P, R.o, R.d, Hit{R.t, R.tri, R.back});
                sr.add(l.x);
                sg.add(l.y);
                sb.add(l.z);
            }
        };
        run_slots<LAYOUT, STATS, MODE == 1 ? COUNTING_VISITS : LOCKED_VISITS>(
            A, st, lig, valid ? rays : 0u, (uint32_t)L, cnt, start, shade);
        group_ksum(sr, L);
        group_ksum(sg, L);
        group_ksum(sb, L);
        back = group_sum(back, L);
        if (valid && lig == 0u)
            A.out[i] = make_float4(sr.value() * PI / (float)rays, sg.value() * PI / (float)rays,
                                   sb.value() * PI / (float)rays, back / (float)rays);
    }
    if constexpr (MODE == 2) flush_counts(cnt, A.stats);
}

// trace.wgsl:272-324 over rays + g^2 sun slots; WAVES as for the gather
template <int LAYOUT, bool STATS, int WAVES>
__global__ __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1))) void k_volume_tuned(
    TunedArgs A) {
    StackView st = make_stack(A.spill, A.depth);
    const lb_params &P = A.P;
    int L = (int)A.lanes_per_item;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t per_wave = 64u / (uint32_t)L;
    uint32_t grp = lane / (uint32_t)L, lig = lane % (uint32_t)L;
    uint32_t n_units = (A.n_items + per_wave - 1u) / per_wave;
    uint32_t seed = fbits(P.reach[3]);
    uint32_t rays = P.counts[2];
    uint32_t g = P.counts[0], gg = g * g;
    f3 sund = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    frame_t sf = frame(sund);
    Counts cnt{0, 0, 0, 0, 0, 0};
    uint32_t range = blockIdx.x % XCDS;
    for (;;) {
        uint32_t unit = take_unit(A, n_units, range);
        if (unit == 0xffffffffu) break;
        uint32_t i = unit * per_wave + grp;
        bool valid = i < A.n_items;
        ItemIn pr = load_item(A.items, valid ? i : 0u);
        f3 o = pr.pos;
        uint32_t key = pr.w * 7919u + 77u;
        float s1 = rnd(key, 29u, seed), s2 = rnd(key, 31u, seed);
        KSum c[12];
#pragma unroll
        for (int k = 0; k < 12; k++) c[k] = KSum{0.0f, 0.0f};
        float back = 0.0f, sun = 0.0f;
        auto start = [&](uint32_t slot, Ray &R) {
            if (slot < rays) {
                uint32_t r = az_slot(slot, rays, s2);
                float u = fract(((float)r + 0.5f) / (float)rays + s1);
                float v = fract(radical(r) + s2);
                float z = 1.0f - 2.0f * u;
                float rr = sqrtf(vmax(0.0f, 1.0f - z * z));
                float a = v * 2.0f * PI;
                float sa, ca;
                sincos_lb(a, &sa, &ca);
                ray_start(R, A.S.root, o, mk(rr * ca, z, rr * sa), P.reach[1], false, key + r);
            } else {
                uint32_t s = slot - rays;
                float a = (float)s * 2.399963f;
                float r = P.sun_dir[3] * sqrtf(((float)s + 0.5f) / (float)gg);
                float sa, ca;
                sincos_lb(a, &sa, &ca);
                f3 d = normalize3(add(add(sund, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
                ray_start(R, A.S.root, o, d, FAR, true, key + s + 9000u);
            }
        };
        auto shade = [&](uint32_t slot, const Ray &R) {
            if (slot < rays) {
                f3 d = R.d;
                f3 l = sky(A.S.sh, P, d);
                if (R.tri != NONE) {
                    if (R.back) {
                        back += 1.0f;
                        l = mk(0.0f, 0.0f, 0.0f);
                    } else {
                        if (STATS) {
                            cnt.mats++;
                            cnt.gathers++;
                        }
                        l = bounced(A.S, P, R.o, d, Hit{R.t, R.tri, R.back});
                    }
                }
                f3 l0 = scl(l, 0.282095f);
                f3 lx = scl(scl(l, 0.488603f), d.x);
                f3 ly = scl(scl(l, 0.488603f), d.y);
                f3 lz = scl(scl(l, 0.488603f), d.z);
                c[0].add(l0.x); c[1].add(l0.y); c[2].add(l0.z);
                c[3].add(lx.x); c[4].add(lx.y); c[5].add(lx.z);
                c[6].add(ly.x); c[7].add(ly.y); c[8].add(ly.z);
                c[9].add(lz.x); c[10].add(lz.y); c[11].add(lz.z);
            } else if (R.tri == NONE) {
                sun += 1.0f;
            }
        };
        run_slots<LAYOUT, STATS, LOCKED_VISITS>(A, st, lig, valid ? rays + gg : 0u, (uint32_t)L,
                                                cnt, start, shade);
#pragma unroll
        for (int k = 0; k < 12; k++) group_ksum(c[k], L);
        back = group_sum(back, L);
        sun = group_sum(sun, L);
        if (valid && lig == 0u) {
            float w = 4.0f * PI / (float)rays;
            A.out[i * 4u] = make_float4(c[0].value() * w, c[1].value() * w, c[2].value() * w,
                                        sun / (float)gg);
            A.out[i * 4u + 1u] = make_float4(c[3].value() * w, c[4].value() * w,
                                             c[5].value() * w, back / (float)rays);
            A.out[i * 4u + 2u] =
                make_float4(c[6].value() * w, c[7].value() * w, c[8].value() * w, 0.0f);
            A.out[i * 4u + 3u] =
                make_float4(c[9].value() * w, c[10].value() * w, c[11].value() * w, 0.0f);
        }
    }
    if (STATS) flush_counts(cnt, A.stats);
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }
// the chiplets' counters (XCDS of them, QUEUE_STRIDE words apart: 1 KiB)
__global__ void k_zero_queue(uint32_t *q) {
    if (threadIdx.x < XCDS) q[threadIdx.x * QUEUE_STRIDE] = 0u;
}

using KernelFn = void (*)(TunedArgs);

// The kernel of a pass, layout, counters and residency target (waves a SIMD: 0, 6 or 8).
// MODE as k_gather_tuned's: the other passes have no counting launch (1 is taken as 0)
template <int LAYOUT, int MODE, int WAVES>
static KernelFn pick_waves(int pass) {
    if (pass == LB_PASS_DIRECT) return k_direct_tuned<LAYOUT, MODE == 2, WAVES>;
    if (pass == LB_PASS_GATHER) return k_gather_tuned<LAYOUT, MODE, WAVES>;
    return k_volume_tuned<LAYOUT, MODE == 2, WAVES>;
}
template <int LAYOUT, int MODE>
static KernelFn pick_mode(int pass, int waves) {
    if (waves == 6) return pick_waves<LAYOUT, MODE, 6>(pass);
    if (waves == 8) return pick_waves<LAYOUT, MODE, 8>(pass);
    return pick_waves<LAYOUT, MODE, 0>(pass);
}
template <int LAYOUT>
static KernelFn pick_layout(int pass, int mode, int waves) {
    if (mode == 2) return pick_mode<LAYOUT, 2>(pass, waves);
    if (mode == 1) return pick_mode<LAYOUT, 1>(pass, waves);
    return pick_mode<LAYOUT, 0>(pass, waves);
}
static KernelFn pick_kernel(int pass, int layout, int mode, int waves) {
    return layout == LB_NODES_WIDE ? pick_layout<0>(pass, mode, waves)
                                   : pick_layout<1>(pass, mode, waves);
}

// Lanes sharing one item: the rays an item casts over the rays a lane should take, as a power
// of two in 8..64.
uint32_t lanes_per_item(int pass, const lb_params &P, int rays_per_lane) {
    uint32_t gg = P.counts[0] * P.counts[0];
    uint32_t rays = pass == LB_PASS_DIRECT   ? 2u * gg + P.counts[1]
                    : pass == LB_PASS_GATHER ? P.counts[2]
                                             : P.counts[2] + gg;
    uint32_t want = (rays + (uint32_t)rays_per_lane - 1u) / (uint32_t)rays_per_lane;
    uint32_t L = 8;
    while (L < want && L < 64) L <<= 1;
    return L;
}

struct TunedGrid {
    KernelFn fn;
    uint32_t block, grid, lanes;
    size_t lds;
};

// mode: 0 plain, 1 the gather's counting launch, 2 with the traversal counters. With `spill`
// the launch's spill area in bytes is returned and nothing is reserved.
static hipError_t tuned_grid(lb_gpu *gpu, int pass, int mode, const lb_params &P, uint32_t n,
                             TunedGrid *out, size_t *spill = nullptr) {
    const Tuning &T = gpu->tuning[pass];
    out->fn = pick_kernel(pass, T.nodes, mode, T.waves);
    out->block = (uint32_t)T.workgroup;
    out->lds = stack_lds_bytes(T.lds_stack, (int)out->block);
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(out->fn), (int)out->block, out->lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    out->lanes = lanes_per_item(pass, P, T.rays_per_lane);
    out->lds += (size_t)(out->block / out->lanes) * SLOT_WORDS * 4u;
    uint32_t per_block = (out->block / 64u) * (64u / out->lanes);
    uint32_t need = (n + per_block - 1u) / per_block;
    uint32_t grid = (uint32_t)gpu->cus * (uint32_t)per_cu;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    out->grid = grid;
    size_t spill_bytes = (size_t)grid * out->block * (size_t)(STACK - T.lds_stack) * 4u;
    if (spill_bytes < 16) spill_bytes = 16;
    if (spill) {
        *spill = spill_bytes;
        return hipSuccess;
    }
    if (gpu->spill.reserve(spill_bytes)) return hipErrorOutOfMemory;
    return hipSuccess;
}

hipError_t tuned_spill_bytes(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n,
                             size_t *bytes) {
    TunedGrid tg;
    return tuned_grid(gpu, pass, 0, P, n, &tg, bytes);
}

// Each item's azimuth slots ordered by (recorded cost, slot): a stable counting sort on the 8-bit
// cost, one wave an item. Slot s of an ordered gather then takes the s-th cheapest ray, so the
// lanes of one round trace rays of like length.
//
// Lane l holds slots j * 64 + l (j < 4). Chunk by chunk in slot order, a lane's place among the
// slots of its cost is the count of that cost in the earlier chunks (the wave's 256 bins in LDS)
// plus the lanes below it in this chunk with the same cost (eight ballots, one a cost bit); the
// chunk's highest lane of each cost then adds the chunk's count to the bin. An exclusive scan of
// the bins gives each cost's first place. A wave's LDS requests complete in the order it issues
// them and every wave keeps to its own bins, so no barrier is needed, only the compiler's
// (`order_lds`) between the phases that read what other lanes wrote. With 64 rays or fewer an
// item is one chunk and the ballots alone give each slot its place.
constexpr uint32_t RANK_WAVES = 4;
__device__ __forceinline__ void order_lds() { asm volatile("" ::: "memory"); }
__global__ __launch_bounds__(RANK_WAVES * 64) void k_rank_slots(const uint8_t *__restrict__ cost,
                                                                uint32_t n_items, uint32_t rays,
                                                                uint8_t *__restrict__ perm) {
    __shared__ uint32_t bin[RANK_WAVES][256];
    __shared__ uint8_t ord[RANK_WAVES][256];
    uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    uint64_t below = (1ull << l) - 1ull;
    for (uint32_t item = blockIdx.x * RANK_WAVES + w; item < n_items;
         item += gridDim.x * RANK_WAVES) {
        size_t base = (size_t)item * rays;
        if (rays <= 64u) {
            // one chunk: a lane's place is the count of the slots with a lower cost plus those
            // below it with its own, both from the eight ballots, the cost's bits taken from the
            // highest (no bins, no scan); slot s = lane l
            bool v = l < rays;
            uint32_t x = v ? (uint32_t)cost[base + l] : 0u;
            uint64_t lt = 0ull, eq = __ballot(v);
            for (int b = 7; b >= 0; b--) {
                uint64_t m = __ballot((x >> b) & 1u);
                if ((x >> b) & 1u) {
                    lt |= eq & ~m;
                    eq &= m;
                } else {
                    eq &= ~m;
                }
            }
            if (v) perm[base + __popcll(lt) + __popcll(eq & below)] = (uint8_t)l;
            continue;
        }
        uint32_t c[4], place[4];
        for (uint32_t j = 0; j < 4u; j++) {
            uint32_t s = j * 64u + l;
            c[j] = s < rays ? (uint32_t)cost[base + s] : 0u;
        }
        order_lds();
        for (uint32_t k = l; k < 256u; k += 64u) bin[w][k] = 0u;
        order_lds();
        for (uint32_t j = 0; j < 4u; j++) {
            bool v = j * 64u + l < rays;
            uint32_t x = c[j];
            uint64_t same = __ballot(v);
            for (uint32_t b = 0; b < 8u; b++) {
                uint64_t m = __ballot((x >> b) & 1u);
                same &= ((x >> b) & 1u) ? m : ~m;
            }
            uint32_t prior = v ? bin[w][x] : 0u;
            place[j] = prior + (uint32_t)__popcll(same & below);
            if (v && (same >> l) == 1ull) bin[w][x] = prior + (uint32_t)__popcll(same);
            order_lds();
        }
        // each lane's four bins, then the wave's inclusive scan of their sums
        uint32_t h0 = bin[w][4u * l], h1 = bin[w][4u * l + 1u], h2 = bin[w][4u * l + 2u],
                 h3 = bin[w][4u * l + 3u];
        uint32_t sum = h0 + h1 + h2 + h3, inc = sum;
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            uint32_t up = (uint32_t)__shfl_up((int)inc, d);
            if (l >= d) inc += up;
        }
        uint32_t first = inc - sum;
        bin[w][4u * l] = first;
        bin[w][4u * l + 1u] = first + h0;
        bin[w][4u * l + 2u] = first + h0 + h1;
        bin[w][4u * l + 3u] = first + h0 + h1 + h2;
        order_lds();
        for (uint32_t j = 0; j < 4u; j++) {
            uint32_t s = j * 64u + l;
            if (s < rays) ord[w][bin[w][c[j]] + place[j]] = (uint8_t)s;
        }
        order_lds();
        for (uint32_t j = 0; j < 4u; j++) {
            uint32_t s = j * 64u + l;
            if (s < rays) perm[base + s] = ord[w][s];
        }
        order_lds();
    }
}

hipError_t rank_slots(const uint8_t *cost, uint32_t n_items, uint32_t rays, uint8_t *perm,
                      hipStream_t stream) {
    if (!n_items) return hipSuccess;
    if (rays > 256u) return hipErrorInvalidValue;
    uint32_t blocks = (n_items + RANK_WAVES - 1u) / RANK_WAVES;
    uint32_t grid = blocks < 65536u ? blocks : 65536u;
    hipLaunchKernelGGL(k_rank_slots, dim3(grid), dim3(RANK_WAVES * 64), 0, stream, cost, n_items,
                       rays, perm);
    return hipGetLastError();
}

hipError_t tuned_prepare(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n) {
    if (gpu->tuning[pass].schedule == 1) return stream_prepare(gpu, pass, P, n);
    TunedGrid tg;
    for (int mode = 0; mode < 3; mode++) {
        hipError_t e = tuned_grid(gpu, pass, mode, P, n, &tg);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_tuned(lb_gpu *gpu, int pass, bool stats, const SceneView &S,
                        const lb_params &P, const lb_texel *items, uint32_t n, uint32_t key_base,
                        float4 *out, const GatherOrder *order, const FixedScratch *fixed) {
    if (gpu->tuning[pass].schedule == 1) {
        // the pooled schedule keeps every ray's hit in memory sized by the launch
        if (fixed) return hipErrorNotSupported;
        return launch_stream(gpu, pass, stats, S, P, items, n, key_base, out);
    }
    // the gather's counting launch records each ray's steps and counts nothing else
    bool ordered = order && pass == LB_PASS_GATHER && P.counts[2] <= 256u;
    uint8_t *cost = ordered ? order->cost : nullptr;
    int mode = stats ? 2 : (cost ? 1 : 0);
    TunedGrid tg;
    size_t need = 0;
    hipError_t e = tuned_grid(gpu, pass, mode, P, n, &tg, fixed ? &need : nullptr);
    if (e != hipSuccess) return e;
    if (fixed && need > fixed->spill_bytes) return hipErrorOutOfMemory;
    uint32_t *queue = fixed ? fixed->queue : gpu->queue;
    TunedArgs A;
    A.S = S;
    A.P = P;
    A.items = items;
    A.n_items = n;
    A.key_base = key_base;
    A.out = out;
    A.stats = gpu->counters;
    A.queue = queue;
    A.spill = fixed ? fixed->spill : static_cast<uint32_t *>(gpu->spill.ptr);
    A.lanes_per_item = tg.lanes;
    A.depth = (uint32_t)gpu->tuning[pass].lds_stack;
    A.leaf_batch = gpu->tuning[pass].leaf_batch;
    A.cost = cost;
    A.perm = ordered ? order->perm : nullptr;
    A.rec = order && pass == LB_PASS_GATHER ? order->rec : nullptr;
    A.xcd_split = (uint32_t)gpu->tuning[pass].xcd_split;
    uint32_t rays = P.counts[2];
    A.el_order = pass == LB_PASS_GATHER && !A.perm && gpu->tuning[pass].gather_order >= 1 &&
                 rays >= 64u && rays <= 65536u && (rays & (rays - 1u)) == 0u;
    hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(64), 0, gpu->stream, queue);
    hipLaunchKernelGGL(tg.fn, dim3(tg.grid), dim3(tg.block), tg.lds, gpu->stream, A);
    return hipGetLastError();
}

}  // namespace lbk
