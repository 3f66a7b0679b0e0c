// The traversal of trace.wgsl:83-130 over the TUNED node records, as steps a kernel can schedule:
// a ray's state lives in `Ray`, `inner_visit` and `leaf_tests` take one node visit, `pop` takes
// the next node off the ray's stack. Any sequence of these steps that follows the reference's
// loop visits the same nodes in the same order and accepts the same hit.
//
// Node references: bit 31 marks a leaf; bits 26-30 hold its triangle count (31: read the
// (first, count) pair from `leaves` at the node's index), bits 0-25 its first triangle or, for an
// inner node, its index.
#pragma once

#include "lb_device.h"
#include "lb_kernels.h"

namespace lbk {

constexpr uint32_t LEAF_BIT = 0x80000000u;

// The two halves of a lane's stack carry their address spaces in their types, so an entry moves
// with ds_read / ds_write (LDS) or global_load / global_store, never with a flat access that
// waits on both counters.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(1))) uint32_t glb_u32;

// A lane's entries are consecutive words (an entry's address is a shift and an add, no multiply)
// and lanes sit an odd number of words apart, so the 32 lanes of an LDS access hit 32 distinct
// banks.
struct StackView {
    lds_u32 *lds;    // this lane's entry 0; entry k at lds[k]
    glb_u32 *spill;  // this lane's entry `depth`; entry k at spill[k - depth]
    uint32_t depth;
    // The LDS access is taken whatever k (an entry past the LDS part goes to word `depth`, the
    // lane's spare: lane_stride(depth) > depth), so only the rare spill takes a branch.
    __device__ __forceinline__ void put(uint32_t k, uint32_t v) const {
        lds[k < depth ? k : depth] = v;
        if (k >= depth) spill[k - depth] = v;
    }
    __device__ __forceinline__ uint32_t get(uint32_t k) const {
        uint32_t v = lds[k < depth ? k : depth];
        if (k >= depth) v = spill[k - depth];
        return v;
    }
};

// A workgroup's lanes: LDS stacks of `depth` entries, lane_stride(depth) words apart (the
// kernel's dynamic shared memory), and a spill area of STACK - depth entries per lane of the
// grid.
__device__ __forceinline__ StackView make_stack(uint32_t *spill, uint32_t depth) {
    extern __shared__ uint32_t lds_stack[];
    StackView st;
    st.lds = (lds_u32 *)lds_stack + threadIdx.x * lane_stride(depth);
    uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    st.spill = (glb_u32 *)spill + (size_t)gtid * (uint32_t)(STACK - (int)depth);
    st.depth = depth;
    return st;
}

// The children of inner node n from its WIDE record: boxes, and each child's reference.
__device__ __forceinline__ void children(const SceneView &S, uint32_t n, f3 &amn, f3 &amx,
                                         f3 &bmn, f3 &bmx, uint32_t &ra, uint32_t &rb) {
    const float4 *w = S.wide + n * 4u;
    float4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
    // the references' load goes out with the boxes' (one round trip a visit): without this the
    // compiler defers it past the box tests
    asm volatile("" ::"v"(w3.x), "v"(w3.y));
    amn = mk(w0.x, w0.y, w0.z);
    amx = mk(w0.w, w1.x, w1.y);
    bmn = mk(w1.z, w1.w, w2.x);
    bmx = mk(w2.y, w2.z, w2.w);
    ra = fbits(w3.x);
    rb = fbits(w3.y);
}

// A ray in flight: trace.wgsl:83-130's loop state.
struct Ray {
    f3 o, d, inv;
    float t;       // the nearest accepted hit so far (tmax until one is found)
    uint32_t tri;  // NONE until a hit
    uint32_t n;    // the node to visit next
    uint32_t sp;   // stack entries
    uint32_t key;  // stops()'s ray key
    uint32_t qsel[6];  // quant_slabs' byte selectors: near x y, near z far x, far y z; b's
    bool back, any;
};

__device__ __forceinline__ void ray_start(Ray &R, uint32_t root, f3 o, f3 d, float tmax, bool any,
                                          uint32_t key) {
    R.o = o;
    R.d = d;
    R.inv = inverse_dir(d);
    R.t = tmax;
    R.tri = NONE;
    R.back = false;
    R.n = root;
    R.sp = 0;
    R.key = key;
    R.any = any;
    // on an axis where inv >= 0 the planes' t rise with the byte, so the lower corner's byte
    // gives the nearer plane; else the upper's. Byte indices of child a's corners (quant_slabs),
    // each in the low byte of a 16-bit half whose high byte selector 12 clears
    bool px = R.inv.x >= 0.0f, py = R.inv.y >= 0.0f, pz = R.inv.z >= 0.0f;
    R.qsel[0] = (px ? 0u : 3u) | (py ? 1u : 4u) << 16 | 0x0c000c00u;
    R.qsel[1] = (pz ? 2u : 5u) | (px ? 3u : 0u) << 16 | 0x0c000c00u;
    R.qsel[2] = (py ? 4u : 1u) | (pz ? 5u : 2u) << 16 | 0x0c000c00u;
    R.qsel[3] = R.qsel[0] + 0x00020002u;
    R.qsel[4] = R.qsel[1] + 0x00020002u;
    R.qsel[5] = R.qsel[2] + 0x00020002u;
}

// The triangles of leaf R.n (trace.wgsl:91-104); true when an any-hit ray accepted one.
template <bool STATS, typename C>
__device__ __forceinline__ bool leaf_tests(const SceneView &S, const lb_params &P, Ray &R,
                                           C &cnt) {
    uint32_t first = R.n & 0x03ffffffu, count = (R.n >> 26) & 31u;
    if (count == 31u) {
        uint2 li = S.leaves[first];
        first = li.x;
        count = li.y;
    }
    if (STATS) cnt.leaves++;
    uint32_t end = first + count;
    for (uint32_t i = first; i < end; i++) {
        // one address, the three records at immediate offsets
        const float4 *tp = S.tris_t + (size_t)i * 3u;
        float4 v0 = tp[0], e1 = tp[1], e2 = tp[2];
        // the three loads are issued together and waited for once: without this the compiler
        // sinks v0's load past the determinant test, a second memory round trip
        asm volatile("" ::"v"(v0.x), "v"(v0.y), "v"(v0.z), "v"(e2.w));
        if (STATS) cnt.tris++;
        float det;
        float t = mt_all(R.o, R.d, v0, e1, e2, &det);
        if (t > 1e-4f && t < R.t) {
            if (STATS) cnt.mats++;
            // stops() (trace.wgsl:69-72) with the opacity carried in e2.w
            float op = e2.w;
            if (op >= 1.0f || rnd(i * 9781u, R.key, fbits(P.reach[3])) < op) {
                R.t = t;
                R.tri = i;
                R.back = det < 0.0f;
                // an any-hit ray ends at its first accepted hit: the loop stops there (no
                // return from inside it, which costs the loop's exec-mask bookkeeping)
                if (R.any) end = i;
            }
        }
    }
    return R.any && R.tri != NONE;
}

// Inner node R.n (trace.wgsl:105-123): true when the ray moved on to its near child (the far one
// pushed if it was hit too), false when it missed both.
// A QUANT record's two slab tests without decoding the corners. Per axis the record's word holds
// the origin (a float at or below both children's boxes, its low 9 bits zero) with a biased
// exponent in bits 0-7, so the word shifted left by 23 reads as 2^24 times the step 2^e; the
// plane of byte q lies at t = q * (2^e * inv) + (origin - o) * inv, one fused multiply-add a
// plane after six products a node. The planes' t round differently from (corner - o) * inv on the
// decoded corners, by a few ulps; the encoder keeps every corner outside its FP32 box in exact
// arithmetic as well as after rounding, so the boxes stay conservative.
// On each axis the nearer plane is known from the sign of inv (t is monotone in q, and so after
// rounding), so the ray's selectors (Ray::qsel) gather each box's near and far bytes, two to a
// word, each in the low byte of a 16-bit half: read as an FP16 value that half is q * 2^-24
// (a subnormal, exact), and the mixed-precision fused multiply-add q * 2^-24 * (2^(e+24) * inv)
// + c reads it straight from the word and gives the same rounding as q * (2^e * inv) + c, the
// scaling by 2^24 being exact. The slab test's per-axis min and max are exact selections.
// Record: q0 = (origin words x, y, z, reference a), q1 = 12 corner bytes (a min xyz, a max xyz,
// b min xyz, b max xyz), reference b.
__device__ __forceinline__ float half_lo(uint32_t w) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t w) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}

__device__ __forceinline__ void quant_slabs(const SceneView &S, const Ray &R, float &la, bool &ha,
                                            float &lb, bool &hb, uint32_t &ra, uint32_t &rb) {
    const uint4 *q = S.quant + R.n * 2u;
    uint4 q0 = q[0], q1 = q[1];
    asm volatile("" ::"v"(q1.w));
    float kx = bitsf(q0.x << 23) * R.inv.x, ky = bitsf(q0.y << 23) * R.inv.y,
          kz = bitsf(q0.z << 23) * R.inv.z;
    float cx = (bitsf(q0.x & 0xffffff00u) - R.o.x) * R.inv.x,
          cy = (bitsf(q0.y & 0xffffff00u) - R.o.y) * R.inv.y,
          cz = (bitsf(q0.z & 0xffffff00u) - R.o.z) * R.inv.z;
    // as slab_hit: the entry max(0, nearest planes) in one chain, and lo < t as l0 < t (t > 0).
    // Child a's bytes are selector values 0-5 of q1.y:q1.x; child b's, 2-7 of q1.z:q1.y: its
    // selectors (qsel[3..5]) are a's with 2 added to both data bytes.
    auto slab = [&](uint32_t hi_w, uint32_t lo_w, int b, bool *hit) {
        uint32_t w1 = __builtin_amdgcn_perm(hi_w, lo_w, R.qsel[3 * b]);
        uint32_t w2 = __builtin_amdgcn_perm(hi_w, lo_w, R.qsel[3 * b + 1]);
        uint32_t w3 = __builtin_amdgcn_perm(hi_w, lo_w, R.qsel[3 * b + 2]);
        float l0 = vmax(vmax(vmax(0.0f, __builtin_fmaf(half_lo(w1), kx, cx)),
                             __builtin_fmaf(half_hi(w1), ky, cy)),
                        __builtin_fmaf(half_lo(w2), kz, cz));
        float hi = vmin(vmin(__builtin_fmaf(half_hi(w2), kx, cx),
                             __builtin_fmaf(half_lo(w3), ky, cy)),
                        __builtin_fmaf(half_hi(w3), kz, cz));
        *hit = hi >= l0 && l0 < R.t;
        return l0;
    };
    la = slab(q1.y, q1.x, 0, &ha);
    lb = slab(q1.z, q1.y, 1, &hb);
    ra = q0.w;
    rb = q1.w;
}

template <int LAYOUT, bool STATS, typename C>
__device__ __forceinline__ bool inner_visit(const SceneView &S, Ray &R, const StackView &st,
                                            C &cnt) {
    uint32_t ra, rb;
    if (STATS) cnt.inner++;
    // box() (trace.wgsl:74-80) gives the entry l0 when the box is hit and FAR = 1e30 otherwise,
    // and an entry is always below tmax <= FAR; so with ha, hb the hit tests, the reference's
    // ta <= tb is ha && (!hb || l0a <= l0b), min(ta, tb) < FAR is ha || hb and
    // max(ta, tb) < FAR is ha && hb, taken here on the tests themselves
    bool ha, hb;
    float la, lb;
    if constexpr (LAYOUT == LB_NODES_QUANT) {
        quant_slabs(S, R, la, ha, lb, hb, ra, rb);
    } else {
        f3 amn, amx, bmn, bmx;
        children(S, R.n, amn, amx, bmn, bmx, ra, rb);
        la = slab_hit(R.o, R.inv, amn, amx, R.t, &ha);
        lb = slab_hit(R.o, R.inv, bmn, bmx, R.t, &hb);
    }
    if (!(ha || hb)) return false;
    bool an = ha && (!hb || la <= lb);
    uint32_t near = an ? ra : rb, far = an ? rb : ra;
    if (ha && hb && R.sp < (uint32_t)STACK) {
        st.put(R.sp, far);
        R.sp++;
    }
    R.n = near;
    return true;
}

// trace.wgsl:124-127: the next node off the stack; true when it was empty (the ray ends).
__device__ __forceinline__ bool pop(Ray &R, const StackView &st) {
    if (R.sp == 0u) return true;
    R.sp--;
    R.n = st.get(R.sp);
    return false;
}

// One node visit in the reference's order; true when the ray is finished.
template <int LAYOUT, bool STATS, typename C>
__device__ __forceinline__ bool ray_step(const SceneView &S, const lb_params &P, Ray &R,
                                         const StackView &st, C &cnt) {
    if (R.n & LEAF_BIT) {
        if (leaf_tests<STATS>(S, P, R, cnt)) return true;
    } else if (inner_visit<LAYOUT, STATS>(S, R, st, cnt)) {
        return false;
    }
    return pop(R, st);
}

// ---- where a short ray's traversal can start ---------------------------------------------------
//
// Rays that stay within a known box (the occlusion and inside-test rays of a texel: their origins
// within the texel's footprint, lengths within their reach) meet every node's boxes the same way
// near the root: while one child's box holds the whole region with margin m and the other's lies
// apart from it by m, each ray's slab test hits the first (its planes' t lie behind the origin or
// past the reach) and misses the second (past the reach, or behind), so the reference's loop goes
// to the first child and pushes nothing. Starting such a ray at the deepest node reached that way,
// with an empty stack, it then takes exactly the steps it would take from the root. The margin
// covers the slab arithmetic's rounding: the planes' t err by less than 2^-22 (|q step| +
// |origin - o|) |inv|, m |inv| with m of a centimetre and more for large coordinates.
__device__ __forceinline__ bool region_inside(f3 lo, f3 hi, f3 mn, f3 mx, float m) {
    return lo.x >= mn.x + m && lo.y >= mn.y + m && lo.z >= mn.z + m && hi.x <= mx.x - m &&
           hi.y <= mx.y - m && hi.z <= mx.z - m;
}
__device__ __forceinline__ bool region_apart(f3 lo, f3 hi, f3 mn, f3 mx, float m) {
    return hi.x < mn.x - m || hi.y < mn.y - m || hi.z < mn.z - m || lo.x > mx.x + m ||
           lo.y > mx.y + m || lo.z > mx.z + m;
}

// The children's boxes of inner node n as the TUNED form's records hold them: QUANT corners
// origin + q * 2^e in FP32 (the product exact; within m of what the slab planes stand for),
// WIDE boxes as stored.
template <int LAYOUT>
__device__ __forceinline__ void record_boxes(const SceneView &S, uint32_t n, f3 &amn, f3 &amx,
                                             f3 &bmn, f3 &bmx, uint32_t &ra, uint32_t &rb) {
    if constexpr (LAYOUT == LB_NODES_QUANT) {
        const uint4 *q = S.quant + n * 2u;
        uint4 q0 = q[0], q1 = q[1];
        float ox = bitsf(q0.x & 0xffffff00u), oy = bitsf(q0.y & 0xffffff00u),
              oz = bitsf(q0.z & 0xffffff00u);
        float sx = bitsf(q0.x << 23) * 0x1p-24f, sy = bitsf(q0.y << 23) * 0x1p-24f,
              sz = bitsf(q0.z << 23) * 0x1p-24f;
        auto b = [](uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); };
        amn = mk(ox + b(q1.x, 0) * sx, oy + b(q1.x, 1) * sy, oz + b(q1.x, 2) * sz);
        amx = mk(ox + b(q1.x, 3) * sx, oy + b(q1.y, 0) * sy, oz + b(q1.y, 1) * sz);
        bmn = mk(ox + b(q1.y, 2) * sx, oy + b(q1.y, 3) * sy, oz + b(q1.z, 0) * sz);
        bmx = mk(ox + b(q1.z, 1) * sx, oy + b(q1.z, 2) * sy, oz + b(q1.z, 3) * sz);
        ra = q0.w;
        rb = q1.w;
    } else {
        children(S, n, amn, amx, bmn, bmx, ra, rb);
    }
}

// The node (a child reference) rays within [lo, hi] may start at.
template <int LAYOUT>
__device__ __forceinline__ uint32_t entry_node(const SceneView &S, f3 lo, f3 hi) {
    float big = vmax(vmax(vmax(fabsf(lo.x), fabsf(lo.y)), vmax(fabsf(lo.z), fabsf(hi.x))),
                     vmax(fabsf(hi.y), fabsf(hi.z)));
    float m = 0.01f + 2e-6f * big;
    uint32_t n = S.root;
    while (!(n & LEAF_BIT)) {
        f3 amn, amx, bmn, bmx;
        uint32_t ra, rb;
        record_boxes<LAYOUT>(S, n, amn, amx, bmn, bmx, ra, rb);
        if (region_inside(lo, hi, amn, amx, m) && region_apart(lo, hi, bmn, bmx, m))
            n = ra;
        else if (region_inside(lo, hi, bmn, bmx, m) && region_apart(lo, hi, amn, amx, m))
            n = rb;
        else
            break;
    }
    return n;
}

}  // namespace lbk
