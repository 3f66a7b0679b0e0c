// The PLAIN form of the three ray passes: the reference's own mapping (one item a lane, a serial
// loop over its rays, the 48-entry traversal stack per lane, FP32 nodes read as the reference
// reads them: the node, then both children). It is the baseline the TUNED form is measured
// against and it reproduces crates/lightbaker/src/gpu/trace.wgsl bit for bit.
#include "lb_device.h"
#include "lb_kernels.h"
#include "lb_trace_plain.h"

namespace lbk {

// trace.wgsl:159-168
template <bool STATS, bool WT>
__device__ bool inside_plain(const SceneView &S, const lb_params &P, f3 o, f3 n, uint32_t key,
                             Counts &cnt) {
    frame_t f = frame(n);
    float back = 0.0f;
    for (uint32_t r = 0; r < 16u; r++) {
        f3 d = cosine(((float)r + 0.5f) / 16.0f, radical(r), f);
        Hit h = trace_plain<false, STATS, WT>(S, P, o, d, P.ground[3], key + r + 7000u, cnt);
        if (h.tri != NONE && h.back) back += 1.0f;
    }
    return back / 16.0f > bitsf(P.atlas[3]);
}

// trace.wgsl:170-243
template <bool STATS, bool WT>
__global__ __launch_bounds__(64) void k_direct_plain(SceneView S, lb_params P,
                                                     const lb_texel *__restrict__ items,
                                                     uint32_t n_items, uint32_t key_base,
                                                     float4 *__restrict__ out,
                                                     unsigned long long *stats) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    Counts cnt{0, 0, 0, 0, 0, 0};
    const float4 *tp = reinterpret_cast<const float4 *>(items + i);
    float4 tpos = tp[0], tnrm = tp[1], tdx = tp[2], tdy = tp[3];
    f3 n = xyz(tnrm);
    uint32_t key = (key_base + i) * 7919u;
    uint32_t g = P.counts[0];
    f3 sun = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    frame_t sf = frame(sun);
    float soft = 0.0f;
    uint32_t hard = 0u;
    f3 lift = scl(n, P.reach[2]);
    f3 pos = xyz(tpos), dx = xyz(tdx), dy = xyz(tdy);
    f3 at = pos;
    uint32_t pick = 0u;
    float span = 1.0f;
    if (inside_plain<STATS, WT>(S, P, add(at, lift), n, key, cnt)) {
        for (uint32_t q = 1u; q < 9u; q++) {
            float cx, cy;
            offset(q, &cx, &cy);
            f3 p = add(add(pos, scl(dx, cx)), scl(dy, cy));
            if (!inside_plain<STATS, WT>(S, P, add(p, lift), n, key + q * 31u, cnt)) {
                at = p;
                pick = q;
                span = 0.4f;
                break;
            }
        }
        if (pick == 0u) {
            f3 p = add(pos, scl(n, PUSH));
            if (!inside_plain<STATS, WT>(S, P, add(p, lift), n, key + 331u, cnt)) {
                at = p;
                pick = 9u;
                span = 0.4f;
            }
        }
    }
    uint32_t seed = fbits(P.reach[3]);
    if (dot3(sun, n) > 0.0f) {
        uint32_t gg = g * g;
        for (uint32_t s = 0u; s < gg; s++) {
            float cx = (((float)(s % g) + 0.5f) / (float)g - 0.5f) * span;
            float cy = (((float)(s / g) + 0.5f) / (float)g - 0.5f) * span;
            f3 o = add(add(add(at, scl(dx, cx)), scl(dy, cy)), lift);
            if (trace_plain<true, STATS, WT>(S, P, o, sun, FAR, key + s, cnt).tri == NONE)
                hard |= 1u << s;
            uint32_t k = (s + (uint32_t)(rnd(key, 17u, seed) * (float)gg)) % gg;
            float a = (float)k * 2.399963f;
            float r = P.sun_dir[3] * sqrtf(((float)k + 0.5f) / (float)gg);
            float sa, ca;
            sincos_lb(a, &sa, &ca);
            f3 d = normalize3(add(add(sun, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
            if (trace_plain<true, STATS, WT>(S, P, o, d, FAR, key + s + 1000u, cnt).tri == NONE)
                soft += 1.0f;
        }
        soft /= (float)gg;
    }
    frame_t f = frame(n);
    f3 o = add(at, lift);
    uint32_t rays = P.counts[1];
    float s1 = rnd(key, 3u, seed), s2 = rnd(key, 5u, seed);
    float occ = 0.0f, back = 0.0f;
    for (uint32_t r = 0u; r < rays; r++) {
        f3 d = cosine(fract(((float)r + 0.5f) / (float)rays + s1), fract(radical(r) + s2), f);
        Hit h = trace_plain<false, STATS, WT>(S, P, o, d, P.reach[0], key + r + 2000u, cnt);
        if (h.tri == NONE) continue;
        float k = h.t / P.reach[0];
        occ += 1.0f - k * k;
        if (h.back && h.t < P.ground[3]) back += 1.0f;
    }
    out[i] = make_float4(soft, bitsf(hard | (pick << 24u)), occ / (float)rays, back / (float)rays);
    if (STATS) flush_counts(cnt, stats);
}

// trace.wgsl:245-270
template <bool STATS, bool WT>
__global__ __launch_bounds__(64) void k_gather_plain(SceneView S, lb_params P,
                                                     const lb_texel *__restrict__ items,
                                                     uint32_t n_items, uint32_t key_base,
                                                     float4 *__restrict__ out,
                                                     unsigned long long *stats) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    Counts cnt{0, 0, 0, 0, 0, 0};
    const float4 *tp = reinterpret_cast<const float4 *>(items + i);
    float4 tpos = tp[0], tnrm = tp[1], tdx = tp[2], tdy = tp[3];
    f3 n = xyz(tnrm);
    frame_t f = frame(n);
    uint32_t key = (key_base + i) * 7919u;
    uint32_t rays = P.counts[2];
    uint32_t seed = fbits(P.reach[3]);
    float s1 = rnd(key, 11u, seed), s2 = rnd(key, 13u, seed);
    f3 sum = mk(0.0f, 0.0f, 0.0f);
    float back = 0.0f;
    f3 pos = xyz(tpos), dx = xyz(tdx), dy = xyz(tdy);
    for (uint32_t r = 0u; r < rays; r++) {
        float jx = rnd(key + r, 19u, seed) - 0.5f;
        float jy = rnd(key + r, 23u, seed) - 0.5f;
        f3 o = add(add(add(pos, scl(dx, jx)), scl(dy, jy)), scl(n, P.reach[2]));
        f3 d = cosine(fract(((float)r + 0.5f) / (float)rays + s1), fract(radical(r) + s2), f);
        Hit h = trace_plain<false, STATS, WT>(S, P, o, d, P.reach[1], key + r + 4000u, cnt);
        if (h.tri == NONE) {
            sum = add(sum, sky(S.sh, P, d));
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
        sum = add(sum, bounced(S, P, o, d, h));
    }
    out[i] = make_float4(sum.x * PI / (float)rays, sum.y * PI / (float)rays,
                         sum.z * PI / (float)rays, back / (float)rays);
    if (STATS) flush_counts(cnt, stats);
}

// trace.wgsl:272-324
template <bool STATS, bool WT>
__global__ __launch_bounds__(64) void k_volume_plain(SceneView S, lb_params P,
                                                     const lb_texel *__restrict__ items,
                                                     uint32_t n_items, uint32_t key_base,
                                                     float4 *__restrict__ out,
                                                     unsigned long long *stats) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    Counts cnt{0, 0, 0, 0, 0, 0};
    const float4 *tp = reinterpret_cast<const float4 *>(items + i);
    float4 tpos = tp[0], tnrm = tp[1];
    f3 o = xyz(tpos);
    uint32_t key = fbits(tnrm.w) * 7919u + 77u;
    uint32_t rays = P.counts[2];
    uint32_t seed = fbits(P.reach[3]);
    float s1 = rnd(key, 29u, seed), s2 = rnd(key, 31u, seed);
    f3 l0 = mk(0, 0, 0), l1x = mk(0, 0, 0), l1y = mk(0, 0, 0), l1z = mk(0, 0, 0);
    float back = 0.0f;
    for (uint32_t r = 0u; r < rays; r++) {
        float u = fract(((float)r + 0.5f) / (float)rays + s1);
        float v = fract(radical(r) + s2);
        float z = 1.0f - 2.0f * u;
        float rr = sqrtf(vmax(0.0f, 1.0f - z * z));
        float a = v * 2.0f * PI;
        float sa, ca;
        sincos_lb(a, &sa, &ca);
        f3 d = mk(rr * ca, z, rr * sa);
        Hit h = trace_plain<false, STATS, WT>(S, P, o, d, P.reach[1], key + r, cnt);
        f3 l = sky(S.sh, P, d);
        if (h.tri != NONE) {
            if (h.back) back += 1.0f;
            if (STATS && !h.back) {
                cnt.mats++;
                cnt.gathers++;
            }
            l = h.back ? mk(0.0f, 0.0f, 0.0f) : bounced(S, P, o, d, h);
        }
        l0 = add(l0, scl(l, 0.282095f));
        l1x = add(l1x, scl(scl(l, 0.488603f), d.x));
        l1y = add(l1y, scl(scl(l, 0.488603f), d.y));
        l1z = add(l1z, scl(scl(l, 0.488603f), d.z));
    }
    float w = 4.0f * PI / (float)rays;
    uint32_t g = P.counts[0], gg = g * g;
    f3 sund = mk(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    frame_t sf = frame(sund);
    float sun = 0.0f;
    for (uint32_t s = 0u; s < gg; s++) {
        float a = (float)s * 2.399963f;
        float r = P.sun_dir[3] * sqrtf(((float)s + 0.5f) / (float)gg);
        float sa, ca;
        sincos_lb(a, &sa, &ca);
        f3 d = normalize3(add(add(sund, scl(scl(sf.t, ca), r)), scl(scl(sf.b, sa), r)));
        if (trace_plain<true, STATS, WT>(S, P, o, d, FAR, key + s + 9000u, cnt).tri == NONE)
            sun += 1.0f;
    }
    out[i * 4u] = make_float4(l0.x * w, l0.y * w, l0.z * w, sun / (float)gg);
    out[i * 4u + 1u] = make_float4(l1x.x * w, l1x.y * w, l1x.z * w, back / (float)rays);
    out[i * 4u + 2u] = make_float4(l1y.x * w, l1y.y * w, l1y.z * w, 0.0f);
    out[i * 4u + 3u] = make_float4(l1z.x * w, l1z.y * w, l1z.z * w, 0.0f);
    if (STATS) flush_counts(cnt, stats);
}

hipError_t launch_plain(int pass, bool stats, bool watertight, const SceneView &S,
                        const lb_params &P, const lb_texel *items, uint32_t n, uint32_t key_base,
                        float4 *out, unsigned long long *counters, hipStream_t stream) {
    dim3 grid((n + 63u) / 64u), block(64);
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, S, P, items, n, key_base, out,
                           counters);
    };
    int k = (stats ? 1 : 0) | (watertight ? 2 : 0);
    if (pass == LB_PASS_DIRECT) {
        if (k == 0) go(k_direct_plain<false, false>);
        else if (k == 1) go(k_direct_plain<true, false>);
        else if (k == 2) go(k_direct_plain<false, true>);
        else go(k_direct_plain<true, true>);
    } else if (pass == LB_PASS_GATHER) {
        if (k == 0) go(k_gather_plain<false, false>);
        else if (k == 1) go(k_gather_plain<true, false>);
        else if (k == 2) go(k_gather_plain<false, true>);
        else go(k_gather_plain<true, true>);
    } else {
        if (k == 0) go(k_volume_plain<false, false>);
        else if (k == 1) go(k_volume_plain<true, false>);
        else if (k == 2) go(k_volume_plain<false, true>);
        else go(k_volume_plain<true, true>);
    }
    return hipGetLastError();
}

}  // namespace lbk
