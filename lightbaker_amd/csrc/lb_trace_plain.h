// trace.wgsl:83-130 as the PLAIN form runs it: one ray a lane, the 48-entry stack in private
// memory, FP32 nodes read as the reference reads them (the node, then both children).
#pragma once

#include "lb_device.h"
#include "lb_kernels.h"

namespace lbk {

// trace.wgsl:83-130; WT: the watertight triangle test (wt()) in place of Moller-Trumbore
template <bool ANY, bool STATS, bool WT = false>
__device__ Hit trace_plain(const SceneView &S, const lb_params &P, f3 o, f3 d, float tmax,
                           uint32_t ray, Counts &cnt) {
    Hit hit{tmax, NONE, false};
    f3 inv = inverse_dir(d);
    uint32_t stack[STACK];
    uint32_t sp = 0, n = 0;
    if (STATS) cnt.rays++;
    for (;;) {
        float4 na = S.nodes[n * 2u];
        float4 nb = S.nodes[n * 2u + 1u];
        uint32_t first = fbits(na.w), count = fbits(nb.w);
        if (count > 0u) {
            if (STATS) cnt.leaves++;
            for (uint32_t i = first; i < first + count; i++) {
                float4 v0 = S.tris[i * 3u], e1 = S.tris[i * 3u + 1u], e2 = S.tris[i * 3u + 2u];
                if (STATS) cnt.tris++;
                float det;
                float t = WT ? wt(o, d, v0, e1, e2, &det) : mt(o, d, v0, e1, e2, &det);
                if (t > 1e-4f && t < hit.t) {
                    if (STATS) cnt.mats++;
                    if (stops(S, P, i, v0, ray)) {
                        hit.t = t;
                        hit.tri = i;
                        hit.back = det < 0.0f;
                        if (ANY) return hit;
                    }
                }
            }
        } else {
            float4 a0 = S.nodes[first * 2u], a1 = S.nodes[first * 2u + 1u];
            float4 b0 = S.nodes[first * 2u + 2u], b1 = S.nodes[first * 2u + 3u];
            if (STATS) cnt.inner++;
            float ta = slab(o, inv, xyz(a0), xyz(a1), hit.t);
            float tb = slab(o, inv, xyz(b0), xyz(b1), hit.t);
            bool an = ta <= tb;
            uint32_t near = an ? first : first + 1u;
            uint32_t far = an ? first + 1u : first;
            float tn = vmin(ta, tb), tf = vmax(ta, tb);
            if (tn < FAR) {
                if (tf < FAR && sp < (uint32_t)STACK) stack[sp++] = far;
                n = near;
                continue;
            }
        }
        if (sp == 0u) break;
        sp--;
        n = stack[sp];
    }
    return hit;
}

}  // namespace lbk
