// Launchers of the device passes (host side of the kernels' translation units).
#pragma once

#include <hip/hip_runtime.h>

#include "lb_device.h"

namespace lbk {

struct Tuning {
    int workgroup = 256;     // threads a workgroup (multiple of 64)
    int rays_per_lane = 2;   // rays one lane traces of one item before the wave reduces
    int lds_stack = 16;      // traversal stack entries a lane keeps in LDS
    int nodes = LB_NODES_WIDE;
    int refill = 16;         // idle lanes (of 64) a wave waits for before it starts new rays
    int schedule = 0;        // 0: lock-step slots, 1: flattened with batched refills,
                             // 2: streamed (generate, trace, shade kernels)
    int chunk = 0;           // streamed: rays a chunk holds (0: the default)
};

hipError_t launch_plain(int pass, bool stats, const SceneView &S, const lb_params &P,
                        const lb_texel *items, uint32_t n, uint32_t key_base, float4 *out,
                        unsigned long long *counters, hipStream_t stream);

}  // namespace lbk
