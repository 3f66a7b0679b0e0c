// Launchers of the device passes (host side of the kernels' translation units).
#pragma once

#include <hip/hip_runtime.h>

#include "lb_device.h"

namespace lbk {

struct Tuning {
    int workgroup = 256;     // threads a workgroup (multiple of 64)
    int rays_per_lane = 2;   // rays one lane traces of one item before the wave reduces
    int lds_stack = 16;      // traversal stack entries a lane keeps in LDS
    int nodes = LB_NODES_QUANT;
    int refill = 32;         // pooled: idle lanes (of 64) a wave waits for before it refills
    int schedule = 0;        // 0: lock-step slots, 1: pooled (one ray pool, shaded in order)
    int chunk = 0;           // pooled: hits a chunk holds (0: the default)
    int leaf_batch = 16;     // lanes at a leaf that end a wave's inner phase (64: all)
    int waves = 6;           // lock-step: residency target (0: the compiler's; 6; 8)
    int xcd_split = 0;       // lock-step: items split over the 8 chiplets' queues (0: one queue)
    int gather_order = 0;    // lock-step gather: 0 azimuth, 1 elevation bands then ranked, 2 bands
    bool operator==(const Tuning &o) const {
        return workgroup == o.workgroup && rays_per_lane == o.rays_per_lane &&
               lds_stack == o.lds_stack && nodes == o.nodes && refill == o.refill &&
               schedule == o.schedule && chunk == o.chunk && leaf_batch == o.leaf_batch &&
               waves == o.waves && xcd_split == o.xcd_split && gather_order == o.gather_order;
    }
};
// every field of Tuning is an int and operator== names each one
static_assert(sizeof(Tuning) == 11 * sizeof(int), "a new Tuning field must join operator==");

// LDS words between two lanes' traversal stacks of `depth` entries: odd, so a wave's lanes
// spread over all banks.
__host__ __device__ inline uint32_t lane_stride(uint32_t depth) { return (depth + 1u) | 1u; }
// LDS bytes of a workgroup's traversal stacks
inline size_t stack_lds_bytes(int depth, int workgroup) {
    return (size_t)lane_stride((uint32_t)depth) * (size_t)workgroup * 4u;
}

hipError_t launch_plain(int pass, bool stats, bool watertight, const SceneView &S,
                        const lb_params &P, const lb_texel *items, uint32_t n, uint32_t key_base,
                        float4 *out, unsigned long long *counters, hipStream_t stream);

}  // namespace lbk
