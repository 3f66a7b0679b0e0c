// Host-side state behind the C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <string>

#include "lb_kernels.h"

namespace lbk {

int fail(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

// A device allocation that only grows.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int reserve(size_t bytes);
    int upload(const void *src, size_t bytes, hipStream_t stream);
    void release();
};

}  // namespace lbk

// An lb_gpu is shared by its bakes and adjoints: each holds a reference (gpu_retain) until it is
// destroyed, and lb_close drops the caller's, so the handle outlives every child whatever order the
// caller closes them in.
struct lb_gpu {
    std::atomic<int> refs{1};
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string name;
    unsigned long long *counters = nullptr;  // lb_stats block
    uint32_t *queue = nullptr;                // TUNED: the work queue's head
    size_t lds_per_block = 0;
    lbk::DevBuf run_items, run_light, run_out;  // lb_run's staging
    lbk::DevBuf zero;                           // one zero vec4 (no light)
    lbk::DevBuf word;                           // lb_encode_volume's scale
    lbk::DevBuf spill;                          // TUNED: stack entries beyond the LDS part
    lbk::DevBuf stream_hits;                    // TUNED, pooled: a chunk's hits
    lbk::DevBuf stream_aux;                     // TUNED, pooled: picks, item lists, counts
    int form = LB_FORM_PLAIN;
    lbk::Tuning tuning[3];  // TUNED: the knobs of each pass (LB_PASS_*)
    uint8_t *scratch = nullptr;  // this device's copy of the static scratch (lb_run_device_noalloc)
};

struct lb_scene {
    lb_gpu *gpu = nullptr;  // valid during lb_scene_create only
    int device = 0;
    lbk::DevBuf nodes, tris, materials, charts;
    lbk::DevBuf wide, quant, leaves;  // TUNED node layouts
    lbk::DevBuf tris_t;               // TUNED triangles: opacity in e2.w
    uint32_t n_nodes = 0, n_tris = 0, n_materials = 0, n_charts = 0, n_inner = 0;
    uint32_t root_ref = 0;  // the root as a TUNED child reference
    float sky_sh[12] = {};  // the sky's L1 term, c0 rgb, c1 rgb, c2 rgb, c3 rgb (all 0: none)
};

extern "C" int lb_set_form(lb_gpu *gpu, int form);

namespace lbk {

void gpu_retain(lb_gpu *gpu);
// drops one reference; the last one synchronizes the stream and frees the handle
void gpu_release(lb_gpu *gpu);

bool validate_bvh(const lb_node *nodes, size_t n_nodes, const lb_tri *tris, size_t n_tris,
                  size_t n_materials, size_t n_charts, std::string *err);
int build_tuned_nodes(lb_scene *s, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                      size_t n_tris, const float *materials);
SceneView scene_view(const lb_scene *s, const float *light);
// the kernels' form of 12 sky coefficients (c0 rgb, c1 rgb, c2 rgb, c3 rgb)
SkySH sky_sh_of(const float c[12]);
// The order a TUNED lock-step gather takes each item's rays in (one byte a ray, item-major): with
// `cost`, the launch records each ray's traversal steps by its azimuth slot (a counting launch);
// with `perm`, slot s of item i traces azimuth slot perm[i * rays + s]. Either needs rays <= 256.
// With `rec`, the launch also writes each ray's record (lb_records.h) at r * items + i.
struct GatherOrder {
    uint8_t *cost = nullptr;
    const uint8_t *perm = nullptr;
    uint2 *rec = nullptr;
};
// Device memory a TUNED lock-step launch takes in place of the gpu's own queue and spill area
// (which it grows on demand): the launch then allocates nothing, and fails if the spill area it
// needs exceeds `spill_bytes`.
struct FixedScratch {
    uint32_t *queue = nullptr;  // the chiplets' queue heads, 1 KiB
    uint32_t *spill = nullptr;
    size_t spill_bytes = 0;
};
int run_pass(lb_gpu *gpu, int pass, int form, const lb_scene *scene, const lb_params *params,
             const lb_texel *items, size_t n, uint32_t key_base, const float *light,
             size_t n_light, float *out, bool stats, const GatherOrder *order = nullptr,
             const FixedScratch *fixed = nullptr);
// the spill area a TUNED lock-step launch of `pass` over n items needs with the gpu's tuning
hipError_t tuned_spill_bytes(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n,
                             size_t *bytes);
// each item's azimuth slots ranked by recorded cost, into perm (rays <= 256)
hipError_t rank_slots(const uint8_t *cost, uint32_t n_items, uint32_t rays, uint8_t *perm,
                      hipStream_t stream);
bool is_empty_bvh(const lb_node *nodes, size_t n_nodes);
// sizes the tuned form's spill area for a launch over n items (before a graph capture)
hipError_t tuned_prepare(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n);
hipError_t stream_prepare(lb_gpu *gpu, int pass, const lb_params &P, uint32_t n);
hipError_t launch_stream(lb_gpu *gpu, int pass, bool stats, const SceneView &S,
                         const lb_params &P, const lb_texel *items, uint32_t n, uint32_t key_base,
                         float4 *out);
hipError_t launch_tuned(lb_gpu *gpu, int pass, bool stats, const SceneView &S,
                        const lb_params &P, const lb_texel *items, uint32_t n, uint32_t key_base,
                        float4 *out, const GatherOrder *order = nullptr,
                        const FixedScratch *fixed = nullptr);

}  // namespace lbk
