// The C-ABI of include/lightbaker.h: the reference's `Gpu` (crates/lightbaker/src/gpu/mod.rs)
// on one HIP device, plus the device-resident extensions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lb_host.h"
#include "lb_kernels.h"

using namespace lbk;

namespace {
thread_local std::string g_error;

// lb_run_device_noalloc's scratch: the TUNED lock-step launch's work-queue heads (the first
// 1 KiB), then its traversal spill area of 64 MiB, which holds the spill of every lane a 256-CU
// device keeps resident (32 waves a CU: 524 288 lanes x 32 entries x 4 B) with 16 or more of
// the 48 stack entries in LDS. Each device holds its own copy for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress gives on that device
// (lb_gpu.scratch), never through this name.
constexpr size_t SCRATCH_QUEUE_BYTES = 1024;
constexpr size_t SCRATCH_SPILL_BYTES = size_t(64) << 20;
constexpr size_t SCRATCH_BYTES = SCRATCH_QUEUE_BYTES + SCRATCH_SPILL_BYTES;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];
}

namespace lbk {

int fail(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return 1;
}

int hip_fail(hipError_t e, const char *what) {
    return fail("GPU: %s: %s", what, hipGetErrorString(e));
}

}  // namespace lbk

#define HIP_TRY(expr, what)                           \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

extern "C" {

const char *lb_last_error(void) { return g_error.c_str(); }

int lb_open(lb_gpu **out) {
    if (!out) return fail("lb_open: null handle");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return fail("no GPU adapter");
    auto *g = new lb_gpu();
    g->device = 0;
    if (hipSetDevice(0) != hipSuccess) {
        delete g;
        return fail("no GPU device");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        delete g;
        return fail("no GPU device");
    }
    // the marketing name where the runtime reports one, else the device's PCI id
    std::string dev = prop.name;
    if (dev.empty()) {
        char id[64];
        std::snprintf(id, sizeof id, "AMD GPU %04x:%02x:%02x", prop.pciDomainID, prop.pciBusID,
                      prop.pciDeviceID);
        dev = id;
    }
    g->name = dev + " (HIP " + prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) +
              " CUs)";
    g->cus = prop.multiProcessorCount;
    g->lds_per_block = prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess ||
        hipMalloc(&g->counters, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&g->queue, 1024) != hipSuccess || g->zero.reserve(64) ||
        hipMemset(g->zero.ptr, 0, 64) != hipSuccess || hipMemset(g->queue, 0, 1024) != hipSuccess) {
        delete g;
        return fail("no GPU device");
    }
    // this device's copy of the no-allocation entry point's scratch (device 0 is current)
    hipError_t se = hipGetSymbolAddress(reinterpret_cast<void **>(&g->scratch), HIP_SYMBOL(scratch));
    if (se != hipSuccess || !g->scratch) {
        delete g;
        return fail("GPU: the static scratch's device address: %s", hipGetErrorString(se));
    }
    *out = g;
    return 0;
}

const char *lb_adapter_name(const lb_gpu *gpu) { return gpu ? gpu->name.c_str() : ""; }

void lb_close(lb_gpu *gpu) { gpu_release(gpu); }

}  // extern "C"

namespace lbk {

void gpu_retain(lb_gpu *gpu) {
    if (gpu) gpu->refs.fetch_add(1);
}

void gpu_release(lb_gpu *gpu) {
    if (!gpu || gpu->refs.fetch_sub(1) != 1) return;
    hipSetDevice(gpu->device);
    hipStreamSynchronize(gpu->stream);
    gpu->run_items.release();
    gpu->run_light.release();
    gpu->run_out.release();
    gpu->zero.release();
    gpu->word.release();
    gpu->spill.release();
    gpu->stream_hits.release();
    gpu->stream_aux.release();
    if (gpu->counters) hipFree(gpu->counters);
    if (gpu->queue) hipFree(gpu->queue);
    hipEventDestroy(gpu->ev0);
    hipEventDestroy(gpu->ev1);
    hipStreamDestroy(gpu->stream);
    delete gpu;
}

}  // namespace lbk

extern "C" {

int lb_scene_create(lb_gpu *gpu, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                    size_t n_tris, const float *materials, size_t n_materials, const float *charts,
                    size_t n_charts, lb_scene **out) {
    if (!gpu || !out) return fail("lb_scene_create: null handle");
    *out = nullptr;
    if (!nodes || n_nodes == 0) return fail("lb_scene_create: a BVH has at least one node");
    if (!tris || n_tris == 0) return fail("lb_scene_create: a BVH has at least one triangle");
    if (!materials || n_materials == 0) return fail("lb_scene_create: no materials");
    if (n_nodes > 0x7fffffffu || n_tris > 0x07ffffffu)
        return fail("lb_scene_create: %zu nodes / %zu triangles exceed the 32-bit indices",
                    n_nodes, n_tris);
    std::string err;
    if (!validate_bvh(nodes, n_nodes, tris, n_tris, n_materials, n_charts, &err))
        return fail("lb_scene_create: %s", err.c_str());
    hipSetDevice(gpu->device);
    auto *s = new lb_scene();
    s->gpu = gpu;
    s->device = gpu->device;
    static const float zero_chart[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const float *ch = n_charts ? charts : zero_chart;
    size_t nch = n_charts ? n_charts : 1;
    s->n_nodes = (uint32_t)n_nodes;
    s->n_tris = (uint32_t)n_tris;
    s->n_materials = (uint32_t)n_materials;
    s->n_charts = (uint32_t)nch;
    int rc = 0;
    if (is_empty_bvh(nodes, n_nodes)) {
        // the reference's empty marker (bvh.rs:76-80) as one leaf over the zero triangle, which
        // no ray meets: read as the reference reads it the marker would be an inner node whose
        // children lie past the array
        lb_node leaf = nodes[0];
        leaf.count = 1;
        rc |= s->nodes.upload(&leaf, sizeof leaf, gpu->stream);
    } else {
        rc |= s->nodes.upload(nodes, n_nodes * sizeof(lb_node), gpu->stream);
    }
    rc |= s->tris.upload(tris, n_tris * sizeof(lb_tri), gpu->stream);
    rc |= s->materials.upload(materials, n_materials * 16, gpu->stream);
    rc |= s->charts.upload(ch, nch * 32, gpu->stream);
    if (rc == 0) rc = build_tuned_nodes(s, nodes, n_nodes, tris, n_tris, materials);
    if (rc == 0 && hipStreamSynchronize(gpu->stream) != hipSuccess)
        rc = fail("GPU: scene upload failed");
    if (rc) {
        delete s;
        return 1;
    }
    *out = s;
    return 0;
}

int lb_scene_set_sky(lb_scene *scene, const float *sh) {
    if (!scene) return fail("lb_scene_set_sky: null handle");
    for (int k = 0; k < 12; k++) {
        float v = sh ? sh[k] : 0.0f;
        if (!std::isfinite(v)) return fail("lb_scene_set_sky: coefficient %d is not finite", k);
        scene->sky_sh[k] = v;
    }
    return 0;
}

void lb_scene_destroy(lb_scene *scene) {
    if (!scene) return;
    // the scene may outlive its lb_gpu handle: touch only the device, never the handle
    hipSetDevice(scene->device);
    hipDeviceSynchronize();
    delete scene;
}

int lb_buffer_alloc(lb_gpu *gpu, size_t bytes, void **dev_out) {
    if (!gpu || !dev_out) return fail("lb_buffer_alloc: null handle");
    hipSetDevice(gpu->device);
    HIP_TRY(hipMalloc(dev_out, bytes ? bytes : 16), "hipMalloc");
    return 0;
}

int lb_buffer_free(lb_gpu *gpu, void *dev) {
    if (!gpu) return fail("lb_buffer_free: null handle");
    if (!dev) return 0;
    hipSetDevice(gpu->device);
    hipStreamSynchronize(gpu->stream);
    HIP_TRY(hipFree(dev), "hipFree");
    return 0;
}

int lb_upload(lb_gpu *gpu, void *dst, const void *src, size_t bytes) {
    if (!gpu) return fail("lb_upload: null handle");
    if (!bytes) return 0;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, gpu->stream), "upload");
    HIP_TRY(hipStreamSynchronize(gpu->stream), "upload");
    return 0;
}

int lb_download(lb_gpu *gpu, void *dst, const void *src, size_t bytes) {
    if (!gpu) return fail("lb_download: null handle");
    if (!bytes) return 0;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, gpu->stream), "GPU read back");
    HIP_TRY(hipStreamSynchronize(gpu->stream), "GPU read back");
    return 0;
}

int lb_sync(lb_gpu *gpu) {
    if (!gpu) return fail("lb_sync: null handle");
    HIP_TRY(hipStreamSynchronize(gpu->stream), "GPU");
    return 0;
}

int lb_run_device(lb_gpu *gpu, int pass, int form, const lb_scene *scene, const lb_params *params,
                  const lb_texel *items_dev, size_t n, uint32_t key_base, const float *light_dev,
                  size_t n_light, float *out_dev) {
    return run_pass(gpu, pass, form, scene, params, items_dev, n, key_base, light_dev, n_light,
                    out_dev, false);
}

int lb_run_device_noalloc(lb_gpu *gpu, int pass, int form, const lb_scene *scene,
                          const lb_params *params, const lb_texel *items_dev, size_t n,
                          uint32_t key_base, const float *light_dev, size_t n_light,
                          float *out_dev) {
    if (!gpu || !scene || !params) return fail("lb_run_device_noalloc: null handle");
    if (pass < LB_PASS_DIRECT || pass > LB_PASS_VOLUME)
        return fail("lb_run_device_noalloc: no pass %d", pass);
    if (form != LB_FORM_TUNED)  // the other forms' launches use no scratch
        return run_pass(gpu, pass, form, scene, params, items_dev, n, key_base, light_dev,
                        n_light, out_dev, false);
    if (gpu->tuning[pass].schedule != 0)
        return fail("lb_run_device_noalloc: the pooled schedule keeps every ray's hit in memory "
                    "sized by the launch; set schedule 0 (lock-step) for this pass");
    if (n == 0) return 0;
    if (n > 0xffffffffu) return fail("lb_run_device_noalloc: %zu items exceed 32-bit indices", n);
    hipSetDevice(gpu->device);
    size_t need = 0;
    HIP_TRY(tuned_spill_bytes(gpu, pass, *params, (uint32_t)n, &need), "occupancy");
    const size_t have = SCRATCH_SPILL_BYTES;
    if (need > have)
        return fail("lb_run_device_noalloc: this launch spills %zu bytes of traversal stack, the "
                    "static scratch holds %zu: keep more stack entries in LDS (lds_stack) or use "
                    "a smaller workgroup",
                    need, have);
    FixedScratch fx;
    fx.queue = reinterpret_cast<uint32_t *>(gpu->scratch);
    fx.spill = reinterpret_cast<uint32_t *>(gpu->scratch + SCRATCH_QUEUE_BYTES);
    fx.spill_bytes = have;
    return run_pass(gpu, pass, form, scene, params, items_dev, n, key_base, light_dev, n_light,
                    out_dev, false, nullptr, &fx);
}

int lb_run_stats(lb_gpu *gpu, int pass, int form, const lb_scene *scene, const lb_params *params,
                 const lb_texel *items_dev, size_t n, uint32_t key_base, const float *light_dev,
                 size_t n_light, float *out_dev, lb_stats *stats) {
    if (!gpu || !stats) return fail("lb_run_stats: null handle");
    HIP_TRY(hipMemsetAsync(gpu->counters, 0, 8 * sizeof(unsigned long long), gpu->stream),
            "counters");
    if (run_pass(gpu, pass, form, scene, params, items_dev, n, key_base, light_dev, n_light,
                 out_dev, true))
        return 1;
    unsigned long long c[8];
    HIP_TRY(hipMemcpyAsync(c, gpu->counters, sizeof c, hipMemcpyDeviceToHost, gpu->stream),
            "counters");
    HIP_TRY(hipStreamSynchronize(gpu->stream), "GPU");
    stats->rays = c[0];
    stats->inner_visits = c[1];
    stats->tri_tests = c[2];
    stats->material_loads = c[3];
    stats->light_gathers = c[4];
    stats->leaf_visits = c[5];
    return 0;
}

int lb_run(lb_gpu *gpu, int pass, const lb_scene *scene, const lb_params *params,
           const lb_texel *items, size_t n, const float *light, size_t n_light, float *out,
           size_t out_cap) {
    if (!gpu || !scene || !params) return fail("lb_run: null handle");
    if (pass < LB_PASS_DIRECT || pass > LB_PASS_VOLUME) return fail("lb_run: no pass %d", pass);
    if (n == 0) return 0;
    if (!items || !out) return fail("lb_run: null items or output");
    size_t width = pass == LB_PASS_VOLUME ? 4 : 1;
    if (out_cap < n * width)
        return fail("lb_run: output holds %zu vec4s, %zu items need %zu", out_cap, n, n * width);
    hipSetDevice(gpu->device);
    static const float zero_light[4] = {0, 0, 0, 0};
    const float *lt = (light && n_light) ? light : zero_light;
    size_t nl = (light && n_light) ? n_light : 1;
    if (gpu->run_items.reserve(n * sizeof(lb_texel)) || gpu->run_light.reserve(nl * 16) ||
        gpu->run_out.reserve(n * width * 16))
        return 1;
    HIP_TRY(hipMemcpyAsync(gpu->run_items.ptr, items, n * sizeof(lb_texel), hipMemcpyHostToDevice,
                           gpu->stream),
            "upload");
    HIP_TRY(hipMemcpyAsync(gpu->run_light.ptr, lt, nl * 16, hipMemcpyHostToDevice, gpu->stream),
            "upload");
    if (run_pass(gpu, pass, gpu->form, scene, params,
                 static_cast<const lb_texel *>(gpu->run_items.ptr), n, 0,
                 static_cast<const float *>(gpu->run_light.ptr), nl,
                 static_cast<float *>(gpu->run_out.ptr), false))
        return 1;
    HIP_TRY(hipMemcpyAsync(out, gpu->run_out.ptr, n * width * 16, hipMemcpyDeviceToHost,
                           gpu->stream),
            "GPU read back");
    HIP_TRY(hipStreamSynchronize(gpu->stream), "GPU");
    return 0;
}

int lb_set_form(lb_gpu *gpu, int form) {
    if (!gpu) return fail("lb_set_form: null handle");
    if (form != LB_FORM_PLAIN && form != LB_FORM_TUNED && form != LB_FORM_WATERTIGHT)
        return fail("lb_set_form: no form %d", form);
    gpu->form = form;
    return 0;
}

static int check_tuning(const lb_gpu *gpu, const lb_tuning *t, const char *who) {
    if (!gpu || !t) return fail("%s: null handle", who);
    if (t->workgroup < 64 || t->workgroup > 1024 || t->workgroup % 64)
        return fail("%s: workgroup %d is not a multiple of 64 in 64..1024", who, t->workgroup);
    if (t->rays_per_lane < 1 || t->rays_per_lane > 16)
        return fail("%s: rays per lane %d outside 1..16", who, t->rays_per_lane);
    if (t->lds_stack < 8 || t->lds_stack > 48)
        return fail("%s: LDS stack %d outside 8..48", who, t->lds_stack);
    if (t->nodes != LB_NODES_WIDE && t->nodes != LB_NODES_QUANT)
        return fail("%s: no node layout %d", who, t->nodes);
    if (t->refill < 1 || t->refill > 64) return fail("%s: refill %d outside 1..64", who, t->refill);
    if (t->schedule < 0 || t->schedule > 1) return fail("%s: no schedule %d", who, t->schedule);
    if (t->chunk < 0) return fail("%s: chunk %d is negative", who, t->chunk);
    if (t->leaf_batch < 1 || t->leaf_batch > 64)
        return fail("%s: leaf batch %d outside 1..64", who, t->leaf_batch);
    if (t->waves != 0 && t->waves != 6 && t->waves != 8)
        return fail("%s: waves %d is not 0, 6 or 8", who, t->waves);
    if (t->xcd_split != 0 && t->xcd_split != 1)
        return fail("%s: xcd_split %d is not 0 or 1", who, t->xcd_split);
    if (t->gather_order < 0 || t->gather_order > 2)
        return fail("%s: gather_order %d is not 0, 1 or 2", who, t->gather_order);
    if (stack_lds_bytes(t->lds_stack, t->workgroup) > gpu->lds_per_block)
        return fail("%s: an LDS stack of %d entries for %d lanes exceeds %zu bytes", who,
                    t->lds_stack, t->workgroup, gpu->lds_per_block);
    return 0;
}

static void put_tuning(Tuning &d, const lb_tuning *t) {
    d.workgroup = t->workgroup;
    d.rays_per_lane = t->rays_per_lane;
    d.lds_stack = t->lds_stack;
    d.nodes = t->nodes;
    d.refill = t->refill;
    d.schedule = t->schedule;
    d.chunk = t->chunk;
    d.leaf_batch = t->leaf_batch;
    d.waves = t->waves;
    d.xcd_split = t->xcd_split;
    d.gather_order = t->gather_order;
}

static void get_tuning(const Tuning &s, lb_tuning *t) {
    t->workgroup = s.workgroup;
    t->rays_per_lane = s.rays_per_lane;
    t->lds_stack = s.lds_stack;
    t->nodes = s.nodes;
    t->refill = s.refill;
    t->schedule = s.schedule;
    t->chunk = s.chunk;
    t->leaf_batch = s.leaf_batch;
    t->waves = s.waves;
    t->xcd_split = s.xcd_split;
    t->gather_order = s.gather_order;
}

int lb_set_tuning(lb_gpu *gpu, const lb_tuning *t) {
    if (check_tuning(gpu, t, "lb_set_tuning")) return 1;
    for (auto &d : gpu->tuning) put_tuning(d, t);
    return 0;
}

int lb_get_tuning(const lb_gpu *gpu, lb_tuning *t) {
    if (!gpu || !t) return fail("lb_get_tuning: null handle");
    get_tuning(gpu->tuning[LB_PASS_GATHER], t);
    return 0;
}

int lb_set_pass_tuning(lb_gpu *gpu, int pass, const lb_tuning *t) {
    if (check_tuning(gpu, t, "lb_set_pass_tuning")) return 1;
    if (pass < LB_PASS_DIRECT || pass > LB_PASS_VOLUME)
        return fail("lb_set_pass_tuning: no pass %d", pass);
    put_tuning(gpu->tuning[pass], t);
    return 0;
}

int lb_get_pass_tuning(const lb_gpu *gpu, int pass, lb_tuning *t) {
    if (!gpu || !t) return fail("lb_get_pass_tuning: null handle");
    if (pass < LB_PASS_DIRECT || pass > LB_PASS_VOLUME)
        return fail("lb_get_pass_tuning: no pass %d", pass);
    get_tuning(gpu->tuning[pass], t);
    return 0;
}

int lb_timer_begin(lb_gpu *gpu) {
    if (!gpu) return fail("lb_timer_begin: null handle");
    HIP_TRY(hipEventRecord(gpu->ev0, gpu->stream), "event");
    return 0;
}

int lb_timer_end(lb_gpu *gpu, float *ms) {
    if (!gpu || !ms) return fail("lb_timer_end: null handle");
    HIP_TRY(hipEventRecord(gpu->ev1, gpu->stream), "event");
    HIP_TRY(hipEventSynchronize(gpu->ev1), "GPU");
    HIP_TRY(hipEventElapsedTime(ms, gpu->ev0, gpu->ev1), "event");
    return 0;
}

}  // extern "C"

namespace lbk {

int DevBuf::reserve(size_t bytes) {
    if (bytes <= cap) return 0;
    release();
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    cap = bytes;
    return 0;
}

int DevBuf::upload(const void *src, size_t bytes, hipStream_t stream) {
    if (reserve(bytes ? bytes : 16)) return 1;
    if (!bytes) return 0;
    hipError_t e = hipMemcpyAsync(ptr, src, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "upload");
    // the source is borrowed for the call only: finish the copy before returning
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "upload");
    return 0;
}

void DevBuf::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

SkySH sky_sh_of(const float c[12]) {
    SkySH K;
    K.r = make_float4(c[0], c[3], c[6], c[9]);
    K.g = make_float4(c[1], c[4], c[7], c[10]);
    K.b = make_float4(c[2], c[5], c[8], c[11]);
    K.on = 0;
    for (int k = 0; k < 12; k++) K.on |= c[k] != 0.0f;
    return K;
}

SceneView scene_view(const lb_scene *s, const float *light) {
    SceneView v;
    v.nodes = static_cast<const float4 *>(s->nodes.ptr);
    v.tris = static_cast<const float4 *>(s->tris.ptr);
    v.materials = static_cast<const float4 *>(s->materials.ptr);
    v.charts = static_cast<const float4 *>(s->charts.ptr);
    v.light = reinterpret_cast<const float4 *>(light);
    v.wide = static_cast<const float4 *>(s->wide.ptr);
    v.quant = static_cast<const uint4 *>(s->quant.ptr);
    v.leaves = static_cast<const uint2 *>(s->leaves.ptr);
    v.tris_t = static_cast<const float4 *>(s->tris_t.ptr);
    v.n_nodes = s->n_nodes;
    v.root = s->root_ref;
    v.n_tris = s->n_tris;
    v.sh = sky_sh_of(s->sky_sh);
    return v;
}

int check_params(const lb_scene *scene, const lb_params *P, int pass, size_t n_light) {
    uint32_t g = P->counts[0];
    if (pass != LB_PASS_GATHER && (g < 1 || g * g > 24))
        return fail("params: sun grid side %u (g*g must be 1..24: bits 24.. hold the pick)", g);
    if (pass == LB_PASS_DIRECT && P->counts[1] < 1) return fail("params: no ao rays");
    if (pass != LB_PASS_DIRECT && P->counts[2] < 1) return fail("params: no gather rays");
    if (pass != LB_PASS_DIRECT) {
        uint64_t texels = (uint64_t)P->atlas[0] * P->atlas[1];
        if (P->atlas[0] < 1 || P->atlas[1] < 1) return fail("params: empty atlas");
        if (texels > n_light)
            return fail("params: atlas %ux%u but the light holds %zu texels", P->atlas[0],
                        P->atlas[1], n_light);
    }
    (void)scene;
    return 0;
}

int run_pass(lb_gpu *gpu, int pass, int form, const lb_scene *scene, const lb_params *params,
             const lb_texel *items, size_t n, uint32_t key_base, const float *light,
             size_t n_light, float *out, bool stats, const GatherOrder *order,
             const FixedScratch *fixed) {
    if (!gpu || !scene || !params) return fail("run: null handle");
    if (pass < LB_PASS_DIRECT || pass > LB_PASS_VOLUME) return fail("run: no pass %d", pass);
    if (form != LB_FORM_PLAIN && form != LB_FORM_TUNED && form != LB_FORM_WATERTIGHT)
        return fail("run: no form %d", form);
    if (n == 0) return 0;
    if (n > 0xffffffffu) return fail("run: %zu items exceed 32-bit item indices", n);
    if (!items || !out) return fail("run: null items or output");
    if (pass != LB_PASS_DIRECT && !light) return fail("run: the gather and volume read a light");
    if (check_params(scene, params, pass, pass == LB_PASS_DIRECT ? 1 : n_light)) return 1;
    if (form == LB_FORM_TUNED &&
        stack_lds_bytes(gpu->tuning[pass].lds_stack, gpu->tuning[pass].workgroup) >
            gpu->lds_per_block)
        return fail("run: an LDS stack of %d entries for %d lanes exceeds %zu bytes a workgroup",
                    gpu->tuning[pass].lds_stack, gpu->tuning[pass].workgroup, gpu->lds_per_block);
    hipSetDevice(gpu->device);
    lb_params P = *params;
    P.counts[3] = (uint32_t)n;
    P.atlas[2] = key_base;
    SceneView S = scene_view(scene, light ? light : static_cast<const float *>(gpu->zero.ptr));
    hipError_t e;
    if (form != LB_FORM_TUNED)
        e = launch_plain(pass, stats, form == LB_FORM_WATERTIGHT, S, P, items, (uint32_t)n,
                         key_base, reinterpret_cast<float4 *>(out), gpu->counters, gpu->stream);
    else
        e = launch_tuned(gpu, pass, stats, S, P, items, (uint32_t)n, key_base,
                         reinterpret_cast<float4 *>(out), order, fixed);
    if (e != hipSuccess) return hip_fail(e, "launch");
    return 0;
}

}  // namespace lbk
