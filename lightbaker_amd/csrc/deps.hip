// What a tile's light depends on (crates/lightbaker/src/deps.rs:72-121): every casting triangle
// whose box meets the tile's zone swept towards the sun. The reference tests every triangle for
// every tile on the host (O(tiles x triangles)); here one launch tests every (zone, triangle)
// pair, a lane a zone and 32 triangles, and writes one word of hit bits, with `swept`'s FP32
// operations in the reference's order, so the bits equal the CPU restatement's.
#include <hip/hip_runtime.h>

#include "lb_bake.h"
#include "lb_host.h"

namespace lbk {
namespace {

constexpr uint32_t BLOCK = 256;

// f32::max / f32::min (a NaN operand is ignored)
__device__ __forceinline__ float max_rs(float a, float b) {
    return a != a ? b : (b != b ? a : (a > b ? a : b));
}
__device__ __forceinline__ float min_rs(float a, float b) {
    return a != a ? b : (b != b ? a : (a < b ? a : b));
}

// deps.rs:104-121: does box b meet zone z swept towards the sun up to `top`?
__device__ bool swept(const float z[6], float3 sun, float top, const float b[6]) {
    float s3[3] = {sun.x, sun.y, sun.z};
    float tmax = sun.y > 1e-4f ? max_rs((top - z[1]) / sun.y, 0.0f) : 0.0f;
    float lo = 0.0f, hi = tmax;
    for (int a = 0; a < 3; a++) {
        float s = s3[a], zmin = z[a], zmax = z[3 + a], bmin = b[a], bmax = b[3 + a];
        if (fabsf(s) < 1e-6f) {
            if (zmin > bmax || zmax < bmin) return false;
            continue;
        }
        float t1 = (bmax - zmin) / s, t2 = (bmin - zmax) / s;
        float a0 = s > 0.0f ? t2 : t1, a1 = s > 0.0f ? t1 : t2;
        lo = max_rs(lo, a0);
        hi = min_rs(hi, a1);
        if (lo > hi) return false;
    }
    return true;
}

__global__ __launch_bounds__(BLOCK) void k_swept(const float *__restrict__ pos,
                                                 const uint8_t *__restrict__ casts, uint32_t n_tris,
                                                 uint32_t words, const float *__restrict__ zones,
                                                 uint32_t n_zones, float3 sun, float top,
                                                 uint32_t *__restrict__ bits) {
    size_t n = (size_t)n_zones * words;
    for (size_t g = (size_t)blockIdx.x * BLOCK + threadIdx.x; g < n; g += (size_t)gridDim.x * BLOCK) {
        uint32_t zi = (uint32_t)(g / words), w = (uint32_t)(g % words);
        float z[6];
        for (int k = 0; k < 6; k++) z[k] = zones[(size_t)zi * 6 + k];
        uint32_t word = 0;
        for (uint32_t k = 0; k < 32; k++) {
            uint32_t t = w * 32 + k;
            if (t >= n_tris) break;
            if (casts && !casts[t]) continue;
            // deps.rs:40-52: the triangle's box (Aabb::empty, then its three corners)
            const float *p = pos + (size_t)t * 9;
            float b[6];
            for (int a = 0; a < 3; a++) {
                b[a] = min_rs(min_rs(min_rs(INFINITY, p[a]), p[3 + a]), p[6 + a]);
                b[3 + a] = max_rs(max_rs(max_rs(-INFINITY, p[a]), p[3 + a]), p[6 + a]);
            }
            if (swept(z, sun, top, b)) word |= 1u << k;
        }
        bits[g] = word;
    }
}

}  // namespace
}  // namespace lbk

extern "C" int lb_deps_swept(lb_gpu *gpu, const float *positions, size_t n_tris,
                             const uint8_t *casts, const float *zones, size_t n_zones,
                             const float sun[3], float top, uint32_t *bits) {
    using namespace lbk;
    if (!gpu) return fail("deps: no gpu");
    if ((n_tris && !positions) || (n_zones && !zones) || !sun || (n_zones && n_tris && !bits))
        return fail("deps: null buffer");
    if (n_tris > 0xffffffe0u || n_zones > 0xffffffffu) return fail("deps: too many triangles or zones");
    size_t words = (n_tris + 31) / 32;
    if (!n_zones || !words) return 0;
    hipSetDevice(gpu->device);
    hipStream_t st = gpu->stream;
    DevBuf d_pos, d_casts, d_zones, d_bits;
    if (d_pos.upload(positions, n_tris * 36, st) || d_zones.upload(zones, n_zones * 24, st) ||
        (casts && d_casts.upload(casts, n_tris, st)) || d_bits.reserve(n_zones * words * 4))
        return 1;
    size_t n = n_zones * words, blocks = (n + BLOCK - 1) / BLOCK;
    dim3 grid((uint32_t)(blocks < (1u << 20) ? blocks : (1u << 20)));
    hipLaunchKernelGGL(k_swept, grid, dim3(BLOCK), 0, st, (const float *)d_pos.ptr,
                       casts ? (const uint8_t *)d_casts.ptr : nullptr, (uint32_t)n_tris,
                       (uint32_t)words, (const float *)d_zones.ptr, (uint32_t)n_zones,
                       make_float3(sun[0], sun[1], sun[2]), top, (uint32_t *)d_bits.ptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(bits, d_bits.ptr, n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : hip_fail(e, "deps");
}
