// The tiles' texels (crates/lightbaker/src/raster.rs:39-114) on the device, equal bit for bit to
// the host restatement (raster.cpp). raster.rs walks each tile's charts and their triangles in
// order twice: first every texel whose centre a triangle covers, then, for texels still empty,
// every texel a triangle only touches, sampled at the triangle's point nearest the centre; a texel
// keeps the first triangle that claims it. Here every triangle (set up on the host exactly as
// raster.rs does: chart-local corners, determinant, normal, texel frame) claims the texels of its
// box in parallel with an atomic minimum of (pass << 31 | its place in that order), and a second
// kernel writes each texel from its winner with raster.rs's arithmetic.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "lb_bake.h"
#include "lb_device.h"
#include "lb_host.h"
#include "lb_raster.h"

namespace lbk {
namespace {

constexpr uint32_t BLOCK = 256;
constexpr uint32_t EMPTY = 0xffffffffu;

struct P2 {
    float x, y;
};

__device__ __forceinline__ float perp(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }
__device__ __forceinline__ P2 sub2(P2 a, P2 b) { return P2{a.x - b.x, a.y - b.y}; }

// raster.rs:103-114
__device__ P2 nearest(P2 p, P2 a, P2 b, P2 c) {
    auto side = [](P2 o, P2 e, P2 q) { return perp(sub2(e, o), sub2(q, o)); };
    float s = signbit(side(a, b, c)) ? -1.0f : 1.0f;
    if (side(a, b, p) * s >= 0.0f && side(b, c, p) * s >= 0.0f && side(c, a, p) * s >= 0.0f)
        return p;
    auto seg = [&](P2 o, P2 e) {
        P2 d = sub2(e, o);
        P2 po = sub2(p, o);
        float len2 = d.x * d.x + d.y * d.y;
        float t = (po.x * d.x + po.y * d.y) / fmaxf(len2, 1e-12f);
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        return P2{o.x + d.x * t, o.y + d.y * t};
    };
    P2 cand[3] = {seg(a, b), seg(b, c), seg(c, a)};
    int best = 0;
    float bd = 0.0f;
    for (int i = 0; i < 3; i++) {
        P2 q = sub2(cand[i], p);
        float d = q.x * q.x + q.y * q.y;
        if (i == 0 || d < bd) {
            best = i;
            bd = d;
        }
    }
    return cand[best];
}

__device__ __forceinline__ void bary(const RasterTri &r, P2 q, float *l1, float *l2, float *l3) {
    *l1 = ((r.ub[1] - r.uc[1]) * (q.x - r.uc[0]) + (r.uc[0] - r.ub[0]) * (q.y - r.uc[1])) / r.det;
    *l2 = ((r.uc[1] - r.ua[1]) * (q.x - r.uc[0]) + (r.ua[0] - r.uc[0]) * (q.y - r.uc[1])) / r.det;
    *l3 = 1.0f - *l1 - *l2;
}

// Triangle k claims the texels of its box: pass 0 where it covers the centre, pass 1 where it
// only touches the texel's square.
__global__ __launch_bounds__(BLOCK) void k_raster_claim(const RasterTri *__restrict__ rec,
                                                        uint32_t n, const uint4 *__restrict__ tiles,
                                                        uint32_t *__restrict__ owner) {
    uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const RasterTri &r = rec[k];
    uint4 t = tiles[r.tile];  // first texel, w, h, atlas corner
    P2 a{r.ua[0], r.ua[1]}, b{r.ub[0], r.ub[1]}, c{r.uc[0], r.uc[1]};
    for (int y = r.lo[1]; y <= r.hi[1]; y++)
        for (int x = r.lo[0]; x <= r.hi[0]; x++) {
            P2 centre{(float)x + 0.5f, (float)y + 0.5f};
            float l1, l2, l3;
            bary(r, centre, &l1, &l2, &l3);
            uint32_t key;
            if (!(l1 < -1e-4f || l2 < -1e-4f || l3 < -1e-4f)) {
                key = k;
            } else {
                P2 q = nearest(centre, a, b, c);
                if (fabsf(q.x - centre.x) > 0.5f || fabsf(q.y - centre.y) > 0.5f) continue;
                key = 0x80000000u | k;
            }
            atomicMin(&owner[t.x + (uint32_t)y * t.y + (uint32_t)x], key);
        }
}

// Texel g from the triangle that claimed it (raster.rs:84-95), or empty.
__global__ __launch_bounds__(BLOCK) void k_raster_write(const RasterTri *__restrict__ rec,
                                                        const uint32_t *__restrict__ owner,
                                                        const uint4 *__restrict__ tiles,
                                                        const uint32_t *__restrict__ tile_of,
                                                        uint32_t n, lb_texel *__restrict__ out) {
    uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
    if (g >= n) return;
    float4 *o = reinterpret_cast<float4 *>(out + g);
    uint32_t key = owner[g];
    if (key == EMPTY) {
        float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        o[0] = z;
        o[1] = z;
        o[2] = z;
        o[3] = z;
        return;
    }
    const RasterTri &r = rec[key & 0x7fffffffu];
    uint4 t = tiles[tile_of[g]];
    uint32_t i = g - t.x;
    P2 centre{(float)(i % t.y) + 0.5f, (float)(i / t.y) + 0.5f};
    P2 q = (key >> 31) ? nearest(centre, P2{r.ua[0], r.ua[1]}, P2{r.ub[0], r.ub[1]},
                                 P2{r.uc[0], r.uc[1]})
                       : centre;
    float l1, l2, l3;
    bary(r, q, &l1, &l2, &l3);
    // pa * l1 + pb * l2 + pc * l3, each a per-component product, added left to right
    float p[3];
    for (int c = 0; c < 3; c++) p[c] = r.pa[c] * l1 + r.pb[c] * l2 + r.pc[c] * l3;
    o[0] = make_float4(p[0], p[1], p[2], 1.0f);
    o[1] = make_float4(r.n[0], r.n[1], r.n[2], __uint_as_float(r.chart));
    o[2] = make_float4(r.dx[0], r.dx[1], r.dx[2], 0.0f);
    o[3] = make_float4(r.dy[0], r.dy[1], r.dy[2], 0.0f);
}

// What the texel passes read of each texel (bake.rs:139-148 with raster.rs:25-36): covered
// (pos.w > 0.5), its chart (none where uncovered) and its normal (w 0).
__global__ __launch_bounds__(BLOCK) void k_raster_derive(const lb_texel *__restrict__ tex,
                                                         uint32_t n, uint32_t *__restrict__ chart,
                                                         float4 *__restrict__ normal,
                                                         uint8_t *__restrict__ covered) {
    uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
    if (g >= n) return;
    const float4 *t = reinterpret_cast<const float4 *>(tex + g);
    float4 p = t[0], nm = t[1];
    bool cv = p.w > 0.5f;
    covered[g] = cv;
    chart[g] = cv ? __float_as_uint(nm.w) : EMPTY;
    normal[g] = make_float4(nm.x, nm.y, nm.z, 0.0f);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

hipError_t raster_derive(const lb_texel *tex, uint32_t n, uint32_t *chart, float4 *normal,
                         uint8_t *covered, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_raster_derive, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream,
                       tex, n, chart, normal, covered);
    return hipGetLastError();
}

hipError_t raster_tiles(const RasterTri *rec, uint32_t n_rec, const uint4 *tiles,
                        const uint32_t *tile_of, uint32_t n_texels, uint32_t *owner,
                        lb_texel *out, hipStream_t stream) {
    if (!n_texels) return hipSuccess;
    dim3 gt((n_texels + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_fill_u32, gt, dim3(BLOCK), 0, stream, owner, EMPTY, n_texels);
    if (n_rec)
        hipLaunchKernelGGL(k_raster_claim, dim3((n_rec + BLOCK - 1) / BLOCK), dim3(BLOCK), 0,
                           stream, rec, n_rec, tiles, owner);
    hipLaunchKernelGGL(k_raster_write, gt, dim3(BLOCK), 0, stream, rec, owner, tiles, tile_of,
                       n_texels, out);
    return hipGetLastError();
}

}  // namespace lbk
