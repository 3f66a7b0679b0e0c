// The adjoint of the bake's bounces (crates/lightbaker/src/bake.rs:243-330 with the gather pass of
// trace.wgsl:245-270): the gradient of a loss on the final light atlas by the materials' albedo,
// the sky's zenith, horizon and ground colours and the sun's irradiance, with visibility held
// fixed. The reference has no gradients; this follows its forward exactly.
//
// Visibility is fixed once: k_record traces every gather ray as the gather does and keeps what
// its light depends on (the sky direction's y, the atlas texel a front hit reads, the material).
// A forward replays those records through the bounces with new parameters and keeps each
// bounce's atlas; the backward runs the bounces in reverse:
//   atlas -> texels      (write_light, bake.rs:435-444): a known texel's light is direct + indirect
//   fill^T, smooth^T     (filter.rs:20-94; weights and masks depend on geometry only)
//   scatter^T            (bake.rs:309-316)
//   gather^T             per ray: sky, albedo and the atlas texel it read (k_gather_t)
// and the sun's share through the direct light (linear in the irradiance: its derivative is the
// direct pass's filled cos x sun share).
//
// The backward's sums run over many terms (an atlas texel's adjoint gathers every ray that read
// it; a parameter's, every ray and texel) and are FP32 with exact compensation: each sum is a
// pair (hi, lo), every addition into hi by an atomic that returns the value it added to, so the
// addition's rounding error is known exactly (Knuth's TwoSum) and goes into lo; block partials and
// the final reduction add pairs the same way. No FP64 on the device. The forward replay repeats the
// gather's FP32 arithmetic in the reference's order.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "lb_bake_state.h"
#include "lb_records.h"
#include "lb_trace_plain.h"

using namespace lbk;

namespace lbk {
namespace {

constexpr int8_t LEVEL_NEVER = 127;  // a texel no ring fills
constexpr uint32_t BLOCK = 256;

__device__ __forceinline__ f3 f3of(float4 v) { return mk(v.x, v.y, v.z); }

// Ray r of gather item i, traced as trace.wgsl:248-262 traces it (item keys from 0, as the bake
// runs the pass), one lane a ray; its record at r * n + i.
__global__ __launch_bounds__(BLOCK) void k_record(SceneView S, lb_params P,
                                                  const lb_texel *__restrict__ items, uint32_t n,
                                                  uint32_t rays, uint2 *__restrict__ rec) {
    // rays a lane over the grid (a launch holds under 2^32 work-items)
    for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t < (uint64_t)n * rays;
         t += (uint64_t)gridDim.x * BLOCK) {
        uint32_t r = (uint32_t)(t / n), i = (uint32_t)(t % n);
        const float4 *tp = reinterpret_cast<const float4 *>(items + i);
        float4 tpos = tp[0], tnrm = tp[1], tdx = tp[2], tdy = tp[3];
        f3 nrm = xyz(tnrm);
        frame_t f = frame(nrm);
        uint32_t key = i * 7919u;
        uint32_t seed = fbits(P.reach[3]);
        float s1 = rnd(key, 11u, seed), s2 = rnd(key, 13u, seed);
        float jx = rnd(key + r, 19u, seed) - 0.5f;
        float jy = rnd(key + r, 23u, seed) - 0.5f;
        f3 o = add(add(add(xyz(tpos), scl(xyz(tdx), jx)), scl(xyz(tdy), jy)), scl(nrm, P.reach[2]));
        f3 d = cosine(fract(((float)r + 0.5f) / (float)rays + s1), fract(radical(r) + s2), f);
        Counts cnt{0, 0, 0, 0, 0, 0};
        Hit h = trace_plain<false, false>(S, P, o, d, P.reach[1], key + r + 4000u, cnt);
        rec[t] = gather_record(S, P, o, d, h.tri, h.back, h.t);
    }
}

// The gather's result from the records: trace.wgsl:256-269's sum in ray order, with its
// arithmetic (sky = mix(horizon, zenith, y) or ground; bounced = albedo / pi * light or the
// fallback where the atlas has none).
__global__ __launch_bounds__(BLOCK) void k_replay(const uint2 *__restrict__ rec, uint32_t n,
                                                  uint32_t rays, const float4 *__restrict__ albedo,
                                                  const float4 *__restrict__ light, SkyP K,
                                                  float4 *__restrict__ out) {
    uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
    if (g >= n) return;
    f3 sum = mk(0.0f, 0.0f, 0.0f);
    float back = 0.0f;
    for (uint32_t r = 0; r < rays; r++) {
        uint2 q = rec[(size_t)r * n + g];
        uint32_t kind = q.y >> 30;
        if (kind == REC_MISS) {
            sum = add(sum, replay_sky(K, q.x, g, r, rays));
            continue;
        }
        if (kind == REC_BACK) {
            back += 1.0f;
            continue;
        }
        f3 e = K.fallback;
        if (kind == REC_ATLAS) {
            float4 l = light[q.x];
            if (l.w > 0.5f) e = f3of(l);
        }
        float4 a = albedo[q.y & 0x3fffffffu];
        sum = add(sum, mul(mk(a.x / PI, a.y / PI, a.z / PI), e));
    }
    out[g] = make_float4(sum.x * PI / (float)rays, sum.y * PI / (float)rays,
                         sum.z * PI / (float)rays, back / (float)rays);
}

}  // namespace

hipError_t record_gather(const SceneView &S, const lb_params &P, const lb_texel *items,
                         uint32_t n, uint32_t rays, uint2 *rec, hipStream_t stream) {
    uint64_t nrays = (uint64_t)n * rays;
    if (!nrays) return hipSuccess;
    uint64_t blocks = (nrays + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_record, dim3((unsigned)(blocks < (1u << 22) ? blocks : (1u << 22))),
                       dim3(BLOCK), 0, stream, S, P, items, n, rays, rec);
    return hipGetLastError();
}

hipError_t replay_gather(const uint2 *rec, uint32_t n, uint32_t rays, const float4 *albedo,
                         const float4 *light, const SkyP &K, float4 *out, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_replay, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, rec, n,
                       rays, albedo, light, K, out);
    return hipGetLastError();
}

namespace {

// One ring of filter.rs:58-94 on the masks alone: the ring (level) at which each texel gets a
// value and how many neighbours it averages. level -1: a value from the start.
__global__ void k_fill_plan(TexelView V, int8_t *level, uint8_t *count, int ring, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || level[g] != LEVEL_NEVER) return;
    Where p = where(V, g);
    uint32_t ci = V.chart[g];
    int no = 0, na = 0;
    int y0 = max((int)p.y - 1, 0), y1 = min((int)p.y + 2, (int)p.h);
    int x0 = max((int)p.x - 1, 0), x1 = min((int)p.x + 2, (int)p.w);
    for (int yy = y0; yy < y1; yy++)
        for (int xx = x0; xx < x1; xx++) {
            uint32_t j = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
            // had a value before this ring (a texel this ring fills reads as not yet)
            if (!(level[j] < ring)) continue;
            if (ci != NONE && V.chart[j] == ci) no++;
            else na++;
        }
    if (no > 0) {
        level[g] = (int8_t)ring;
        count[g] = (uint8_t)no;
    } else if (na > 0 && ci == NONE) {
        level[g] = (int8_t)ring;
        count[g] = (uint8_t)na;
    }
}

__global__ void k_level_init(const uint8_t *__restrict__ valid, int8_t *level, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) level[g] = valid[g] ? (int8_t)-1 : LEVEL_NEVER;
}

// The transpose of one fill ring: a texel the ring filled hands its adjoint to the neighbours it
// averaged, 1 / count each; every other texel keeps its own.
__global__ void k_fill_t(TexelView V, const int8_t *__restrict__ level,
                         const uint8_t *__restrict__ count, int ring,
                         const float4 *__restrict__ din, float4 *__restrict__ dout, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int lv = level[g];
    f3 acc = lv == ring ? mk(0.0f, 0.0f, 0.0f) : f3of(din[g]);
    if (lv < ring) {
        Where p = where(V, g);
        uint32_t cg = V.chart[g];
        int y0 = max((int)p.y - 1, 0), y1 = min((int)p.y + 2, (int)p.h);
        int x0 = max((int)p.x - 1, 0), x1 = min((int)p.x + 2, (int)p.w);
        for (int yy = y0; yy < y1; yy++)
            for (int xx = x0; xx < x1; xx++) {
                uint32_t i = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
                if (level[i] != ring) continue;
                uint32_t ci = V.chart[i];
                // a chart's texel averages only its chart's neighbours; an empty one any
                if (ci != NONE && ci != cg) continue;
                acc = add(acc, scl(f3of(din[i]), 1.0f / (float)count[i]));
            }
    }
    dout[g] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

// filter.rs:20-53's weight sums: texel g's Gaussian weights over its valid same-surface window
__global__ void k_wsum(TexelView V, const uint8_t *__restrict__ valid, int r,
                       const float *__restrict__ w, float *__restrict__ wsum, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float s = 1.0f;
    if (valid[g] && r > 0) {
        Where p = where(V, g);
        s = 0.0f;
        int y0 = max((int)p.y - r, 0), y1 = min((int)p.y + r + 1, (int)p.h);
        int x0 = max((int)p.x - r, 0), x1 = min((int)p.x + r + 1, (int)p.w);
        for (int yy = y0; yy < y1; yy++)
            for (int xx = x0; xx < x1; xx++) {
                uint32_t j = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
                if (!valid[j] || !same_surface(V, g, j)) continue;
                s += w[(yy - (int)p.y + r) * (2 * r + 1) + (xx - (int)p.x + r)];
            }
    }
    wsum[g] = s;
}

// The transpose of filter.rs:20-53: a valid texel j gets k(g, j) / wsum(g) of each valid
// same-surface g whose window holds it (windows and the surface test are symmetric); an invalid
// texel passed through unchanged.
__global__ void k_smooth_t(TexelView V, const uint8_t *__restrict__ valid, int r,
                           const float *__restrict__ w, const float *__restrict__ wsum,
                           const float4 *__restrict__ din, float4 *__restrict__ dout, uint32_t n) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (!valid[j] || r <= 0) {
        float4 v = din[j];
        dout[j] = make_float4(v.x, v.y, v.z, 0.0f);
        return;
    }
    Where p = where(V, j);
    f3 acc = mk(0.0f, 0.0f, 0.0f);
    int y0 = max((int)p.y - r, 0), y1 = min((int)p.y + r + 1, (int)p.h);
    int x0 = max((int)p.x - r, 0), x1 = min((int)p.x + r + 1, (int)p.w);
    for (int yy = y0; yy < y1; yy++)
        for (int xx = x0; xx < x1; xx++) {
            uint32_t g = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
            if (!valid[g] || !same_surface(V, g, j)) continue;
            float k = w[((int)p.y - yy + r) * (2 * r + 1) + ((int)p.x - xx + r)];
            acc = add(acc, scl(f3of(din[g]), k / wsum[g]));
        }
    dout[j] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

// The rounding error of a + b when s = fl(a + b) (Knuth's TwoSum: exact in round-to-nearest)
__device__ __forceinline__ float two_sum_err(float a, float b, float s) {
    float bb = s - a;
    return (a - (s - bb)) + (b - bb);
}

// Adds x to the compensated sum (hi, lo) at one address: the atomic returns the value it added
// to, so the addition's rounding error is known exactly and joins lo.
__device__ __forceinline__ void comp_add(float *hi, float *lo, float x) {
    if (x == 0.0f) return;
    float old = atomicAdd(hi, x);
    float e = two_sum_err(old, x, old + x);
    if (e != 0.0f) atomicAdd(lo, e);
}

// Neumaier's running sum in a register pair
struct Comp {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        float t = s + x;
        c += two_sum_err(s, x, t);
        s = t;
    }
};

// write_light's transpose (bake.rs:435-444): a known texel's direct and indirect each get the
// adjoint of its atlas texel (the pair hi + lo); `di` (when given) is set, `dd` accumulates.
__global__ void k_light_t(TexelView V, const float *__restrict__ dl_hi,
                          const float *__restrict__ dl_lo, uint32_t aw, float4 *__restrict__ di,
                          float4 *__restrict__ dd, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    f3 v = mk(0.0f, 0.0f, 0.0f);
    if (V.known[g] >= 1) {
        Where p = where(V, g);
        uint32_t corner = V.tiles[p.tile].w;
        size_t a = (size_t)((corner >> 16) + p.y) * aw + (corner & 0xffffu) + p.x;
        v = mk(dl_hi[a * 3] + dl_lo[a * 3], dl_hi[a * 3 + 1] + dl_lo[a * 3 + 1],
               dl_hi[a * 3 + 2] + dl_lo[a * 3 + 2]);
    }
    if (di) di[g] = make_float4(v.x, v.y, v.z, 0.0f);
    float4 o = dd[g];
    dd[g] = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, 0.0f);
}

// the loss's adjoint by the atlas as a compensated pair: hi its rgb, lo zero
__global__ void k_seed(const float4 *__restrict__ src, float *__restrict__ hi,
                       float *__restrict__ lo, uint32_t n) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    float4 v = src[a];
    hi[(size_t)a * 3] = v.x;
    hi[(size_t)a * 3 + 1] = v.y;
    hi[(size_t)a * 3 + 2] = v.z;
    lo[(size_t)a * 3] = lo[(size_t)a * 3 + 1] = lo[(size_t)a * 3 + 2] = 0.0f;
}

__global__ void k_zero_f32(float *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_zero_f4(float4 *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The gather's transpose for one bounce. Item g's result E = pi / rays * sum over its rays; with
// dE its adjoint, each ray hands on
//   sky:        pi / rays * dE * (y, 1 - y) to zenith and horizon (y >= 0), to ground (y < 0), and
//               pi / rays * dE * (0.282095, 0.488603 d) to the sky's L1 coefficients (c0..c3)
//   atlas hit:  dE * light / rays to its material's albedo, dE * albedo / rays to the texel
//   fallback:   dE * fallback / rays to the albedo, dE * albedo / rays * dfallback to zenith and
//               horizon (pi / 2 each: fallback = pi (zenith + horizon) / 2) and to c0 and c2 (the
//               L1 term's irradiance on an upward face)
// Parameter sums go to this block's row of `part` as (hi, lo) pairs, slots albedo rgb per
// material, then zenith, horizon, ground, c0..c3 rgb (width slots); the rows of several bounces
// add up. LDS holds the block's pairs; each item's sky and fallback terms are summed in
// registers first (Neumaier), the albedo and atlas terms go straight to compensated atomics.
constexpr uint32_t SKY_SLOTS = 21;  // zenith, horizon, ground, c0, c1, c2, c3 (rgb each)
__global__ __launch_bounds__(BLOCK) void k_gather_t(
    const uint2 *__restrict__ rec, uint32_t n, uint32_t rays, const uint32_t *__restrict__ gat,
    const float4 *__restrict__ de, const float4 *__restrict__ albedo,
    const float4 *__restrict__ light, SkyP K, float *__restrict__ dl_hi,
    float *__restrict__ dl_lo, float *__restrict__ part, uint32_t width) {
    extern __shared__ float acc[];  // width hi, then width lo
    for (uint32_t k = threadIdx.x; k < 2u * width; k += BLOCK) acc[k] = 0.0f;
    __syncthreads();
    float *ahi = acc, *alo = acc + width;
    uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
    if (g < n) {
        f3 dE = f3of(de[gat[g]]);
        float inv = 1.0f / (float)rays;
        float c = PI / (float)rays;
        Comp sk[SKY_SLOTS];
        uint32_t M3 = width - SKY_SLOTS;
        for (uint32_t r = 0; r < rays; r++) {
            uint2 q = rec[(size_t)r * n + g];
            uint32_t kind = q.y >> 30;
            if (kind == REC_BACK) continue;
            if (kind == REC_MISS) {
                float y = bitsf(q.x);
                float dv[3] = {dE.x * c, dE.y * c, dE.z * c};
                for (int k = 0; k < 3; k++) {
                    if (y < 0.0f) {
                        sk[6 + k].add(dv[k]);
                    } else {
                        sk[k].add(dv[k] * y);
                        sk[3 + k].add(dv[k] * (1.0f - y));
                    }
                }
                {  // the L1 coefficients' share (their derivative at any value, 0 included)
                    f3 d = gather_dir(K.items, g, r, rays, K.seed);
                    float basis[4] = {0.282095f, 0.488603f * d.x, 0.488603f * d.y,
                                      0.488603f * d.z};
                    for (int j = 0; j < 4; j++)
                        for (int k = 0; k < 3; k++) sk[9 + 3 * j + k].add(dv[k] * basis[j]);
                }
                continue;
            }
            uint32_t m = q.y & 0x3fffffffu;
            f3 a = f3of(albedo[m]);
            f3 da = scl(dE, inv);  // dE / rays
            bool lit = false;
            f3 ga;
            if (kind == REC_ATLAS) {
                float4 l = light[q.x];
                if (l.w > 0.5f) {
                    lit = true;
                    ga = mul(da, f3of(l));
                    f3 gl = mul(da, a);
                    comp_add(&dl_hi[(size_t)q.x * 3], &dl_lo[(size_t)q.x * 3], gl.x);
                    comp_add(&dl_hi[(size_t)q.x * 3 + 1], &dl_lo[(size_t)q.x * 3 + 1], gl.y);
                    comp_add(&dl_hi[(size_t)q.x * 3 + 2], &dl_lo[(size_t)q.x * 3 + 2], gl.z);
                }
            }
            if (!lit) {
                ga = mul(da, K.fallback);
                f3 w = mul(da, a);
                float wv[3] = {w.x, w.y, w.z};
                for (int k = 0; k < 3; k++) {
                    sk[k].add(wv[k] * (0.5f * PI));
                    sk[3 + k].add(wv[k] * (0.5f * PI));
                    sk[9 + k].add(wv[k] * SH_FALLBACK_C0);
                    sk[15 + k].add(wv[k] * SH_FALLBACK_C2);
                }
            }
            comp_add(&ahi[m * 3u], &alo[m * 3u], ga.x);
            comp_add(&ahi[m * 3u + 1u], &alo[m * 3u + 1u], ga.y);
            comp_add(&ahi[m * 3u + 2u], &alo[m * 3u + 2u], ga.z);
        }
        for (uint32_t k = 0; k < SKY_SLOTS; k++) {
            comp_add(&ahi[M3 + k], &alo[M3 + k], sk[k].s);
            comp_add(&ahi[M3 + k], &alo[M3 + k], sk[k].c);
        }
    }
    __syncthreads();
    // the block's pairs into its row (this kernel alone writes the row, bounce after bounce)
    float *row = part + (size_t)blockIdx.x * width * 2u;
    for (uint32_t k = threadIdx.x; k < width; k += BLOCK) {
        float h = row[2u * k], s = h + ahi[k];
        row[2u * k] = s;
        row[2u * k + 1u] += two_sum_err(h, ahi[k], s) + alo[k];
    }
}

// The sun's share: sum over texels of dD . G (G the filled cos x sun share), each block's three
// channels as (hi, lo) pairs in its row of `part`: a tree of compensated pair additions.
__global__ __launch_bounds__(BLOCK) void k_sun_t(const float4 *__restrict__ dd,
                                                 const float4 *__restrict__ G, uint32_t n,
                                                 float *__restrict__ part) {
    __shared__ float hs[3][BLOCK], ls[3][BLOCK];
    uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
    float v[3] = {0.0f, 0.0f, 0.0f};
    if (g < n) {
        float4 d = dd[g];
        float k = G[g].x;
        v[0] = d.x * k;
        v[1] = d.y * k;
        v[2] = d.z * k;
    }
    for (int c = 0; c < 3; c++) {
        hs[c][threadIdx.x] = v[c];
        ls[c][threadIdx.x] = 0.0f;
    }
    __syncthreads();
    for (uint32_t h = BLOCK / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int c = 0; c < 3; c++) {
                float a = hs[c][threadIdx.x], b = hs[c][threadIdx.x + h], s = a + b;
                hs[c][threadIdx.x] = s;
                ls[c][threadIdx.x] += ls[c][threadIdx.x + h] + two_sum_err(a, b, s);
            }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        part[((size_t)blockIdx.x * 3 + threadIdx.x) * 2] = hs[threadIdx.x][0];
        part[((size_t)blockIdx.x * 3 + threadIdx.x) * 2 + 1] = ls[threadIdx.x][0];
    }
}

// out[k] = the sum over rows of the pairs part[row][k], rows in order, compensated
__global__ void k_rows(const float *__restrict__ part, uint32_t rows, uint32_t width,
                       float *__restrict__ out) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= width) return;
    Comp s;
    for (uint32_t b = 0; b < rows; b++) {
        s.add(part[((size_t)b * width + k) * 2]);
        s.c += part[((size_t)b * width + k) * 2 + 1];
    }
    out[k] = s.s + s.c;
}

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

}  // namespace
}  // namespace lbk

struct lb_adjoint {
    lb_bake *B = nullptr;
    uint32_t ng = 0, rays = 0, bounces = 0, M = 0, T = 0;
    size_t atlas = 0;
    lbk::DevBuf rec, level, count, wsum, soft1, G, D, I, F4, E, m1, m2, albedo, ckpt, dl, dI, dI2,
        dD, part_g, part_s;
    lbk::SkyP sky{};
    uint32_t blocks_g = 0, blocks_s = 0;

    float4 *ck(uint32_t k) const { return (float4 *)ckpt.ptr + (size_t)k * atlas; }
};

namespace lbk {
namespace {

// filter.rs:58-94 over `rings` rings, 3 channels of float4 `v`, from `valid0`, masks in m1/m2
void fill3(lb_adjoint *A, const TexelView &V, float4 *v, uint32_t rings, const uint8_t *valid0,
           hipStream_t st) {
    uint8_t *ma = (uint8_t *)A->m1.ptr, *mb = (uint8_t *)A->m2.ptr;
    tp_copy_u8(valid0, ma, A->T, st);
    for (uint32_t r = 0; r < rings; r++) {
        tp_fill_ring(V, 3, (float *)v, ma, mb, st);
        std::swap(ma, mb);
    }
}

TexelView view_with(const lb_adjoint *A, float4 *direct, float4 *indirect) {
    TexelView V = A->B->V;
    V.direct = direct;
    V.indirect = indirect;
    return V;
}

int adjoint_create(lb_bake *B, lb_adjoint *A) {
    lb_gpu *gpu = B->gpu;
    hipStream_t st = gpu->stream;
    hipSetDevice(gpu->device);
    if (!B->scene) return fail("lb_adjoint_create: the bake is not prepared");
    if (B->sc.n_materials >= (1u << 30)) return fail("lb_adjoint_create: too many materials");
    A->B = B;
    A->ng = (uint32_t)B->ng;
    A->rays = B->P.counts[2];
    A->bounces = B->sc.gi_bounces;
    A->M = (uint32_t)B->sc.n_materials;
    A->T = (uint32_t)B->T;
    A->atlas = (size_t)B->aw * B->ah;
    if (B->grings >= (uint32_t)LEVEL_NEVER) return fail("lb_adjoint_create: %u fill rings", B->grings);
    uint32_t width = A->M * 3u + SKY_SLOTS;
    if ((size_t)width * 2u * sizeof(float) > 64 * 1024)
        return fail("lb_adjoint_create: %u materials exceed the LDS accumulators", A->M);
    uint64_t nrays = (uint64_t)A->ng * A->rays;
    if (nrays >= (1ull << 40)) return fail("lb_adjoint_create: %llu rays", (unsigned long long)nrays);
    size_t Tn = std::max<size_t>(A->T, 1), Gn = std::max<size_t>(A->ng, 1);
    if (A->rec.reserve(std::max<uint64_t>(nrays, 1) * 8) || A->level.reserve(Tn) ||
        A->count.reserve(Tn) || A->wsum.reserve(Tn * 4) || A->soft1.reserve(Tn * 4) ||
        A->G.reserve(Tn * 16) || A->D.reserve(Tn * 16) || A->I.reserve(Tn * 16) ||
        A->F4.reserve(Tn * 16) || A->E.reserve(Gn * 16) || A->m1.reserve(Tn) ||
        A->m2.reserve(Tn) || A->albedo.reserve(std::max<uint32_t>(A->M, 1) * 16) ||
        A->ckpt.reserve(A->atlas * 16 * (A->bounces + 1)) || A->dl.reserve(A->atlas * 24) ||
        A->dI.reserve(Tn * 16) || A->dI2.reserve(Tn * 16) || A->dD.reserve(Tn * 16))
        return 1;
    A->blocks_g = (uint32_t)((Gn + BLOCK - 1) / BLOCK);
    A->blocks_s = (uint32_t)((Tn + BLOCK - 1) / BLOCK);
    if (A->part_g.reserve((size_t)A->blocks_g * width * 8) || A->part_s.reserve((size_t)A->blocks_s * 24))
        return 1;  // (hi, lo) FP32 pairs: 8 B a slot
    const TexelView &V = B->V;

    // the records (the gather items are the bake's, resident since its execute)
    SceneView S = scene_view(B->scene, nullptr);
    if (nrays)
        hipLaunchKernelGGL(k_record, dim3((unsigned)std::min<uint64_t>((nrays + BLOCK - 1) / BLOCK, 1u << 22)), dim3(BLOCK), 0, st, S, B->P,
                           (const lb_texel *)B->d_gitems.ptr, A->ng, A->rays, (uint2 *)A->rec.ptr);

    // the bounce fill's plan and the bounce smoothing's weight sums (bake.rs:318-320)
    if (A->T) {
        hipLaunchKernelGGL(k_level_init, grid_for(A->T), dim3(BLOCK), 0, st,
                           (const uint8_t *)B->d_gmask.ptr, (int8_t *)A->level.ptr, A->T);
        for (uint32_t r = 0; r < B->grings; r++)
            hipLaunchKernelGGL(k_fill_plan, grid_for(A->T), dim3(BLOCK), 0, st, V,
                               (int8_t *)A->level.ptr, (uint8_t *)A->count.ptr, (int)r, A->T);
        hipLaunchKernelGGL(k_wsum, grid_for(A->T), dim3(BLOCK), 0, st, V,
                           (const uint8_t *)B->d_gmask.ptr, (int)B->rg, (const float *)B->d_wg.ptr,
                           (float *)A->wsum.ptr, A->T);
        // the smoothed sun share, and G: the direct light at unit irradiance (bake.rs:243-256)
        tp_smooth(V, 1, V.soft, (float *)A->soft1.ptr, V.valid, 1, (const float *)B->d_w1.ptr, st);
        TexelView VG = view_with(A, (float4 *)A->G.ptr, (float4 *)A->I.ptr);
        tp_direct_light(VG, (const float *)A->soft1.ptr, make_float3(1.0f, 1.0f, 1.0f),
                        make_float3(B->sun.x, B->sun.y, B->sun.z), st);
        fill3(A, VG, (float4 *)A->G.ptr, B->rings, V.valid, st);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : hip_fail(e, "lb_adjoint_create");
}

int adjoint_forward(lb_adjoint *A, const lb_light_params &p, float *light) {
    lb_bake *B = A->B;
    hipStream_t st = B->gpu->stream;
    hipSetDevice(B->gpu->device);
    if (p.n_materials != A->M)
        return fail("lb_adjoint_forward: %zu materials, the scene has %u", p.n_materials, A->M);
    if (A->M && !p.albedo) return fail("lb_adjoint_forward: no albedo");
    std::vector<float4> alb(std::max<uint32_t>(A->M, 1), make_float4(0, 0, 0, 0));
    for (uint32_t m = 0; m < A->M; m++)
        alb[m] = make_float4(p.albedo[m * 3], p.albedo[m * 3 + 1], p.albedo[m * 3 + 2], 0.0f);
    if (A->albedo.upload(alb.data(), alb.size() * 16, st)) return 1;
    const float PIF = 3.14159265358979323846f;
    SkyP &K = A->sky;
    K.zenith = f3{p.zenith[0], p.zenith[1], p.zenith[2]};
    K.horizon = f3{p.horizon[0], p.horizon[1], p.horizon[2]};
    K.ground = f3{p.ground[0], p.ground[1], p.ground[2]};
    // bake.rs:106-116: fallback = (zenith + horizon) * 0.5 * pi, in that order; then the sky's L1
    // term's irradiance on an upward face, as the bake's prepare adds it
    float fb[3];
    for (int c = 0; c < 3; c++) {
        fb[c] = (p.zenith[c] + p.horizon[c]) * 0.5f * PIF;
        fb[c] += SH_FALLBACK_C0 * p.sky_sh[c] + SH_FALLBACK_C2 * p.sky_sh[6 + c];
    }
    K.fallback = f3{fb[0], fb[1], fb[2]};
    K.sh = sky_sh_of(p.sky_sh);
    K.items = (const lb_texel *)B->d_gitems.ptr;
    std::memcpy(&K.seed, &B->P.reach[3], 4);
    const TexelView &V0 = B->V;
    TexelView V = view_with(A, (float4 *)A->D.ptr, (float4 *)A->I.ptr);
    uint32_t T = A->T;
    size_t bytes = A->atlas * 16;
    // the direct light (bake.rs:247-256) at this irradiance
    if (T) {
        tp_direct_light(V, (const float *)A->soft1.ptr, make_float3(p.sun[0], p.sun[1], p.sun[2]),
                        make_float3(B->sun.x, B->sun.y, B->sun.z), st);
        fill3(A, V, (float4 *)A->D.ptr, B->rings, V0.valid, st);
        tp_fill_f32((float *)A->I.ptr, 0.0f, T * 4u, st);
    }
    // the first atlas: the clean tiles' light, then every dirty texel's (bake.rs:276-278)
    hipError_t e;
    if (B->any_clean)
        e = hipMemcpyAsync(A->ck(0), B->d_light0.ptr, bytes, hipMemcpyDeviceToDevice, st);
    else
        e = hipMemsetAsync(A->ck(0), 0, bytes, st);
    if (e != hipSuccess) return hip_fail(e, "lb_adjoint_forward");
    tp_write_light(V, A->ck(0), B->aw, st);
    for (uint32_t k = 1; k <= A->bounces && A->ng; k++) {
        e = hipMemcpyAsync(A->ck(k), A->ck(k - 1), bytes, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return hip_fail(e, "lb_adjoint_forward");
        hipLaunchKernelGGL(k_replay, grid_for(A->ng), dim3(BLOCK), 0, st,
                           (const uint2 *)A->rec.ptr, A->ng, A->rays, (const float4 *)A->albedo.ptr,
                           A->ck(k - 1), K, (float4 *)A->E.ptr);
        // bake.rs:301-328 as the bake's bounce runs it
        tp_fill_f32((float *)A->I.ptr, 0.0f, T * 4u, st);
        tp_scatter((const float4 *)A->E.ptr, (const uint32_t *)B->d_gat.ptr, A->ng,
                   (float4 *)A->I.ptr, st);
        tp_smooth(V, 4, (const float *)A->I.ptr, (float *)A->F4.ptr,
                  (const uint8_t *)B->d_gmask.ptr, (int)B->rg, (const float *)B->d_wg.ptr, st);
        tp_copy_f32((const float *)A->F4.ptr, (float *)A->I.ptr, T * 4u, st);
        fill3(A, V, (float4 *)A->I.ptr, B->grings, (const uint8_t *)B->d_gmask.ptr, st);
        tp_write_light(V, A->ck(k), B->aw, st);
    }
    uint32_t last = A->ng ? A->bounces : 0;
    if (light) {
        e = hipMemcpyAsync(light, A->ck(last), bytes, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return hip_fail(e, "lb_adjoint_forward");
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : hip_fail(e, "lb_adjoint_forward");
}

int adjoint_backward(lb_adjoint *A, const float *dlight, float *grad) {
    lb_bake *B = A->B;
    hipStream_t st = B->gpu->stream;
    hipSetDevice(B->gpu->device);
    const TexelView &V = B->V;
    uint32_t T = A->T, width = A->M * 3u + SKY_SLOTS;
    float *dl_hi = (float *)A->dl.ptr, *dl_lo = dl_hi + A->atlas * 3;
    float4 *dI = (float4 *)A->dI.ptr, *dI2 = (float4 *)A->dI2.ptr, *dD = (float4 *)A->dD.ptr;
    size_t pg = (size_t)A->blocks_g * width * 2, ps = (size_t)A->blocks_s * 3 * 2;
    hipLaunchKernelGGL(k_zero_f32, grid_for(pg), dim3(BLOCK), 0, st, (float *)A->part_g.ptr, pg);
    if (T) hipLaunchKernelGGL(k_zero_f4, grid_for(T), dim3(BLOCK), 0, st, dD, T);
    hipLaunchKernelGGL(k_seed, grid_for(A->atlas), dim3(BLOCK), 0, st, (const float4 *)dlight,
                       dl_hi, dl_lo, (uint32_t)A->atlas);
    uint32_t last = A->ng ? A->bounces : 0;
    for (uint32_t k = last; k >= 1; k--) {
        if (T) {
            hipLaunchKernelGGL(k_light_t, grid_for(T), dim3(BLOCK), 0, st, V,
                               (const float *)dl_hi, (const float *)dl_lo, B->aw, dI, dD, T);
            for (int r = (int)B->grings - 1; r >= 0; r--) {
                hipLaunchKernelGGL(k_fill_t, grid_for(T), dim3(BLOCK), 0, st, V,
                                   (const int8_t *)A->level.ptr, (const uint8_t *)A->count.ptr, r,
                                   (const float4 *)dI, dI2, T);
                std::swap(dI, dI2);
            }
            hipLaunchKernelGGL(k_smooth_t, grid_for(T), dim3(BLOCK), 0, st, V,
                               (const uint8_t *)B->d_gmask.ptr, (int)B->rg,
                               (const float *)B->d_wg.ptr, (const float *)A->wsum.ptr,
                               (const float4 *)dI, dI2, T);
            std::swap(dI, dI2);
        }
        hipLaunchKernelGGL(k_zero_f32, grid_for(A->atlas * 6), dim3(BLOCK), 0, st, dl_hi,
                           A->atlas * 6);
        hipLaunchKernelGGL(k_gather_t, dim3(A->blocks_g), dim3(BLOCK),
                           width * 2u * sizeof(float), st, (const uint2 *)A->rec.ptr, A->ng,
                           A->rays, (const uint32_t *)B->d_gat.ptr, (const float4 *)dI,
                           (const float4 *)A->albedo.ptr, (const float4 *)A->ck(k - 1), A->sky,
                           dl_hi, dl_lo, (float *)A->part_g.ptr, width);
    }
    if (T) {
        hipLaunchKernelGGL(k_light_t, grid_for(T), dim3(BLOCK), 0, st, V, (const float *)dl_hi,
                           (const float *)dl_lo, B->aw, (float4 *)nullptr, dD, T);
        hipLaunchKernelGGL(k_sun_t, dim3(A->blocks_s), dim3(BLOCK), 0, st, (const float4 *)dD,
                           (const float4 *)A->G.ptr, T, (float *)A->part_s.ptr);
    } else {
        hipLaunchKernelGGL(k_zero_f32, grid_for(ps), dim3(BLOCK), 0, st, (float *)A->part_s.ptr, ps);
    }
    hipLaunchKernelGGL(k_rows, grid_for(width), dim3(BLOCK), 0, st, (const float *)A->part_g.ptr,
                       A->blocks_g, width, grad);
    hipLaunchKernelGGL(k_rows, dim3(1), dim3(BLOCK), 0, st, (const float *)A->part_s.ptr,
                       T ? A->blocks_s : 1u, 3u, grad + width);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : hip_fail(e, "lb_adjoint_backward");
}

}  // namespace
}  // namespace lbk

extern "C" {

int lb_adjoint_create(lb_bake *bake, lb_adjoint **out) {
    if (!bake || !out) return fail("lb_adjoint_create: null handle");
    *out = nullptr;
    auto *a = new lb_adjoint();
    if (lbk::adjoint_create(bake, a)) {
        delete a;
        return 1;
    }
    bake->refs.fetch_add(1);  // the bake lives until this adjoint is destroyed
    *out = a;
    return 0;
}

void lb_adjoint_destroy(lb_adjoint *adj) {
    if (!adj) return;
    lb_bake *B = adj->B;
    if (B && B->gpu) hipStreamSynchronize(B->gpu->stream);
    delete adj;
    lb_bake_destroy(B);  // drops the adjoint's reference
}

int lb_adjoint_info(const lb_adjoint *adj, size_t *atlas_texels, size_t *n_params, size_t *items,
                    size_t *rays) {
    if (!adj) return fail("lb_adjoint_info: null handle");
    if (atlas_texels) *atlas_texels = adj->atlas;
    if (n_params) *n_params = (size_t)adj->M * 3 + lbk::SKY_SLOTS + 3;
    if (items) *items = adj->ng;
    if (rays) *rays = adj->rays;
    return 0;
}

int lb_adjoint_records(const lb_adjoint *adj, uint32_t *out, size_t cap) {
    if (!adj || !out) return fail("lb_adjoint_records: null handle");
    size_t words = (size_t)adj->ng * adj->rays * 2;
    if (cap < words) return fail("lb_adjoint_records: %zu words, %zu needed", cap, words);
    if (!words) return 0;
    hipError_t e = hipMemcpy(out, adj->rec.ptr, words * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : hip_fail(e, "lb_adjoint_records");
}

int lb_adjoint_forward(lb_adjoint *adj, const lb_light_params *params, float *light) {
    if (!adj || !params) return fail("lb_adjoint_forward: null handle");
    return lbk::adjoint_forward(adj, *params, light);
}

int lb_adjoint_backward(lb_adjoint *adj, const float *dlight, float *grad) {
    if (!adj || !dlight || !grad) return fail("lb_adjoint_backward: null handle");
    return lbk::adjoint_backward(adj, dlight, grad);
}

}  // extern "C"
