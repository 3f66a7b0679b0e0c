// The texel passes between the ray passes, on the device: the direct pass's decode and AO tone
// (crates/lightbaker/src/bake.rs:207-241), the edge-aware smoothing, the ring fill and the sun's
// distance field (filter.rs:13-141), the direct light (bake.rs:247-252), the gathered light's
// scatter (bake.rs:309-316) and the light atlas write (bake.rs:435-444). In the reference they run
// on the CPU between dispatches; here they keep the bounce loop on the GPU so it can be captured
// in a graph.
//
// Layout: the dirty tiles' texels back to back (tile-major, row-major within a tile); a texel
// finds its tile through `tile_of`. One lane a texel. Every output is a pure function of the
// previous step's arrays in the reference's summation order, so the results are bit-exact with
// the CPU restatement. Gaussian weights come from a host table (expf is the host's).
#include "lb_device.h"
#include "lb_texel.h"

namespace lbk {

// bake.rs:207-241 for the covered item k (texel cov[k])
__global__ void k_decode(TexelView V, const float4 *__restrict__ res, const uint32_t *cov,
                         uint32_t n_cov, DecodeArgs a) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_cov) return;
    uint32_t g = cov[k];
    float4 r = res[k];
    if (r.w > a.inside_share) {
        V.valid[g] = 0;
        return;
    }
    V.soft[g] = r.x;
    uint32_t bits = fbits(r.y);
    V.hard[g] = bits & 0x00ffffffu;
    uint32_t pick = bits >> 24;
    V.state[g] = pick == 0 ? VALIDITY_CENTRE : (pick == 9 ? VALIDITY_PUSHED : VALIDITY_OFFSET);
    float4 *t = reinterpret_cast<float4 *>(V.texels + g);
    if (pick == 9) {
        float4 p = t[0], n = t[1];
        t[0] = make_float4(p.x + n.x * PUSH, p.y + n.y * PUSH, p.z + n.z * PUSH, p.w);
    } else if (pick > 0) {
        float cx, cy;
        offset(pick, &cx, &cy);
        float4 p = t[0], dx = t[2], dy = t[3];
        // bake.rs:237: pos += dx * cx + dy * cy (the offset summed first)
        t[0] = make_float4(p.x + (dx.x * cx + dy.x * cy), p.y + (dx.y * cx + dy.y * cy),
                           p.z + (dx.z * cx + dy.z * cy), p.w);
    }
    float v = vmax(1.0f - a.ao_strength * r.z, 0.0f);
    if (a.ao_contrast != 1.0f) v = powf(v, a.ao_contrast);
    V.ao[g] = vmax(v, a.ao_floor);
}

// filter.rs:20-53 over C channels (stride C), weights w[(dy + r) * (2r + 1) + dx + r]
template <int C>
__global__ void k_smooth(TexelView V, const float *__restrict__ in, float *__restrict__ out,
                         const uint8_t *__restrict__ valid, int r, const float *__restrict__ w,
                         uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float acc[C];
    for (int c = 0; c < C; c++) acc[c] = in[g * C + c];
    if (valid[g] && r > 0) {
        Where p = where(V, g);
        float sum[C];
        for (int c = 0; c < C; c++) sum[c] = 0.0f;
        float wsum = 0.0f;
        int y0 = max((int)p.y - r, 0), y1 = min((int)p.y + r + 1, (int)p.h);
        int x0 = max((int)p.x - r, 0), x1 = min((int)p.x + r + 1, (int)p.w);
        for (int yy = y0; yy < y1; yy++)
            for (int xx = x0; xx < x1; xx++) {
                uint32_t j = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
                if (!valid[j] || !same_surface(V, g, j)) continue;
                float k = w[(yy - (int)p.y + r) * (2 * r + 1) + (xx - (int)p.x + r)];
                for (int c = 0; c < C; c++) sum[c] += in[j * C + c] * k;
                wsum += k;
            }
        for (int c = 0; c < C; c++) acc[c] = sum[c] / wsum;
    }
    for (int c = 0; c < C; c++) out[g * C + c] = acc[c];
}

// filter.rs:106-141: distance from each texel's centre to the nearest sub-sample of the other
// state on the same surface, over the window the field can reach
// The nearest opposite sample is found on squared distances and rooted once: sqrt and the
// subtraction of `half` are monotone in FP32 as in exact arithmetic, so the minimum of
// sqrt(d2) - half over the samples is sqrt(min d2) - half, bit for bit. The sample offsets
// ((c % g) + 0.5) / g and ((c / g) + 0.5) / g are computed once a workgroup, in LDS.
__global__ void k_sun_field(TexelView V, const uint8_t *__restrict__ covered, uint32_t g_side,
                            float radius, uint32_t n) {
    __shared__ float offs[2][32];
    uint32_t gg = g_side * g_side;
    if (threadIdx.x < gg) {
        float gf = (float)g_side;
        offs[0][threadIdx.x] = ((float)(threadIdx.x % g_side) + 0.5f) / gf;
        offs[1][threadIdx.x] = ((float)(threadIdx.x / g_side) + 0.5f) / gf;
    }
    __syncthreads();
    uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= n) return;
    if (!covered[gi]) {
        V.field[gi] = -1.0f;
        return;
    }
    Where p = where(V, gi);
    int reach = (int)ceilf(radius) + 1;
    bool lit = (uint32_t)__popc(V.hard[gi]) * 2u >= gg;
    float gf = (float)g_side;
    float half = 0.5f / gf;
    int y0 = max((int)p.y - reach, 0), y1 = min((int)p.y + reach + 1, (int)p.h);
    int x0 = max((int)p.x - reach, 0), x1 = min((int)p.x + reach + 1, (int)p.w);
    float cxf = (float)p.x + 0.5f, cyf = (float)p.y + 0.5f;
    uint32_t mask = gg >= 32u ? 0xffffffffu : ((1u << gg) - 1u);
    float d2min = FAR;
    for (int yy = y0; yy < y1; yy++)
        for (int xx = x0; xx < x1; xx++) {
            uint32_t j = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
            // the samples of the opposite state, tested first: most neighbours have none, and
            // then their chart and normal are not read
            uint32_t opp = (lit ? ~V.hard[j] : V.hard[j]) & mask;
            if (!opp || !covered[j] || !same_surface(V, gi, j)) continue;
            while (opp) {
                uint32_t c = (uint32_t)__builtin_ctz(opp);
                opp &= opp - 1u;
                float px = (float)xx + offs[0][c] - cxf;
                float py = (float)yy + offs[1][c] - cyf;
                d2min = vmin(d2min, px * px + py * py);
            }
        }
    float best = d2min < FAR ? vmin(radius, sqrtf(d2min) - half) : radius;
    float d = vmin(vmax(best, 0.0f), radius);
    V.field[gi] = 0.5f + 0.5f * (lit ? d : -d) / radius;
}

// bake.rs:247-252
__global__ void k_direct_light(TexelView V, const float *__restrict__ soft, float3 irr, float3 sun,
                               uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float4 nr = V.normal[g];
    float c = vmax(nr.x * sun.x + nr.y * sun.y + nr.z * sun.z, 0.0f);
    float s = soft[g];
    V.direct[g] = make_float4(irr.x * c * s, irr.y * c * s, irr.z * c * s, 0.0f);
}

// One array of a fill: C channels at stride S (C = 0: no array), with the two running sums.
template <int C, int S>
struct FillArr {
    float *v;
    float own[C > 0 ? C : 1], any[C > 0 ? C : 1];
    __device__ __forceinline__ void clear() {
        for (int c = 0; c < C; c++) own[c] = any[c] = 0.0f;
    }
    __device__ __forceinline__ void add_own(uint32_t j) {
        for (int c = 0; c < C; c++) own[c] += v[j * S + c];
    }
    __device__ __forceinline__ void add_any(uint32_t j) {
        for (int c = 0; c < C; c++) any[c] += v[j * S + c];
    }
    __device__ __forceinline__ void put_own(uint32_t g, int no) {
        for (int c = 0; c < C; c++) v[g * S + c] = own[c] / (float)no;
    }
    __device__ __forceinline__ void put_any(uint32_t g, int na) {
        for (int c = 0; c < C; c++) v[g * S + c] = any[c] / (float)na;
    }
};

// filter.rs:58-94, one ring: texels without a value take the mean of their 3 x 3 neighbours that
// had one (their own chart's first; an empty texel any chart's). Up to three arrays that fill from
// the same mask (the direct phase's occlusion, sun field and light) take one pass over the
// neighbours, each channel summed in the same order as alone. Updated in place: a ring writes only
// texels without a value, and reads only texels that had one before it (`was`), whose values no
// ring changes; `now` out.
template <int C0, int S0, int C1, int S1, int C2, int S2>
__global__ void k_fill_ring(TexelView V, float *v0, float *v1, float *v2,
                            const uint8_t *__restrict__ was, uint8_t *__restrict__ now,
                            uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    uint8_t known = was[g];
    if (!known) {
        Where p = where(V, g);
        FillArr<C0, S0> a0;
        FillArr<C1, S1> a1;
        FillArr<C2, S2> a2;
        a0.v = v0;
        a1.v = v1;
        a2.v = v2;
        a0.clear();
        a1.clear();
        a2.clear();
        int no = 0, na = 0;
        uint32_t ci = V.chart[g];
        int y0 = max((int)p.y - 1, 0), y1 = min((int)p.y + 2, (int)p.h);
        int x0 = max((int)p.x - 1, 0), x1 = min((int)p.x + 2, (int)p.w);
        for (int yy = y0; yy < y1; yy++)
            for (int xx = x0; xx < x1; xx++) {
                uint32_t j = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
                if (!was[j]) continue;
                if (ci != NONE && V.chart[j] == ci) {
                    a0.add_own(j);
                    a1.add_own(j);
                    a2.add_own(j);
                    no++;
                } else {
                    a0.add_any(j);
                    a1.add_any(j);
                    a2.add_any(j);
                    na++;
                }
            }
        if (no > 0) {
            a0.put_own(g, no);
            a1.put_own(g, no);
            a2.put_own(g, no);
            known = 1;
        } else if (na > 0 && ci == NONE) {
            a0.put_any(g, na);
            a1.put_any(g, na);
            a2.put_any(g, na);
            known = 1;
        }
    }
    now[g] = known;
}

// One texel's ring step in k_fill_ring's arithmetic (one array): the mean of its 3 x 3 neighbours
// for which `was(j)` holds, its own chart's first; true when it took a value.
template <int C, int S, typename Was>
__device__ __forceinline__ bool fill_texel(const TexelView &V, float *v, uint32_t g, Was was) {
    Where p = where(V, g);
    FillArr<C, S> a;
    a.v = v;
    a.clear();
    int no = 0, na = 0;
    uint32_t ci = V.chart[g];
    int y0 = max((int)p.y - 1, 0), y1 = min((int)p.y + 2, (int)p.h);
    int x0 = max((int)p.x - 1, 0), x1 = min((int)p.x + 2, (int)p.w);
    for (int yy = y0; yy < y1; yy++)
        for (int xx = x0; xx < x1; xx++) {
            uint32_t j = p.off + (uint32_t)yy * p.w + (uint32_t)xx;
            if (!was(j)) continue;
            if (ci != NONE && V.chart[j] == ci) {
                a.add_own(j);
                no++;
            } else {
                a.add_any(j);
                na++;
            }
        }
    if (no > 0) {
        a.put_own(g, no);
        return true;
    }
    if (na > 0 && ci == NONE) {
        a.put_any(g, na);
        return true;
    }
    return false;
}

// Which texels a ring fills follows from the starting mask alone, not from the values, and every
// bounce fills the indirect light from the same mask (the gather set). So the first bounce's fill
// files each texel's ring (0: known before the first ring, r + 1: filled by ring r, RING_NEVER:
// never; ring 0 writes it for every texel, a later ring for the texels it fills), and the later
// bounces' rings visit only their own texels, reading the neighbours known before them from the
// filed rings: the same texels filled from the same neighbours in the same order, bit for bit.
template <int C, int S>
__global__ void k_fill_ring_file(TexelView V, float *v, const uint8_t *__restrict__ was,
                                 uint8_t *__restrict__ now, uint8_t *__restrict__ ring, uint32_t r,
                                 uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    uint8_t known = was[g];
    if (!known && fill_texel<C, S>(V, v, g, [&](uint32_t j) { return was[j] != 0; })) {
        known = 1;
        ring[g] = (uint8_t)(r + 1u);
    } else if (r == 0u) {
        ring[g] = known ? 0 : RING_NEVER;
    }
    now[g] = known;
}
template <int C, int S>
__global__ void k_fill_ring_filed(TexelView V, float *v, const uint8_t *__restrict__ ring,
                                  uint32_t r, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || ring[g] != r + 1u) return;
    fill_texel<C, S>(V, v, g, [&](uint32_t j) { return ring[j] <= r; });
}

// bake.rs:257-265: the state of covered texels filled from neighbours, and known = 2 own, 1 filled
__global__ void k_known(TexelView V, const uint8_t *__restrict__ filled, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    bool covered = V.texels[g].pos[3] > 0.5f;
    if (covered && V.state[g] == VALIDITY_NONE && filled[g]) V.state[g] = VALIDITY_FILLED;
    V.known[g] = V.valid[g] ? 2 : filled[g];
}

// bake.rs:309-316: the gathered items' irradiance at their texels, zero elsewhere
__global__ void k_scatter(const float4 *__restrict__ res, const uint32_t *__restrict__ idx,
                          uint32_t n_items, float4 *__restrict__ e) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_items) return;
    float4 r = res[k];
    e[idx[k]] = make_float4(r.x, r.y, r.z, 0.0f);
}

// bake.rs:435-444 for every texel of the dirty tiles
__global__ void k_write_light(TexelView V, float4 *__restrict__ light, uint32_t atlas_w,
                              uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    Where p = where(V, g);
    uint32_t corner = V.tiles[p.tile].w;
    uint32_t ax = (corner & 0xffffu) + p.x, ay = (corner >> 16) + p.y;
    float4 d = V.direct[g], e = V.indirect[g];
    light[ay * atlas_w + ax] = V.known[g] >= 1
                                   ? make_float4(d.x + e.x, d.y + e.y, d.z + e.z, 1.0f)
                                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// items[k] = texels[idx[k]]
__global__ void k_collect(const lb_texel *__restrict__ texels, const uint32_t *__restrict__ idx,
                          uint32_t n, lb_texel *__restrict__ items) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 *s = reinterpret_cast<const float4 *>(texels + idx[k]);
    float4 *d = reinterpret_cast<float4 *>(items + k);
    d[0] = s[0];
    d[1] = s[1];
    d[2] = s[2];
    d[3] = s[3];
}

__global__ void k_fill_f32(float *p, float v, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = v;
}

__global__ void k_copy_f32(const float *s, float *d, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) d[g] = s[g];
}

__global__ void k_copy_u8(const uint8_t *s, uint8_t *d, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) d[g] = s[g];
}

// bake.rs:286-297, per texel: valid texels on the downscale grid (a block's top-left texel), and
// the first valid texel (row-major) of each chart in a ds x ds block whose grid texel is not of
// that chart. A block's texels are all in its tile, so the test is local.
__global__ void k_gather_flags(TexelView V, uint32_t ds, uint8_t *__restrict__ flag, uint32_t n) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    uint8_t f = 0;
    if (V.valid[g]) {
        Where p = where(V, g);
        uint32_t x0 = p.x - p.x % ds, y0 = p.y - p.y % ds;
        if (p.x == x0 && p.y == y0) {
            f = 1;
        } else {
            uint32_t c = V.chart[g];
            uint32_t grid = p.off + y0 * p.w + x0;
            bool taken = V.valid[grid] && V.chart[grid] == c;
            uint32_t y1 = min(y0 + ds, p.h), x1 = min(x0 + ds, p.w);
            for (uint32_t yy = y0; yy < y1 && !taken; yy++)
                for (uint32_t xx = x0; xx < x1; xx++) {
                    if (yy > p.y || (yy == p.y && xx >= p.x)) break;
                    uint32_t j = p.off + yy * p.w + xx;
                    if (j != grid && V.valid[j] && V.chart[j] == c) {
                        taken = true;
                        break;
                    }
                }
            f = taken ? 0 : 1;
        }
    }
    flag[g] = f;
}

// Stream compaction of the flagged texels' indices, in order: per-block counts, one block's scan
// of them, then each block writes its texels at its offset.
constexpr uint32_t CBLOCK = 1024;

__global__ __launch_bounds__(1024) void k_count(const uint8_t *__restrict__ flag, uint32_t n,
                                                uint32_t *__restrict__ counts) {
    __shared__ uint32_t part[16];
    uint32_t g = blockIdx.x * CBLOCK + threadIdx.x;
    uint32_t v = (g < n && flag[g]) ? 1u : 0u;
    uint64_t ball = __ballot(v);
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = (uint32_t)__popcll(ball);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < CBLOCK / 64; w++) s += part[w];
        counts[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(1024) void k_scan_counts(uint32_t *__restrict__ counts, uint32_t nb,
                                                      uint32_t *__restrict__ total) {
    // exclusive scan of nb counts by one block, 1024 at a time
    __shared__ uint32_t buf[CBLOCK];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += CBLOCK) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? counts[i] : 0u;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < CBLOCK; off <<= 1) {
            uint32_t t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0u;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) counts[i] = carry + buf[threadIdx.x] - v;
        uint32_t last = buf[CBLOCK - 1];
        __syncthreads();
        carry += last;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void k_compact(const uint8_t *__restrict__ flag, uint32_t n,
                                                  const uint32_t *__restrict__ offsets,
                                                  uint32_t *__restrict__ out) {
    __shared__ uint32_t part[16];
    uint32_t g = blockIdx.x * CBLOCK + threadIdx.x;
    uint32_t v = (g < n && flag[g]) ? 1u : 0u;
    uint64_t ball = __ballot(v);
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = (uint32_t)__popcll(ball);
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += part[w];
    uint64_t below = lane ? (ball & ((~0ull) >> (64u - lane))) : 0ull;
    if (v) out[offsets[blockIdx.x] + before + (uint32_t)__popcll(below)] = g;
}

// ---- launchers ---------------------------------------------------------------------------------

static inline dim3 grid_for(uint32_t n) { return dim3((n + 255u) / 256u); }

void tp_decode(const TexelView &V, const float4 *res, const uint32_t *cov, uint32_t n_cov,
               const DecodeArgs &a, hipStream_t s) {
    if (n_cov) hipLaunchKernelGGL(k_decode, grid_for(n_cov), dim3(256), 0, s, V, res, cov, n_cov, a);
}
void tp_smooth(const TexelView &V, int channels, const float *in, float *out,
               const uint8_t *valid, int r, const float *w, hipStream_t s) {
    if (!V.n) return;
    if (channels == 1)
        hipLaunchKernelGGL(k_smooth<1>, grid_for(V.n), dim3(256), 0, s, V, in, out, valid, r, w, V.n);
    else
        hipLaunchKernelGGL(k_smooth<4>, grid_for(V.n), dim3(256), 0, s, V, in, out, valid, r, w, V.n);
}
void tp_sun_field(const TexelView &V, const uint8_t *covered, uint32_t g, float radius,
                  hipStream_t s) {
    if (V.n) hipLaunchKernelGGL(k_sun_field, grid_for(V.n), dim3(256), 0, s, V, covered, g, radius, V.n);
}
void tp_direct_light(const TexelView &V, const float *soft, float3 irr, float3 sun, hipStream_t s) {
    if (V.n) hipLaunchKernelGGL(k_direct_light, grid_for(V.n), dim3(256), 0, s, V, soft, irr, sun, V.n);
}
void tp_fill_ring(const TexelView &V, int channels, float *v, const uint8_t *was, uint8_t *now,
                  hipStream_t s) {
    if (!V.n) return;
    float *no_array = nullptr;
    if (channels == 1)
        hipLaunchKernelGGL((k_fill_ring<1, 1, 0, 1, 0, 1>), grid_for(V.n), dim3(256), 0, s, V, v,
                           no_array, no_array, was, now, V.n);
    else
        hipLaunchKernelGGL((k_fill_ring<3, 4, 0, 1, 0, 1>), grid_for(V.n), dim3(256), 0, s, V, v,
                           no_array, no_array, was, now, V.n);
}
void tp_fill_ring_file(const TexelView &V, float *v, const uint8_t *was, uint8_t *now,
                       uint8_t *ring, uint32_t r, hipStream_t s) {
    if (V.n)
        hipLaunchKernelGGL((k_fill_ring_file<3, 4>), grid_for(V.n), dim3(256), 0, s, V, v, was,
                           now, ring, r, V.n);
}
void tp_fill_ring_filed(const TexelView &V, float *v, const uint8_t *ring, uint32_t r,
                        hipStream_t s) {
    if (V.n)
        hipLaunchKernelGGL((k_fill_ring_filed<3, 4>), grid_for(V.n), dim3(256), 0, s, V, v, ring,
                           r, V.n);
}
void tp_fill_ring_direct(const TexelView &V, float *ao, float *field, float *direct,
                         const uint8_t *was, uint8_t *now, hipStream_t s) {
    if (V.n)
        hipLaunchKernelGGL((k_fill_ring<1, 1, 1, 1, 3, 4>), grid_for(V.n), dim3(256), 0, s, V, ao,
                           field, direct, was, now, V.n);
}
void tp_known(const TexelView &V, const uint8_t *filled, hipStream_t s) {
    if (V.n) hipLaunchKernelGGL(k_known, grid_for(V.n), dim3(256), 0, s, V, filled, V.n);
}
void tp_scatter(const float4 *res, const uint32_t *idx, uint32_t n_items, float4 *e, hipStream_t s) {
    if (n_items) hipLaunchKernelGGL(k_scatter, grid_for(n_items), dim3(256), 0, s, res, idx, n_items, e);
}
void tp_write_light(const TexelView &V, float4 *light, uint32_t atlas_w, hipStream_t s) {
    if (V.n) hipLaunchKernelGGL(k_write_light, grid_for(V.n), dim3(256), 0, s, V, light, atlas_w, V.n);
}
void tp_collect(const lb_texel *texels, const uint32_t *idx, uint32_t n, lb_texel *items,
                hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_collect, grid_for(n), dim3(256), 0, s, texels, idx, n, items);
}
void tp_fill_f32(float *p, float v, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_f32, grid_for(n), dim3(256), 0, s, p, v, n);
}
void tp_gather_set(const TexelView &V, uint32_t ds, uint8_t *flag, uint32_t *counts,
                   uint32_t *total, uint32_t *out, hipStream_t s) {
    uint32_t n = V.n;
    if (!n) {
        hipMemsetAsync(total, 0, 4, s);
        return;
    }
    uint32_t nb = (n + CBLOCK - 1) / CBLOCK;
    hipLaunchKernelGGL(k_gather_flags, grid_for(n), dim3(256), 0, s, V, ds, flag, n);
    hipLaunchKernelGGL(k_count, dim3(nb), dim3(CBLOCK), 0, s, flag, n, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(CBLOCK), 0, s, counts, nb, total);
    hipLaunchKernelGGL(k_compact, dim3(nb), dim3(CBLOCK), 0, s, flag, n, counts, out);
}
uint32_t tp_gather_blocks(uint32_t n) { return (n + CBLOCK - 1) / CBLOCK; }
void tp_compact(const uint8_t *flag, uint32_t n, uint32_t *counts, uint32_t *total, uint32_t *out,
                hipStream_t s) {
    uint32_t nb = (n + CBLOCK - 1) / CBLOCK;
    if (nb) hipLaunchKernelGGL(k_count, dim3(nb), dim3(CBLOCK), 0, s, flag, n, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(CBLOCK), 0, s, counts, nb, total);
    if (nb) hipLaunchKernelGGL(k_compact, dim3(nb), dim3(CBLOCK), 0, s, flag, n, counts, out);
}
void tp_copy_f32(const float *src, float *dst, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_copy_f32, grid_for(n), dim3(256), 0, s, src, dst, n);
}
void tp_copy_u8(const uint8_t *src, uint8_t *dst, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_copy_u8, grid_for(n), dim3(256), 0, s, src, dst, n);
}

}  // namespace lbk
