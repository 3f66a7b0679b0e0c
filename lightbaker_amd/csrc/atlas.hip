// The bake's tail on the device: every tile's TileLight (bake.rs:337-362), the atlas assembled
// from it (bake.rs:364-421) and the output codecs (encode.rs:7-69, output.rs:101,108,165-181).
// One lane a texel (TileLight), a smaller texel (atlas, RGB9E5), a 4 x 4 block (BC5) or a probe;
// every value is computed with the reference's FP32 operations in its order (no contraction), so
// the results equal the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include "lb_atlas.h"
#include "lb_bake.h"
#include "lb_host.h"

namespace lbk {
namespace {

constexpr uint32_t BLOCK = 256;
constexpr uint8_t VALIDITY_NONE = 0;    // output.rs:29-30
constexpr uint8_t VALIDITY_VOLUME = 5;  // output.rs:39-40

__device__ __forceinline__ bool isnan_(float v) { return v != v; }
// f32::max (a NaN operand is ignored)
__device__ __forceinline__ float max_rs(float a, float b) {
    return isnan_(a) ? b : (isnan_(b) ? a : (a > b ? a : b));
}
// f32::clamp (NaN stays NaN)
__device__ __forceinline__ float clamp_rs(float v, float lo, float hi) {
    v = v < lo ? lo : v;
    return v > hi ? hi : v;
}
// `as u32` of a float: NaN and negatives to 0, saturating above
__device__ __forceinline__ uint32_t f2u_rs(float v) {
    if (isnan_(v) || v <= 0.0f) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

// ---- TileLight ---------------------------------------------------------------------------------

// volume.rs:125-147
__device__ void volume_irradiance(const VolumeView &V, float3 p, float3 n, float e[3], float *sun) {
    float fx = clamp_rs((p.x - V.ox) / V.step, 0.0f, (float)(V.nx - 1));
    float fy = clamp_rs((p.y - V.oy) / V.step, 0.0f, (float)(V.ny - 1));
    float fz = clamp_rs((p.z - V.oz) / V.step, 0.0f, (float)(V.nz - 1));
    uint32_t x0 = min(f2u_rs(fx), V.nx - 2), y0 = min(f2u_rs(fy), V.ny - 2),
             z0 = min(f2u_rs(fz), V.nz - 2);
    float tx = fx - (float)x0, ty = fy - (float)y0, tz = fz - (float)z0;
    float p4[4][4] = {};
    for (int k = 0; k < 8; k++) {
        uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        float w = (dx ? tx : 1.0f - tx) * (dy ? ty : 1.0f - ty);
        w = w * (dz ? tz : 1.0f - tz);
        size_t q = (((size_t)(y0 + dy) * V.nz + (z0 + dz)) * V.nx + (x0 + dx)) * 4;
        for (int a = 0; a < 4; a++) {
            float4 r = V.probes[q + a];
            p4[a][0] += r.x * w;
            p4[a][1] += r.y * w;
            p4[a][2] += r.z * w;
            p4[a][3] += r.w * w;
        }
    }
    const float PIF = 3.14159265358979323846f;
    const float a0 = PIF * 0.282095f, a1 = 2.0f * PIF / 3.0f * 0.488603f;
    for (int c = 0; c < 3; c++) {
        float l1 = p4[1][c] * n.x + p4[2][c] * n.y;
        l1 = l1 + p4[3][c] * n.z;
        float lobe = a0 * p4[0][c] + a1 * l1;
        e[c] = 0.5f * a0 * p4[0][c] + 0.5f * max_rs(lobe, 0.0f);
    }
    *sun = p4[0][3];
}

__global__ __launch_bounds__(BLOCK) void k_tile_light(BakedView B, VolumeView V, TileLightView T) {
    for (uint32_t g = blockIdx.x * BLOCK + threadIdx.x; g < B.n; g += gridDim.x * BLOCK) {
        size_t dst = (size_t)((int64_t)g + B.dst_base[B.tile_of[g]]);
        uint8_t kn = B.known[g], st = B.state[g];
        float f = B.field[g], a = B.ao[g];
        float4 d = B.direct[g], e = B.indirect[g];
        if (V.probes && B.covered[g] && kn == 0) {
            // bake.rs:337-356: no light of its own, the volume's where it lies
            float4 p4 = B.pos[(size_t)g * 4];
            float4 n4 = B.normal[g];
            float3 n = make_float3(n4.x, n4.y, n4.z);
            float3 p = make_float3(p4.x + n.x * 0.15f, p4.y + n.y * 0.15f, p4.z + n.z * 0.15f);
            float ev[3], s;
            volume_irradiance(V, p, n, ev, &s);
            float dd = n.x * V.sun_dir.x + n.y * V.sun_dir.y;
            dd = dd + n.z * V.sun_dir.z;
            float c = max_rs(dd, 0.0f);
            e = make_float4(ev[0], ev[1], ev[2], 0.0f);
            d = make_float4(V.sun_irr.x * c * s, V.sun_irr.y * c * s, V.sun_irr.z * c * s, 0.0f);
            f = s > 0.5f ? 1.0f : 0.0f;
            a = 1.0f;
            kn = 1;
            st = VALIDITY_VOLUME;
        }
        T.known[dst] = kn;
        T.state[dst] = st;
        T.field[dst] = f;
        T.ao[dst] = a;
        T.direct[dst * 3] = d.x;
        T.direct[dst * 3 + 1] = d.y;
        T.direct[dst * 3 + 2] = d.z;
        T.indirect[dst * 3] = e.x;
        T.indirect[dst * 3 + 1] = e.y;
        T.indirect[dst * 3 + 2] = e.z;
    }
}

// The all-tile texel index of atlas texel (ax, ay), or ATLAS_NONE; its tile in *t.
__device__ __forceinline__ uint32_t texel_at(const AtlasMap &M, uint32_t ax, uint32_t ay,
                                             const AtlasTile **t) {
    uint32_t k = M.block_tile[(size_t)(ay >> 2) * M.bw + (ax >> 2)];
    if (k == ATLAS_NONE) return ATLAS_NONE;
    const AtlasTile &r = M.tiles[k];
    *t = &r;
    return r.first + (ay - r.y) * r.w + (ax - r.x);
}

// bake.rs:435-444 for the clean tiles; every other texel (0, 0, 0, 0)
__global__ __launch_bounds__(BLOCK) void k_clean_light(AtlasMap M, TileLightView T, float4 *light) {
    size_t n = (size_t)M.aw * M.ah;
    for (size_t g = (size_t)blockIdx.x * BLOCK + threadIdx.x; g < n; g += (size_t)gridDim.x * BLOCK) {
        uint32_t ax = (uint32_t)(g % M.aw), ay = (uint32_t)(g / M.aw);
        const AtlasTile *t = nullptr;
        uint32_t i = texel_at(M, ax, ay, &t);
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i != ATLAS_NONE && !t->dirty && T.known[i] >= 1) {
            size_t j = (size_t)i * 3;
            v = make_float4(T.direct[j] + T.indirect[j], T.direct[j + 1] + T.indirect[j + 1],
                            T.direct[j + 2] + T.indirect[j + 2], 1.0f);
        }
        light[g] = v;
    }
}

// ---- the atlas (bake.rs:364-421) ---------------------------------------------------------------

// Smaller texel (sx, sy): its DS x DS atlas texels' direct and validity, and the weighted mean
// of their indirect light. A smaller texel never spans two tiles (tiles are 4-aligned and DS
// divides 4), so the reference's order (tile, then row, then column) is the block's row-major
// order.
template <int DS>
__global__ __launch_bounds__(BLOCK) void k_atlas(AtlasMap M, TileLightView T, float2 *direct,
                                                 uint8_t *validity, float *ind, uint8_t *has,
                                                 uint32_t iw, uint32_t ih) {
    size_t n = (size_t)iw * ih;
    for (size_t g = (size_t)blockIdx.x * BLOCK + threadIdx.x; g < n; g += (size_t)gridDim.x * BLOCK) {
        uint32_t sx = (uint32_t)(g % iw), sy = (uint32_t)(g / iw);
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        for (int dy = 0; dy < DS; dy++) {
            uint32_t ay = sy * DS + dy;
            if (ay >= M.ah) break;
            for (int dx = 0; dx < DS; dx++) {
                uint32_t ax = sx * DS + dx;
                if (ax >= M.aw) break;
                size_t a = (size_t)ay * M.aw + ax;
                const AtlasTile *t = nullptr;
                uint32_t i = texel_at(M, ax, ay, &t);
                uint8_t kn = i != ATLAS_NONE ? T.known[i] : 0;
                if (kn == 0) {
                    direct[a] = make_float2(1.0f, 1.0f);
                    validity[a] = VALIDITY_NONE;
                    continue;
                }
                direct[a] = make_float2(T.field[i], T.ao[i]);
                validity[a] = T.state[i];
                float k = kn == 2 ? 1.0f : 1e-3f;
                size_t j = (size_t)i * 3;
                s0 += T.indirect[j] * k;
                s1 += T.indirect[j + 1] * k;
                s2 += T.indirect[j + 2] * k;
                s3 += k;
            }
        }
        bool h = s3 > 0.0f;
        ind[g * 3] = h ? s0 / s3 : 0.0f;
        ind[g * 3 + 1] = h ? s1 / s3 : 0.0f;
        ind[g * 3 + 2] = h ? s2 / s3 : 0.0f;
        has[g] = h;
    }
}

// One ring: an empty smaller texel takes the mean of the 3 x 3 neighbours that had light
__global__ __launch_bounds__(BLOCK) void k_atlas_ring(const float *src, const uint8_t *was,
                                                      float *dst, uint8_t *now, uint32_t iw,
                                                      uint32_t ih) {
    size_t n = (size_t)iw * ih;
    for (size_t g = (size_t)blockIdx.x * BLOCK + threadIdx.x; g < n; g += (size_t)gridDim.x * BLOCK) {
        if (was[g]) {
            dst[g * 3] = src[g * 3];
            dst[g * 3 + 1] = src[g * 3 + 1];
            dst[g * 3 + 2] = src[g * 3 + 2];
            now[g] = 1;
            continue;
        }
        uint32_t x = (uint32_t)(g % iw), y = (uint32_t)(g / iw);
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
        int c = 0;
        for (uint32_t yy = y > 0 ? y - 1 : 0; yy < min(y + 2, ih); yy++)
            for (uint32_t xx = x > 0 ? x - 1 : 0; xx < min(x + 2, iw); xx++) {
                size_t j = (size_t)yy * iw + xx;
                if (!was[j]) continue;
                a0 += src[j * 3];
                a1 += src[j * 3 + 1];
                a2 += src[j * 3 + 2];
                c++;
            }
        if (c > 0) {
            float fc = (float)c;
            dst[g * 3] = a0 / fc;
            dst[g * 3 + 1] = a1 / fc;
            dst[g * 3 + 2] = a2 / fc;
        } else {
            dst[g * 3] = src[g * 3];
            dst[g * 3 + 1] = src[g * 3 + 1];
            dst[g * 3 + 2] = src[g * 3 + 2];
        }
        now[g] = c > 0;
    }
}

// ---- codecs ------------------------------------------------------------------------------------

// output.rs:101 `q`, :167 `byte`: (v.clamp(0, 1) * 255).round() as u8
__device__ __forceinline__ uint32_t unorm8(float v) {
    float r = roundf(clamp_rs(v, 0.0f, 1.0f) * 255.0f);
    return min(f2u_rs(r), 255u);
}

// encode.rs:7-34: one channel's 4 x 4 block (texel k at v[k], row-major) -> 8 bytes
__device__ uint2 bc4_block(const uint32_t v[16]) {
    uint32_t hi = v[0], lo = v[0];
#pragma unroll
    for (int k = 1; k < 16; k++) {
        hi = max(hi, v[k]);
        lo = min(lo, v[k]);
    }
    uint2 out = make_uint2(hi | (lo << 8), 0u);
    if (hi == lo) return out;
    float fh = (float)hi, fl = (float)lo;
    float pal[8];
    pal[0] = fh;
    pal[1] = fl;
#pragma unroll
    for (int i = 2; i < 8; i++) pal[i] = ((float)(8 - i) * fh + (float)(i - 1) * fl) / 7.0f;
    uint64_t bits = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        float x = (float)v[k];
        uint32_t best = 0;
        float bd = fabsf(pal[0] - x);
#pragma unroll
        for (int i = 1; i < 8; i++) {
            float d = fabsf(pal[i] - x);
            if (d < bd) {
                best = i;
                bd = d;
            }
        }
        bits |= (uint64_t)best << (3 * k);
    }
    out.x |= (uint32_t)(bits << 16);
    out.y = (uint32_t)(bits >> 16);
    return out;
}

// encode.rs:37-50: a block a lane, 16 bytes (red block, green block), blocks row-major
template <bool FLOAT>
__global__ __launch_bounds__(BLOCK) void k_bc5(const void *rg, uint32_t w, uint32_t h,
                                               uint4 *out) {
    uint32_t bw = w / 4;
    size_t nb = (size_t)bw * (h / 4);
    for (size_t b = (size_t)blockIdx.x * BLOCK + threadIdx.x; b < nb; b += (size_t)gridDim.x * BLOCK) {
        uint32_t bx = (uint32_t)(b % bw) * 4, by = (uint32_t)(b / bw) * 4;
        uint32_t r[16], g[16];
#pragma unroll
        for (int y = 0; y < 4; y++) {
            size_t row = (size_t)(by + y) * w + bx;
            if (FLOAT) {
                const float4 *p = reinterpret_cast<const float4 *>(
                    static_cast<const float2 *>(rg) + row);
                float4 a = p[0], c = p[1];
                r[y * 4] = unorm8(a.x);
                g[y * 4] = unorm8(a.y);
                r[y * 4 + 1] = unorm8(a.z);
                g[y * 4 + 1] = unorm8(a.w);
                r[y * 4 + 2] = unorm8(c.x);
                g[y * 4 + 2] = unorm8(c.y);
                r[y * 4 + 3] = unorm8(c.z);
                g[y * 4 + 3] = unorm8(c.w);
            } else {
                uint2 q = *reinterpret_cast<const uint2 *>(static_cast<const uint8_t *>(rg) + row * 2);
#pragma unroll
                for (int x = 0; x < 4; x++) {
                    uint32_t word = x < 2 ? q.x : q.y;
                    uint32_t sh = (x & 1) * 16;
                    r[y * 4 + x] = (word >> sh) & 0xffu;
                    g[y * 4 + x] = (word >> (sh + 8)) & 0xffu;
                }
            }
        }
        uint2 a = bc4_block(r), c = bc4_block(g);
        out[b] = make_uint4(a.x, a.y, c.x, c.y);
    }
}

// encode.rs:53-69. floor(log2(m)) is taken from m's exponent field: the reference's log2 may
// round up to an integer just below a power of two, but then m / scale rounds to 512 and the
// carry (encode.rs:62-65) gives the same exponent, so the words agree.
__device__ uint32_t rgb9e5(float r, float g, float b) {
    const float MAX = 65408.0f;
    r = clamp_rs(r, 0.0f, MAX);
    g = clamp_rs(g, 0.0f, MAX);
    b = clamp_rs(b, 0.0f, MAX);
    float m = max_rs(max_rs(r, g), b);
    if (m <= 0.0f) return 0u;
    int32_t lg;
    if (isnan_(m))
        lg = 0;  // log2(NaN).floor() as i32
    else if (m < 1.52587890625e-05f)
        lg = -16;  // below 2^-16: clamped (encode.rs:60)
    else
        lg = (int32_t)((__float_as_uint(m) >> 23) & 0xffu) - 127;
    int32_t e = max(lg, -16) + 1 + 15;
    float scale = __uint_as_float((uint32_t)(e - 24 + 127) << 23);  // 2^(e - 15 - 9), exact
    if (f2u_rs(floorf(m / scale + 0.5f)) == 512u) {
        e += 1;
        scale *= 2.0f;
    }
    uint32_t qr = min(f2u_rs(floorf(r / scale + 0.5f)), 511u);
    uint32_t qg = min(f2u_rs(floorf(g / scale + 0.5f)), 511u);
    uint32_t qb = min(f2u_rs(floorf(b / scale + 0.5f)), 511u);
    return qr | (qg << 9) | (qb << 18) | ((uint32_t)e << 27);
}

__global__ __launch_bounds__(BLOCK) void k_rgb9e5(const float *rgb, size_t n, uint32_t *out) {
    for (size_t g = (size_t)blockIdx.x * BLOCK + threadIdx.x; g < n; g += (size_t)gridDim.x * BLOCK)
        out[g] = rgb9e5(rgb[g * 3], rgb[g * 3 + 1], rgb[g * 3 + 2]);
}

// output.rs:108: the brightest L0 channel (at least 1e-6). Positive floats order as their bits,
// so the maximum is an atomic max of the bits; NaN and anything below 1e-6 never win.
__global__ __launch_bounds__(BLOCK) void k_probe_scale(const float4 *probes, size_t n,
                                                       uint32_t *bits) {
    float m = 1e-6f;
    for (size_t g = (size_t)blockIdx.x * BLOCK + threadIdx.x; g < n; g += (size_t)gridDim.x * BLOCK) {
        float4 p = probes[g * 4];
        m = max_rs(max_rs(max_rs(m, p.x), p.y), p.z);
    }
    for (int o = 32; o > 0; o >>= 1) m = max_rs(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(bits, __float_as_uint(m));
}

// output.rs:165-181: 16 bytes a probe
__global__ __launch_bounds__(BLOCK) void k_volume_bytes(const float4 *probes, size_t n,
                                                        const uint32_t *scale_bits, uint4 *out) {
    const float SQRT3 = 1.7320508f;
    float scale = __uint_as_float(*scale_bits);
    for (size_t g = (size_t)blockIdx.x * BLOCK + threadIdx.x; g < n; g += (size_t)gridDim.x * BLOCK) {
        float4 l0 = probes[g * 4];
        float l0c[3] = {l0.x, l0.y, l0.z};
        uint32_t b[16];
        for (int c = 0; c < 3; c++) b[c] = unorm8(sqrtf(max_rs(l0c[c], 0.0f) / scale));
        b[3] = unorm8(l0.w);
        for (int axis = 1; axis < 4; axis++) {
            float4 l1 = probes[g * 4 + axis];
            float l1c[3] = {l1.x, l1.y, l1.z};
            for (int c = 0; c < 3; c++) {
                float r = l0c[c] > 1e-6f ? l1c[c] / (l0c[c] * SQRT3) : 0.0f;
                b[4 + (axis - 1) * 3 + c] = unorm8(r * 0.5f + 0.5f);
            }
        }
        b[13] = b[14] = b[15] = 0;
        uint4 o;
        o.x = b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24);
        o.y = b[4] | (b[5] << 8) | (b[6] << 16) | (b[7] << 24);
        o.z = b[8] | (b[9] << 8) | (b[10] << 16) | (b[11] << 24);
        o.w = b[12] | (b[13] << 8) | (b[14] << 16) | (b[15] << 24);
        out[g] = o;
    }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

// a grid for n work items of one lane each: at most 2^20 blocks, the kernels stride beyond
inline dim3 grid_of(size_t n) {
    size_t b = (n + BLOCK - 1) / BLOCK;
    return dim3((uint32_t)(b < (1u << 20) ? (b ? b : 1) : (1u << 20)));
}

}  // namespace

hipError_t atlas_tile_light(const BakedView &B, const VolumeView &V, const TileLightView &T,
                            hipStream_t s) {
    if (!B.n) return hipSuccess;
    hipLaunchKernelGGL(k_tile_light, grid_of(B.n), dim3(BLOCK), 0, s, B, V, T);
    return hipGetLastError();
}

hipError_t atlas_clean_light(const AtlasMap &M, const TileLightView &T, float4 *light,
                             hipStream_t s) {
    size_t n = (size_t)M.aw * M.ah;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_clean_light, grid_of(n), dim3(BLOCK), 0, s, M, T, light);
    return hipGetLastError();
}

hipError_t atlas_assemble(const AtlasMap &M, const TileLightView &T, uint32_t ds, float2 *direct,
                          uint8_t *validity, float *indirect, float *tmp, uint8_t *has0,
                          uint8_t *has1, hipStream_t s) {
    uint32_t iw = (M.aw + ds - 1) / ds, ih = (M.ah + ds - 1) / ds;
    size_t n = (size_t)iw * ih;
    if (!n) return hipSuccess;
    dim3 gr = grid_of(n);
    // the sums into `indirect`, ring 1 into `tmp`, ring 2 back into `indirect`
    if (ds == 1)
        hipLaunchKernelGGL(k_atlas<1>, gr, dim3(BLOCK), 0, s, M, T, direct, validity, indirect, has0, iw, ih);
    else if (ds == 2)
        hipLaunchKernelGGL(k_atlas<2>, gr, dim3(BLOCK), 0, s, M, T, direct, validity, indirect, has0, iw, ih);
    else if (ds == 4)
        hipLaunchKernelGGL(k_atlas<4>, gr, dim3(BLOCK), 0, s, M, T, direct, validity, indirect, has0, iw, ih);
    else
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_atlas_ring, gr, dim3(BLOCK), 0, s, indirect, has0, tmp, has1, iw, ih);
    hipLaunchKernelGGL(k_atlas_ring, gr, dim3(BLOCK), 0, s, tmp, has1, indirect, has0, iw, ih);
    return hipGetLastError();
}

hipError_t encode_bc5_f32(const float2 *rg, uint32_t w, uint32_t h, uint8_t *out, hipStream_t s) {
    if (w % 4 || h % 4) return hipErrorInvalidValue;
    size_t nb = (size_t)(w / 4) * (h / 4);
    if (!nb) return hipSuccess;
    hipLaunchKernelGGL(k_bc5<true>, grid_of(nb), dim3(BLOCK), 0, s, (const void *)rg, w, h,
                       reinterpret_cast<uint4 *>(out));
    return hipGetLastError();
}

hipError_t encode_bc5_u8(const uint8_t *rg, uint32_t w, uint32_t h, uint8_t *out, hipStream_t s) {
    if (w % 4 || h % 4) return hipErrorInvalidValue;
    size_t nb = (size_t)(w / 4) * (h / 4);
    if (!nb) return hipSuccess;
    hipLaunchKernelGGL(k_bc5<false>, grid_of(nb), dim3(BLOCK), 0, s, (const void *)rg, w, h,
                       reinterpret_cast<uint4 *>(out));
    return hipGetLastError();
}

hipError_t encode_rgb9e5(const float *rgb, size_t n, uint32_t *out, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_rgb9e5, grid_of(n), dim3(BLOCK), 0, s, rgb, n, out);
    return hipGetLastError();
}

hipError_t encode_volume(const float4 *probes, size_t n, uint8_t *out, uint32_t *scale_bits,
                         hipStream_t s) {
    // the scale (with no probe, the floor alone), then the bytes
    size_t b = (n + BLOCK - 1) / BLOCK;
    dim3 gr((uint32_t)(b < 1 ? 1 : (b < 1024 ? b : 1024)));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, scale_bits);
    hipLaunchKernelGGL(k_probe_scale, gr, dim3(BLOCK), 0, s, probes, n, scale_bits);
    if (n)
        hipLaunchKernelGGL(k_volume_bytes, grid_of(n), dim3(BLOCK), 0, s, probes, n,
                           (const uint32_t *)scale_bits, reinterpret_cast<uint4 *>(out));
    return hipGetLastError();
}

}  // namespace lbk

// ---- C-ABI: the codecs over device memory ------------------------------------------------------

extern "C" {

int lb_encode_bc5(lb_gpu *gpu, const uint8_t *rg, uint32_t w, uint32_t h, uint8_t *out) {
    if (!gpu) return lbk::fail("bc5: no gpu");
    if (w % 4 || h % 4) return lbk::fail("bc5: %u x %u is not a multiple of 4 x 4", w, h);
    if ((size_t)w * h && (!rg || !out)) return lbk::fail("bc5: null buffer");
    if (((uintptr_t)rg & 7) || ((uintptr_t)out & 15))
        return lbk::fail("bc5: rg must be 8-byte and out 16-byte aligned");
    hipError_t e = lbk::encode_bc5_u8(rg, w, h, out, gpu->stream);
    return e == hipSuccess ? 0 : lbk::hip_fail(e, "bc5");
}

int lb_encode_rgb9e5(lb_gpu *gpu, const float *rgb, size_t n, uint32_t *out) {
    if (!gpu) return lbk::fail("rgb9e5: no gpu");
    if (n && (!rgb || !out)) return lbk::fail("rgb9e5: null buffer");
    if (((uintptr_t)rgb | (uintptr_t)out) & 3)
        return lbk::fail("rgb9e5: rgb and out must be 4-byte aligned");
    hipError_t e = lbk::encode_rgb9e5(rgb, n, out, gpu->stream);
    return e == hipSuccess ? 0 : lbk::hip_fail(e, "rgb9e5");
}

int lb_encode_volume(lb_gpu *gpu, const float *probes, size_t n, uint8_t *out, float *scale) {
    if (!gpu) return lbk::fail("volume bytes: no gpu");
    if (n && (!probes || !out)) return lbk::fail("volume bytes: null buffer");
    if (!scale) return lbk::fail("volume bytes: null scale");
    if (((uintptr_t)probes | (uintptr_t)out) & 15)
        return lbk::fail("volume bytes: probes and out must be 16-byte aligned");
    hipStream_t s = gpu->stream;
    if (gpu->word.reserve(16)) return -1;
    uint32_t *bits = static_cast<uint32_t *>(gpu->word.ptr);
    hipError_t e = lbk::encode_volume(reinterpret_cast<const float4 *>(probes), n, out, bits, s);
    if (e == hipSuccess) e = hipMemcpyAsync(scale, bits, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : lbk::hip_fail(e, "volume bytes");
}

}  // extern "C"
