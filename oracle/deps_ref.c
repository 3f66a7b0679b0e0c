/* ORACLE — test infrastructure, not product code.
 *
 * Scalar C restatement of deps.rs's triangle filter: the box of a triangle (deps.rs:40-52:
 * Aabb::empty, then its three corners) and `swept` (deps.rs:104-121: do the box and the zone
 * swept along the sun up to `top` overlap), in FP32 with the reference's operations and order;
 * f32::max / f32::min ignore a NaN operand.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "oracle.h"

static float mx(float a, float b) { return isnan(a) ? b : (isnan(b) ? a : (a > b ? a : b)); }
static float mn(float a, float b) { return isnan(a) ? b : (isnan(b) ? a : (a < b ? a : b)); }

/* deps.rs:104-121; z and b: min xyz, max xyz */
int orc_swept(const float z[6], const float sun[3], float top, const float b[6]) {
    float tmax = sun[1] > 1e-4f ? mx((top - z[1]) / sun[1], 0.0f) : 0.0f;
    float lo = 0.0f, hi = tmax;
    for (int a = 0; a < 3; a++) {
        float s = sun[a], zmin = z[a], zmax = z[3 + a], bmin = b[a], bmax = b[3 + a];
        if (fabsf(s) < 1e-6f) {
            if (zmin > bmax || zmax < bmin) return 0;
            continue;
        }
        float t1 = (bmax - zmin) / s, t2 = (bmin - zmax) / s;
        float a0 = s > 0.0f ? t2 : t1, a1 = s > 0.0f ? t1 : t2;
        lo = mx(lo, a0);
        hi = mn(hi, a1);
        if (lo > hi) return 0;
    }
    return 1;
}

/* every (zone, triangle) pair: bits as lb_deps_swept lays them out */
void orc_deps_bits(const float *pos, size_t n_tris, const uint8_t *casts, const float *zones,
                   size_t n_zones, const float sun[3], float top, uint32_t *bits) {
    size_t words = (n_tris + 31) / 32;
    for (size_t i = 0; i < n_zones * words; i++) bits[i] = 0;
    for (size_t t = 0; t < n_tris; t++) {
        if (casts && !casts[t]) continue;
        float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int v = 0; v < 3; v++)
            for (int a = 0; a < 3; a++) {
                b[a] = mn(b[a], pos[t * 9 + v * 3 + a]);
                b[3 + a] = mx(b[3 + a], pos[t * 9 + v * 3 + a]);
            }
        for (size_t z = 0; z < n_zones; z++)
            if (orc_swept(&zones[z * 6], sun, top, b)) bits[z * words + t / 32] |= 1u << (t % 32);
    }
}
