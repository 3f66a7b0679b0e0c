/*
 * ORACLE — test infrastructure, not product code: the CPU restatement that the HIP kernels are
 * checked against. Nothing in lightbaker_amd/ includes or links it.
 */
#ifndef LB_ORACLE_H
#define LB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/lightbaker.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_counters {
    uint64_t rays;
    uint64_t inner_visits;
    uint64_t tri_tests;
    uint64_t material_loads;
    uint64_t light_gathers;
    uint64_t leaf_visits;
} orc_counters;

void lb_sincos(float x, float *s, float *c);
uint32_t orc_pcg(uint32_t v);
float orc_rand(uint32_t a, uint32_t b, float seed);
float orc_radical(uint32_t i);
void orc_cosine(float u, float v, const float n[3], float out[3]);
int orc_sat_floor(float x, int hi);

int orc_run(int pass, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
            const float *materials, const float *charts, const lb_params *params,
            const lb_texel *items, size_t n, uint32_t key_base, const float *light, float *out,
            orc_counters *cnt);
int orc_run_quant(int pass, const lb_node *nodes, size_t n_nodes, const uint32_t *quant,
                  const lb_tri *tris, const float *materials, const float *charts,
                  const lb_params *params, const lb_texel *items, size_t n, uint32_t key_base,
                  const float *light, float *out, orc_counters *cnt);
int orc_quant_compare(int pass, const lb_node *nodes, size_t n_nodes, const uint32_t *quant,
                      const lb_tri *tris, const float *materials, const float *charts,
                      const lb_params *params, const lb_texel *items, size_t n,
                      uint32_t key_base, const float *light, float *out,
                      unsigned long long counts[3]);
void orc_set_sky_sh(const float *sh);
int orc_gather_dirs(const lb_params *params, const lb_texel *items, size_t n, uint32_t key_base,
                    float *out);
int orc_gather_records(const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                       const float *materials, const float *charts, const lb_params *params,
                       const lb_texel *items, size_t n, uint32_t key_base, uint32_t *out);
int orc_gather_terms(const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                     const float *materials, const float *charts, const lb_params *params,
                     const lb_texel *items, size_t n, uint32_t key_base, const float *light,
                     float *out);
int orc_gather_steps(const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                     const float *materials, const lb_params *params, const lb_texel *items,
                     size_t n, uint32_t key_base, uint8_t *out);
void orc_smooth64(int w, int h, const uint32_t *chart, const float *normal, const double *v,
                  int C, const uint8_t *valid, int r, double *out);
void orc_fill64(int w, int h, const uint32_t *chart, double *v, int C, uint8_t *valid, int rings);
void orc_trace(const lb_node *nodes, size_t n_nodes, const lb_tri *tris, const float *materials,
               const lb_params *params, const float o[3], const float d[3], float tmax, int any,
               uint32_t ray, float *t_out, uint32_t *tri_out, int *back_out, orc_counters *cnt);
int orc64_run(int pass, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
              const float *materials, const float *charts, const lb_params *params,
              const lb_texel *items, size_t n, uint32_t key_base, const float *light, double *out);
void orc_trace_brute(const lb_tri *tris, size_t n_tris, const float *materials,
                     const lb_params *params, const float o[3], const float d[3], float tmax,
                     uint32_t ray, float *t_out, uint32_t *tri_out, int *back_out);

/* the bake's tail and the output codecs (finish_ref.c) */
void orc_probe_fill(float *probes, size_t nx, size_t ny, size_t nz);
void orc_volume_irr(const float *probes, const uint32_t size[3], const float origin[3], float step,
                    const float p[3], const float n[3], float e[3], float *sun);
size_t orc_volume_lit(const float *probes, const uint32_t size[3], const float origin[3],
                      float step, const float sun_dir[3], const float sun_irr[3], size_t n,
                      const float *texels, uint8_t *known, uint8_t *state, float *field, float *ao,
                      float *direct, float *indirect);
void orc_atlas(const int32_t *tiles, size_t nt, const uint8_t *known, const uint8_t *state,
               const float *field, const float *ao, const float *indirect, uint32_t aw,
               uint32_t ah, uint32_t ds, float *direct_out, uint8_t *validity, float *ind_out);
uint8_t orc_unorm8(float v);
void orc_bc4_block(const uint8_t v[16], uint8_t out[8]);
int orc_bc5(const uint8_t *rg, size_t w, size_t h, uint8_t *out);
uint32_t orc_rgb9e5(const float c[3]);
void orc_rgb9e5_decode(uint32_t p, float out[3]);
float orc_volume_bytes(const float *probes, size_t n, uint8_t *out);
void orc_unorm8_n(const float *v, size_t n, uint8_t *out);
void orc_rgb9e5_n(const float *rgb, size_t n, uint32_t *out);
void orc_rgb9e5_decode_n(const uint32_t *p, size_t n, float *out);

/* deps.rs's triangle filter (deps_ref.c) */
int orc_swept(const float z[6], const float sun[3], float top, const float b[6]);
void orc_deps_bits(const float *pos, size_t n_tris, const uint8_t *casts, const float *zones,
                   size_t n_zones, const float sun[3], float top, uint32_t *bits);

#ifdef __cplusplus
}
#endif

#endif
