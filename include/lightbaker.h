/*
 * lightbaker_amd: the ray-cast bake of jer-nc/lightbaker on one AMD MI355X (gfx950).
 *
 * This header is the drop-in boundary. The first group of entry points replaces the wgpu
 * `Gpu` of the reference (crates/lightbaker/src/gpu/mod.rs); a Rust crate would bind it with
 * `extern "C"` declarations over the same #[repr(C)] records. Every record here is byte-identical
 * to the reference's Pod type it names (little-endian, no padding).
 *
 * Conventions: a return of 0 is success; any other value is an error whose text
 * lb_last_error() returns (thread-local, valid until the next call on the same thread).
 * Handles are owned by the library; input arrays are borrowed for the duration of a call;
 * output arrays are allocated by the caller. Calls on one lb_gpu are serialized by the caller,
 * as the reference's single-threaded driver does (bake.rs:151,302, volume.rs:77).
 *
 * Lifetimes: handles may be released in any order. A bake (lb_bake_*) holds a reference to its
 * lb_gpu and an adjoint holds one to its bake, so lb_close and lb_bake_destroy only drop the
 * caller's reference; the handle is freed with the last one. An lb_scene keeps only its device,
 * never the lb_gpu. Buffers from lb_buffer_alloc are freed through an open lb_gpu.
 */
#ifndef LIGHTBAKER_H
#define LIGHTBAKER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- records (byte-identical to the reference's Pod types) ---------------------------------- */

/* bvh.rs:12-22 `Node` (32 B): box; inner: first child (the second follows it); leaf: first
 * triangle and count (count 0: inner node). */
typedef struct lb_node {
    float min[3];
    uint32_t first;
    float max[3];
    uint32_t count;
} lb_node;

/* bvh.rs:24-33 `Tri` (48 B): v0 (w: material index as bits), e1 = v1 - v0 (w: chart index as
 * bits, 0xffffffff none), e2 = v2 - v0 (w: 0). */
typedef struct lb_tri {
    float v0[4];
    float e1[4];
    float e2[4];
} lb_tri;

/* raster.rs:13-22 `Texel` (64 B): pos (w: 1 covered), normal (w: chart index as bits; for a
 * volume probe the probe's index as bits), dx, dy (metres per texel along x and y). */
typedef struct lb_texel {
    float pos[4];
    float normal[4];
    float dx[4];
    float dy[4];
} lb_texel;

/* gpu/mod.rs:11-23 `Params` (144 B), filled as bake.rs:109-120 does. counts = (sun grid side g,
 * ao rays, gather or volume rays, items); reach = (ao reach, gather reach, lift, seed);
 * atlas = (width, height, first item, inside share as f32 bits). counts[3] and atlas[2] are
 * written by the library per launch, as Gpu::run does (gpu/mod.rs:153-154). */
typedef struct lb_params {
    float sun_dir[4];
    float sun_irr[4];
    float zenith[4];
    float horizon[4];
    float ground[4];
    float fallback[4];
    uint32_t counts[4];
    float reach[4];
    uint32_t atlas[4];
} lb_params;

/* gpu/mod.rs:25-33 `Pass`; the discriminants index the pipelines as in the reference. */
enum lb_pass { LB_PASS_DIRECT = 0, LB_PASS_GATHER = 1, LB_PASS_VOLUME = 2 };

/* Kernel form. PLAIN: one item per lane, the reference's serial ray loop, FP32 node records.
 * TUNED: rays spread over the lanes of a wave with Kahan-compensated wave reductions, an LDS
 * traversal stack, and the node/triangle layout chosen by lb_set_tuning. WATERTIGHT: PLAIN with
 * the watertight ray-triangle test of Woop, Benthin and Wald (2013) in place of the reference's
 * Moller-Trumbore (trace.wgsl:93-104): another hit on rays through shared edges, so it is not
 * bit-exact with the reference; it is there to measure how often the two tests differ. */
enum lb_form { LB_FORM_PLAIN = 0, LB_FORM_TUNED = 1, LB_FORM_WATERTIGHT = 2 };

/* Node layout for the TUNED form. WIDE: both children's FP32 boxes stored in the parent (64 B
 * per inner node visit instead of the reference's 96 B). QUANT: children's boxes as 8-bit
 * offsets in the parent's box, rounded outward (32 B per visit). */
enum lb_nodes { LB_NODES_WIDE = 0, LB_NODES_QUANT = 1 };

typedef struct lb_gpu lb_gpu;
typedef struct lb_scene lb_scene;

/* ---- the reference's Gpu ---------------------------------------------------------------------- */

/* Gpu::new (gpu/mod.rs:74-127): opens device 0 and loads the kernels. Errors: "no GPU adapter"
 * when no HIP device exists, "no GPU device" when it cannot be opened. */
int lb_open(lb_gpu **out);
/* Gpu.adapter (gpu/mod.rs:62,126): "<device name> (HIP <arch>, <n> CUs)", the device name
 * "AMD GPU <pci domain:bus:device>" where the runtime reports none. */
const char *lb_adapter_name(const lb_gpu *gpu);
void lb_close(lb_gpu *gpu);
const char *lb_last_error(void);

/* Gpu::scene (gpu/mod.rs:129-136): uploads the BVH, the materials (albedo rgb, opacity) and the
 * charts (two affine rows a chart). No charts becomes one zero chart, as the reference does.
 * Also derives the TUNED form's node layouts from the same BVH. */
int lb_scene_create(lb_gpu *gpu, const lb_node *nodes, size_t n_nodes, const lb_tri *tris,
                    size_t n_tris, const float *materials /* [n_materials][4] */,
                    size_t n_materials, const float *charts /* [n_charts][2][4] */,
                    size_t n_charts, lb_scene **out);
void lb_scene_destroy(lb_scene *scene);
/* No counterpart in the reference (its sky is zenith, horizon and ground, trace.wgsl:63-66): an
 * L1 spherical-harmonic term added to the sky the passes see where a ray escapes, per channel
 * 0.282095 c0 + 0.488603 ((c1 d.x + c2 d.y) + c3 d.z), the coefficients as the volume's probes
 * order them (L0, L1 x, y, z): sh[12] = c0 rgb, c1 rgb, c2 rgb, c3 rgb; NULL or all zero: the
 * reference's sky, bit for bit. The term is not clamped (the caller keeps it non-negative). The
 * fallback irradiance in lb_params (bounced() off the charts) is the caller's: the bake adds the
 * term's irradiance on an upward face, pi 0.282095 c0 + 2 pi / 3 0.488603 c2. */
int lb_scene_set_sky(lb_scene *scene, const float *sh);

/* Gpu::run (gpu/mod.rs:140-196): runs `pass` over `n` items (texels, or probes for the volume)
 * reading `light` (rgba per atlas texel; none becomes one zero texel), writes n * width vec4s in
 * item order to `out` (width 4 for the volume, else 1), and returns when they are there. Item i's
 * ray keys derive from i (its index in `items`), exactly as the reference's batched `first + i`.
 * Zero items is a no-op. `out_cap` is the number of vec4s `out` holds. */
int lb_run(lb_gpu *gpu, int pass, const lb_scene *scene, const lb_params *params,
           const lb_texel *items, size_t n, const float *light /* [n_light][4] */, size_t n_light,
           float *out /* [out_cap][4] */, size_t out_cap);

/* ---- device-resident extensions (no counterpart in the reference) -------------------------- */

/* Device buffers owned by the library, for callers that keep inputs resident in HBM. */
int lb_buffer_alloc(lb_gpu *gpu, size_t bytes, void **dev_out);
int lb_buffer_free(lb_gpu *gpu, void *dev);
int lb_upload(lb_gpu *gpu, void *dev_dst, const void *host_src, size_t bytes);
int lb_download(lb_gpu *gpu, void *host_dst, const void *dev_src, size_t bytes);
int lb_sync(lb_gpu *gpu);

/* As lb_run over device pointers, queued on the gpu's stream without waiting. `key_base` is added
 * to every item's index to form its ray key, so a slice of a longer item list keeps the keys of
 * the whole list (gpu/mod.rs:153-154 `atlas.z = first`). */
int lb_run_device(lb_gpu *gpu, int pass, int form, const lb_scene *scene, const lb_params *params,
                  const lb_texel *items_dev, size_t n, uint32_t key_base, const float *light_dev,
                  size_t n_light, float *out_dev);

/* No counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. As lb_run_device, but
 * the TUNED form's lock-step launch takes its work-queue heads and traversal spill area from a
 * static per-device scratch (1 KiB of queue heads, then 64 MiB of spill) instead of buffers
 * the library grows, and no call allocates or frees device memory. The caller still passes every
 * input and output. The PLAIN and WATERTIGHT forms use no scratch and run as lb_run_device.
 * Errors: the pooled schedule (its hits need memory sized by the launch), and a launch whose
 * spill area (resident lanes x stack entries beyond the LDS part x 4 B) exceeds the scratch;
 * 16 or more entries in LDS fit every lane a 256-CU device keeps resident. Every call on a device
 * shares that device's scratch: calls that overlap on one device (two lb_gpu handles' streams at
 * once) race. */
int lb_run_device_noalloc(lb_gpu *gpu, int pass, int form, const lb_scene *scene,
                          const lb_params *params, const lb_texel *items_dev, size_t n,
                          uint32_t key_base, const float *light_dev, size_t n_light,
                          float *out_dev);

/* Tuning of the TUNED form. */
typedef struct lb_tuning {
    int workgroup;      /* threads a workgroup: 64..1024, a multiple of 64 */
    int rays_per_lane;  /* rays a lane takes of one item (1..16): sets the lanes sharing an item */
    int lds_stack;      /* traversal stack entries in LDS (8..48); deeper ones spill to memory */
    int nodes;          /* LB_NODES_WIDE or LB_NODES_QUANT */
    int refill;         /* pooled schedule: idle lanes (1..64) a wave waits for before it
                           starts new rays */
    int schedule;       /* 0: lock-step: a lane traces each ray to the end, then its next;
                           1: pooled: lanes draw rays from the pass's pool as theirs end, the
                           hits are shaded in the reference's order (bit-exact sums) */
    int chunk;          /* pooled schedule: hits one chunk holds in HBM, 0 for 2^28 (2 GiB); a
                           launch runs ceil(rays / chunk) chunks */
    int leaf_batch;     /* lanes (1..64) standing at a leaf before a wave tests leaves rather
                           than stepping its other lanes through inner nodes (64: only once
                           none is at an inner node) */
    int waves;          /* lock-step schedule: waves a SIMD the kernel is compiled to keep
                           resident: 0 (the compiler's choice), 6 or 8 (fewer registers a lane,
                           some spilled) */
    int xcd_split;      /* lock-step schedule: 1: the items split into 8 contiguous ranges, one a
                           chiplet (XCD), each chiplet's waves taking their own range first, so a
                           chiplet's L2 holds the part of the scene its texels see; 0: one queue */
    int gather_order;   /* lock-step gather, the order an item's rays go to its lanes: 0: by
                           azimuth, later bounces of a bake ranked by the first bounce's steps; 1:
                           in elevation bands (32 rays a band, one an azimuth sector), later
                           bounces ranked; 2: elevation bands in every bounce, nothing counted or
                           ranked */
} lb_tuning;

/* Sets the knobs of all three passes; lb_get_tuning reads the gather pass's. */
int lb_set_tuning(lb_gpu *gpu, const lb_tuning *tuning);
int lb_get_tuning(const lb_gpu *gpu, lb_tuning *tuning);
/* The knobs of one pass (LB_PASS_*): each pass is its own kernel with its own best setting. */
int lb_set_pass_tuning(lb_gpu *gpu, int pass, const lb_tuning *tuning);
int lb_get_pass_tuning(const lb_gpu *gpu, int pass, lb_tuning *tuning);

/* No counterpart in the reference (its nodes are the 32 B records alone). The TUNED form's node
 * records lb_scene_create derives from a BVH, built on the host with no device: per node index,
 * the QUANT record (8 u32: for each axis the origin's FP32 bits, their low 9 bits zero, with
 * e + 151 in the low byte for the axis's step 2^e, so the word shifted left by 23 reads as 2^24
 * times the step; child a's reference; 12 corner bytes, child a's min xyz and max xyz then child
 * b's, each corner origin + byte * step; child b's reference) and the WIDE
 * record (16 f32: both children's boxes, then their references as bits). Leaves' records are
 * zero. A reference is an inner child's node index, or bit 31 with a leaf's count in bits 26..30
 * and first triangle in bits 0..25 (count 31: the leaf's node index instead). Either output may
 * be NULL; each holds n_nodes records. For callers that cache or inspect the records. */
int lb_tuned_nodes(const lb_node *nodes, size_t n_nodes, uint32_t *quant /* [n][8] */,
                   float *wide /* [n][16] */);

/* Traversal counters of the last lb_run_stats call: rays traced, inner-node visits (child-box
 * pairs tested), triangle tests, material loads, bounced() light gathers and leaf visits. */
typedef struct lb_stats {
    uint64_t rays;
    uint64_t inner_visits;
    uint64_t tri_tests;
    uint64_t material_loads;
    uint64_t light_gathers;
    uint64_t leaf_visits;
} lb_stats;

/* As lb_run_device with the traversal counted (a separate instantiation of the kernels; the
 * counting adds atomics, so it is not the timed form). Waits for completion. */
int lb_run_stats(lb_gpu *gpu, int pass, int form, const lb_scene *scene, const lb_params *params,
                 const lb_texel *items_dev, size_t n, uint32_t key_base, const float *light_dev,
                 size_t n_light, float *out_dev, lb_stats *stats);

/* Times the last `reps` launches queued by lb_run_device on the gpu's stream with HIP events:
 * lb_timer_begin records a start event, lb_timer_end a stop event and waits; returns ms. */
int lb_timer_begin(lb_gpu *gpu);
int lb_timer_end(lb_gpu *gpu, float *ms);

/* Selects the form lb_run uses (LB_FORM_PLAIN by default). */
int lb_set_form(lb_gpu *gpu, int form);

/* ---- host substrate (the reference's Rust host, restated in C++) ------------------------------ */

/* scene.rs:53-64 `Material` */
typedef struct lb_material_desc {
    float albedo[3];
    float opacity;
    uint32_t receives;
    uint32_t casts;
} lb_material_desc;

/* scene.rs:66-72 `MeshInfo` with its positions (triangles * 3 vertices * xyz, not indexed) */
typedef struct lb_mesh_desc {
    uint32_t material;
    uint32_t triangles;
    const float *positions;
} lb_mesh_desc;

/* A tile's own texel density: tiles whose ground-plan key (layout.rs:347-349: floor of x and z
 * over tileSize) is (x, z) lay their charts out at texels_per_meter. No counterpart in the
 * reference, whose layout has the one density (scene.rs:84). */
typedef struct lb_tile_density {
    int32_t x, z;
    float texels_per_meter;
} lb_tile_density;

/* scene.rs:17-128 `SceneHeader` (sun, sky, materials, meshes, lightmap, bake, volume) */
typedef struct lb_scene_desc {
    float sun_direction[3];
    float sun_radius_deg;
    float sun_irradiance[3];
    float zenith[3];
    float horizon[3];
    float ground[3];
    const lb_material_desc *materials;
    size_t n_materials;
    const lb_mesh_desc *meshes;
    size_t n_meshes;
    /* lightmap (scene.rs:85-97) */
    float texels_per_meter;
    float tile_size;
    float chart_angle_deg;
    uint32_t padding;
    uint32_t atlas_width;
    uint32_t indirect_downscale;
    /* bake (scene.rs:99-118) */
    uint32_t sun_samples;
    uint32_t ao_rays;
    float ao_reach;
    float ao_strength;
    float ao_floor;
    float ao_contrast;
    float sdf_radius;
    uint32_t gi_rays;
    uint32_t gi_bounces;
    float gi_reach;
    float inside_share;
    uint32_t gi_filter;
    /* volume (scene.rs:120-128); has_volume 0: none */
    uint32_t has_volume;
    float volume_min[3];
    float volume_max[3];
    float volume_step;
    uint32_t volume_rays;
    /* per-tile texel densities (n_tile_density entries, distinct keys, each positive); tiles not
     * listed take texels_per_meter. NULL / 0: the reference's layout, byte for byte. A tile's
     * density joins its dependency hash (lb_deps_tile_hashes) only when it is listed. */
    const lb_tile_density *tile_density;
    size_t n_tile_density;
    /* the sky's L1 term (lb_scene_set_sky): c0 rgb, c1 rgb, c2 rgb, c3 rgb; zero: the reference */
    float sky_sh[12];
} lb_scene_desc;

typedef struct lb_bvh lb_bvh;

/* Bvh::build (bvh.rs:50-106) over non-indexed triangles: positions [n_tris][3][3], each
 * triangle's material and chart (0xffffffff: none); triangles whose material does not cast are
 * left out, as are degenerate ones. */
int lb_bvh_build(const float *positions, size_t n_tris, const uint32_t *material,
                 const uint32_t *chart, const uint8_t *casts, size_t n_materials, lb_bvh **out);
/* As lb_bvh_build on up to `threads` host threads (0: the host's): the same nodes and
 * triangles whatever the count (subtrees are built apart and numbered as the single loop would). */
int lb_bvh_build_threads(const float *positions, size_t n_tris, const uint32_t *material,
                         const uint32_t *chart, const uint8_t *casts, size_t n_materials,
                         int threads, lb_bvh **out);
size_t lb_bvh_nodes(const lb_bvh *bvh, const lb_node **nodes);
size_t lb_bvh_tris(const lb_bvh *bvh, const lb_tri **tris);
void lb_bvh_destroy(lb_bvh *bvh);

/* Layout::build (layout.rs:89-115) of a scene, host only: the charts, the tiles and the atlas. */
typedef struct lb_layout lb_layout;
int lb_layout_build(const lb_scene_desc *scene, lb_layout **out);
void lb_layout_destroy(lb_layout *layout);
/* atlas width and height; the counts of charts and tiles */
int lb_layout_size(const lb_layout *layout, uint32_t *width, uint32_t *height, size_t *charts,
                   size_t *tiles);
/* per chart: its affine rows (layout.rs:77-85) [charts][8] and tile; per tile: key x, key z, w, h,
 * atlas x, atlas y [tiles][6]; per triangle (meshes in order): its chart. Any may be NULL. */
int lb_layout_get(const lb_layout *layout, float *affines, uint32_t *chart_tiles, int32_t *tiles,
                  uint32_t *tri_charts);
/* rasterize (raster.rs:39-100) tile `tile` into w * h texels */
int lb_layout_rasterize(const lb_scene_desc *scene, const lb_layout *layout, size_t tile,
                        lb_texel *out, size_t cap);

/* ---- the bake (bake.rs:94-432, volume.rs:34-148) ------------------------------------------- */

typedef struct lb_bake lb_bake;

/* What a rebake takes from the cache (bake.rs:123-135's `lights`, volume.rs:44-65's cached
 * chunks): which tiles and probes to bake, and for the others what an earlier bake left. The
 * per-texel arrays run over every texel of every tile of the layout, tiles in layout order,
 * row-major in a tile (the order of the LB_BAKE_TILE_* arrays); only clean tiles' entries are
 * read. The hashes that decide which tiles are clean (deps.rs) stay with the caller. */
typedef struct lb_bake_cache {
    const uint8_t *tile_dirty;  /* per tile: nonzero = bake it; NULL = every tile */
    /* bake.rs:61-69 `TileLight`, per texel */
    const uint8_t *known;       /* 2 own light, 1 filled from neighbours, 0 none */
    const uint8_t *state;       /* output.rs:29-40 validity */
    const float *field, *ao;    /* f32 */
    const float *direct, *indirect; /* f32 x3 */
    size_t n_texels;            /* the arrays' length: every texel of the layout's tiles */
    const uint8_t *probe_dirty; /* per probe: nonzero = trace it; NULL = every probe */
    const float *probes;        /* f32 x16 per probe as traced (before the inside fill) */
    size_t n_probes;
} lb_bake_cache;

typedef struct lb_bake_options {
    int form;      /* LB_FORM_* of the ray passes */
    int graph;     /* nonzero: the bounce loop (gather + texel passes) replayed from a hipGraph */
    int threads;   /* host threads of the BVH build (0: all, at most 16); the tiles are
                      rasterized on the device */
    const lb_bake_cache *cache; /* NULL: bake every tile and every probe */
    /* nonzero: the first bounce records what each gather ray meets and the later bounces replay
     * those records with the new light instead of tracing the same rays again (every bounce
     * traces the same rays: the same keys and sample points). The results are the same as
     * traced bounces: bit for bit in the PLAIN form, within the TUNED form's sum tolerance.
     * 8 B a gather ray of device memory. */
    int replay;
} lb_bake_options;

/* bake() (bake.rs:94): layout, BVH, the dirty tiles' texels, the direct pass, the texel passes,
 * the bounces, the volume, the atlas. The dirty tiles' items go into each pass as one list
 * (bake.rs:139-149: dirty tiles in layout order, each tile's texels row-major; an item's index
 * in it is its ray key); the clean tiles' light is in the atlas from the start. Errors as the
 * reference's: a tile wider than the atlas, a failing GPU call; and a cache whose arrays do not
 * match the layout. The scene and the cache are borrowed until the call returns. */
int lb_bake_run(lb_gpu *gpu, const lb_scene_desc *scene, const lb_bake_options *options,
                lb_bake **out);
void lb_bake_destroy(lb_bake *bake);

/* lb_bake_run in its three steps, for callers that time or repeat the device part. The scene is
 * borrowed from prepare until finish returns.
 *   prepare  host work before any ray (bake.rs:98-149): layout, BVH, texels, probes, uploaded;
 *   execute  every device pass on resident data (direct, texel passes, gather set, bounces,
 *            volume); repeatable: each call starts from the prepared state. With `per_pass`
 *            (3 entries: direct, gather over all bounces, volume) the counting kernels run and
 *            the traversal counters are returned;
 *   finish   after the rays: the probes' inside fill on the host (volume.rs:86-117), then on the
 *            device every tile's TileLight with the texels lit from the volume (bake.rs:337-362),
 *            the atlas, the indirect atlas and its fill (bake.rs:364-421) and the shipped
 *            encodings (output.rs:95-110); the direct, indirect and validity atlases are read
 *            back, every other array when lb_bake_array asks for it. */
int lb_bake_prepare(lb_gpu *gpu, const lb_scene_desc *scene, const lb_bake_options *options,
                    lb_bake **out);
int lb_bake_execute(lb_bake *bake, lb_stats *per_pass);
int lb_bake_finish(lb_bake *bake);
/* Mean device time of `reps` launches of one pass over the bake's own resident items (after an
 * execute), HIP events on the gpu's stream. */
int lb_bake_time_pass(lb_bake *bake, int pass, int reps, float *ms_avg);

typedef struct lb_bake_info {
    uint32_t atlas_width, atlas_height;        /* layout.rs:56-66 */
    uint32_t indirect_width, indirect_height;  /* bake.rs:365 */
    uint32_t tiles, tiles_baked, charts;       /* bake.rs:49-58 Stats */
    uint64_t texels, texels_baked, gather_items, probes, probes_baked;
    uint32_t volume_size[3];
    float volume_origin[3];
    float volume_step;
    uint32_t bvh_nodes, bvh_tris;
    /* device time of each stage (HIP events on the bake's stream), ms */
    float ms_direct, ms_direct_texel, ms_bounces, ms_volume;
    /* host time of each stage, s */
    double s_layout, s_bvh, s_raster, s_gather_set, s_execute, s_finish, s_total;
    float volume_scale;  /* output.rs:108: the brightest L0 channel, what a volume byte of 255 is */
    float ms_finish;     /* device time of finish's kernels, ms */
    /* bounce graphs captured so far: a graph is captured again when the gather item count, the
     * gather's tuning or a buffer its kernels read (the TUNED form's GPU-wide spill area and
     * pooled hits, which another launch may reallocate) changed since the last capture */
    uint32_t graph_captures;
} lb_bake_info;

int lb_bake_info_get(const lb_bake *bake, lb_bake_info *info);

enum lb_bake_array {
    LB_BAKE_DIRECT = 0,      /* f32 x2 per atlas texel: sun field, occlusion (bake.rs:38-39) */
    LB_BAKE_INDIRECT = 1,    /* f32 x3 per indirect texel (bake.rs:40-42) */
    LB_BAKE_VALIDITY = 2,    /* u8 per atlas texel (output.rs:29-40) */
    LB_BAKE_VOLUME = 3,      /* f32 x16 per probe, x fastest then z then y (volume.rs:25-31) */
    LB_BAKE_CHARTS = 4,      /* f32 x8 per chart: the affine rows (layout.rs:77-85) */
    LB_BAKE_TRI_CHARTS = 5,  /* u32 per triangle, meshes in order (layout.rs:64-65) */
    LB_BAKE_LIGHT = 6,       /* f32 x4 per atlas texel: the light the last bounce left */
    LB_BAKE_CHART_TILES = 7, /* u32 per chart: its tile */
    LB_BAKE_TILES = 8,       /* i32 x6 per tile: key x, key z, w, h, atlas x, atlas y */
    LB_BAKE_NODES = 9,       /* lb_node: the BVH the passes walk */
    LB_BAKE_TRIS = 10,       /* lb_tri */
    LB_BAKE_MATERIALS = 11,  /* f32 x4 per material */
    LB_BAKE_PARAMS = 12,     /* lb_params of the direct and gather passes */
    LB_BAKE_TEXELS = 13,     /* lb_texel per texel of every baked tile, tile after tile, as rasterized */
    LB_BAKE_GATHER_INDEX = 14, /* u32 per gather item: its baked texel (after an execute) */
    /* what bake.rs:360-362 puts in the cache, over every tile (clean tiles as the cache gave
     * them), texels in the lb_bake_cache order (after a finish) */
    LB_BAKE_TILE_KNOWN = 15,    /* u8 */
    LB_BAKE_TILE_STATE = 16,    /* u8 */
    LB_BAKE_TILE_FIELD = 17,    /* f32 */
    LB_BAKE_TILE_AO = 18,       /* f32 */
    LB_BAKE_TILE_DIRECT = 19,   /* f32 x3 */
    LB_BAKE_TILE_INDIRECT = 20, /* f32 x3 */
    LB_BAKE_PROBES = 21,        /* f32 x16 per probe before the inside fill (volume.rs:81-84) */
    LB_BAKE_BAKED_TILES = 22,   /* u32 per baked tile: its index in the layout */
    LB_BAKE_LIGHT0 = 23,        /* f32 x4 per atlas texel: the clean tiles' light, the rest 0 */
    /* the atlases as they ship (output.rs:99-110, after a finish) */
    LB_BAKE_DIRECT_BC5 = 24,       /* BC5 blocks of the direct atlas (encode.rs:37-50), w * h bytes */
    LB_BAKE_INDIRECT_RGB9E5 = 25,  /* u32 per indirect texel (encode.rs:53-69) */
    LB_BAKE_VOLUME_BYTES = 26,     /* 16 bytes per probe (output.rs:165-181) */
    /* the TUNED lock-step gather's ray order (after an execute that ranked it; else empty): u8
     * per gather item and ray, item after item. No counterpart in the reference, whose gather
     * traces each item's rays in index order. */
    LB_BAKE_GATHER_COST = 27,  /* each ray's traversal steps (inner visits + triangle tests, 255
                                * at most) in the first bounce, by azimuth slot */
    LB_BAKE_GATHER_ORDER = 28  /* the azimuth slots ordered by (cost, slot): the later bounces'
                                * slot s traces the ray of azimuth slot order[s] */
};

/* The bake's array `which`: its bytes, with *data pointing at them (owned by the bake, valid
 * until the next execute, finish or destroy). Arrays that stay on the device until asked for are
 * read back here; 0 when the array is empty or the read-back fails (lb_last_error says which). */
size_t lb_bake_array(const lb_bake *bake, int which, const void **data);

/* ---- the output codecs over device memory (encode.rs:7-69, output.rs:101,108,165-181) ------ *
 * Queued on the gpu's stream; lb_encode_volume waits for its scale. Buffers are device
 * pointers: `rg` 8-byte aligned, `out` of BC5 and of the volume 16-byte aligned. */

/* encode.rs:37-50 `bc5`: rg (w * h byte pairs, row-major) -> w * h bytes of blocks (each 4 x 4
 * block's red then green BC4 block, blocks row-major). w and h multiples of 4 (the reference
 * asserts it; here an error). */
int lb_encode_bc5(lb_gpu *gpu, const uint8_t *rg, uint32_t w, uint32_t h, uint8_t *out);
/* encode.rs:53-69 `rgb9e5` over n rgb triples (f32 x3) -> n words */
int lb_encode_rgb9e5(lb_gpu *gpu, const float *rgb, size_t n, uint32_t *out);
/* output.rs:108,165-181: n probes (f32 x16, volume.rs:25-31) -> 16 bytes each, and the scale
 * (host) the L0 bytes are relative to */
int lb_encode_volume(lb_gpu *gpu, const float *probes, size_t n, uint8_t *out, float *scale);

/* ---- what a tile's light depends on (deps.rs:72-121), the filter on the device ------------- *
 * The reference hashes, for every tile and every probe chunk, every casting triangle that can
 * shade it (its box within the zone, or between the zone and the sun) and rebakes the tiles whose
 * hash changed. It tests every triangle against every zone on the host; here one launch tests
 * every pair. Host buffers in and out; the hashes are BLAKE3 as the reference computes them. */

/* BLAKE3 (unkeyed, 32 bytes) of n bytes: the hash the reference keys its cache with. */
int lb_blake3(const void *data, size_t n, uint8_t out[32]);
/* deps.rs:83's filter for many zones: for zone z (n_zones x 6 f32: min xyz, max xyz, already
 * grown by the reach) and triangle t (positions: n_tris x 9 f32, meshes in order), bit t % 32 of
 * bits[z * ceil(n_tris / 32) + t / 32] is set when `casts` is NULL or casts[t] is nonzero, and
 * the triangle's box meets the zone swept along `sun` (normalized) up to `top` (deps.rs:104-121;
 * `top` is deps.rs:47-49's highest box top over every triangle). */
int lb_deps_swept(lb_gpu *gpu, const float *positions, size_t n_tris, const uint8_t *casts,
                  const float *zones, size_t n_zones, const float sun[3], float top,
                  uint32_t *bits);
/* bake.rs:124-128 over every tile of `layout` (NULL: built from the scene): each tile's
 * hash_reach (deps.rs:72-102) of its surfaces' box (bake.rs:447-458) grown by the reach
 * (occlusion's, or the bounces' when further), seeded with tile_seed (bake.rs:461-477), from
 * `global` (deps.rs:60-70's hash of the settings, the caller's). hashes: 32 bytes a tile. */
int lb_deps_tile_hashes(lb_gpu *gpu, const lb_scene_desc *scene, const lb_layout *layout,
                        const uint8_t global[32], uint8_t *hashes);
/* volume.rs:36-56: the probes grouped by ground-plan tile (keys in order), each group's
 * hash_reach with the bounces' reach, seeded with its key and `volume_json` (the volume settings
 * as serde_json writes them). keys (2 i32) and hashes (32 bytes) a chunk, up to cap; *n_chunks
 * the count (also when it exceeds cap, with an error). */
int lb_deps_probe_hashes(lb_gpu *gpu, const lb_scene_desc *scene, const uint8_t global[32],
                         const char *volume_json, size_t json_len, int32_t *keys,
                         uint8_t *hashes, size_t cap, size_t *n_chunks);

/* ---- the adjoint of the bounces ------------------------------------------------------------ *
 * No counterpart in the reference, which has no gradients; it serves callers that fit the
 * scene's light to a target lightmap (inverse lighting: albedos, sky and sun from a bake). The
 * light atlas after the bounces (bake.rs:300-330: the atlas the last bounce writes, direct +
 * indirect per known texel) is a function of the materials' albedo, the sky (trace.wgsl:63-66:
 * zenith, horizon, ground, plus the L1 spherical-harmonic term of lb_scene_set_sky; the fallback
 * irradiance pi (zenith + horizon) / 2 + pi 0.282095 c0 + 2 pi / 3 0.488603 c2 follows them)
 * and the sun's irradiance (bake.rs:247-252), with visibility held fixed: every gather ray's hit
 * is recorded once, and each forward replays those records through the bounces. Every sum the
 * backward takes is FP32 with exact compensation (no FP64 on the device). */

typedef struct lb_adjoint lb_adjoint;

typedef struct lb_light_params {
    const float *albedo;   /* rgb per material, n_materials x 3 (host) */
    size_t n_materials;    /* the scene's */
    float zenith[3], horizon[3], ground[3];
    float sun[3];          /* irradiance */
    float sky_sh[12];      /* the sky's L1 term (lb_scene_set_sky): c0 rgb, c1 rgb, c2 rgb, c3 rgb */
} lb_light_params;

/* After lb_bake_execute: records the gather rays' hits (the bake must outlive the adjoint).
 * Each record is two u32 per ray, ray r of gather item g at (r * items + g) * 2: a word (kind 0:
 * the atlas texel bounced() reads; kind 3: the bits of the direction's y; else 0) and a tag
 * (kind << 30 | material; kinds 0 a front hit on a chart, 1 a front hit off the charts, 2 a back
 * face, 3 the sky). */
int lb_adjoint_create(lb_bake *bake, lb_adjoint **out);
void lb_adjoint_destroy(lb_adjoint *adj);
/* atlas texels, parameters (n_materials * 3 + 24), gather items, rays per item */
int lb_adjoint_info(const lb_adjoint *adj, size_t *atlas_texels, size_t *n_params, size_t *items,
                    size_t *rays);
/* The records, copied to `out` (host, cap u32 words). */
int lb_adjoint_records(const lb_adjoint *adj, uint32_t *out, size_t cap);
/* The bounces replayed with `params`: the final light atlas into `light` (device memory,
 * atlas_texels x 4 f32). With the scene's own parameters it equals the bake's atlas bit for bit
 * in the PLAIN form. The bounces' atlases are kept for lb_adjoint_backward. */
int lb_adjoint_forward(lb_adjoint *adj, const lb_light_params *params, float *light);
/* The gradient of a loss through the last forward: `dlight` (device, atlas_texels x 4 f32, the
 * loss's derivative by each atlas texel's rgb; w unused) in, `grad` (device, n_params f32) out:
 * albedo rgb per material, then zenith, horizon, ground rgb, the sky's L1 coefficients c0..c3
 * rgb (as lb_light_params.sky_sh), and sun rgb. */
int lb_adjoint_backward(lb_adjoint *adj, const float *dlight, float *grad);

#ifdef __cplusplus
}
#endif

#endif /* LIGHTBAKER_H */
