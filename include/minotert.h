/* minotert.h — C ABI of the MI355X-native MinoteRT ray-tracing hot path.
 *
 * The reference renders through one C++ render-stage class, Pathtracer
 * (src/gfx/modules/pathtracer.ixx:18-200), whose two stages are
 *   primaryRays(size, camera, prevCamera) -> GBuffer         (pathtracer.ixx:29)
 *   secondaryRays(gbuffer, camera, atmosphere, ...) -> image (pathtracer.ixx:118-120)
 * and whose GPU work is the compute shaders primaryRay.comp and secondaryRays.comp.
 * This header is the thin C layer under the C++ Pathtracer of include/minotert.hpp:
 * plain structs, uint32_t/float only, device pointers that the caller allocated with
 * mrt_alloc, status codes instead of exceptions, all work ordered on the context's stream.
 *
 * Conventions
 *   - Every call returns mrt_status; on failure mrt_last_error(ctx) describes it.
 *   - Pointers named *_dev are device pointers (mrt_alloc); host pointers are borrowed
 *     for the duration of the call only.
 *   - A context is externally synchronized: one host thread at a time.
 */
#ifndef MINOTERT_H
#define MINOTERT_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define MRT_ABI_VERSION 1u

typedef enum {
    MRT_OK = 0,
    MRT_ERR_INVALID = 1,     /* bad argument (null pointer, size, option) */
    MRT_ERR_HIP = 2,         /* a HIP runtime call failed */
    MRT_ERR_OOM = 3,         /* device allocation failed */
    MRT_ERR_UNSUPPORTED = 4, /* option not available in this build */
    MRT_ERR_STATE = 5        /* call not valid in the current state */
} mrt_status;

/* Ray and hit records (device layout, 32 B and 16 B).
 * A hit accepts t in [tmin, tmax]; the closest hit wins, the smaller triangle ID at equal t
 * (reference acceptance: strict t < best, primaryRay.comp:28). Miss: tri = 0xFFFFFFFF,
 * t = +inf, u = v = 0. u, v are the barycentric weights of v1 and v2. */
typedef struct { float o[3]; float tmin; float d[3]; float tmax; } mrt_ray;
typedef struct { float t; uint32_t tri; float u; float v; } mrt_hit;

/* Camera with the fields of the reference Camera (src/gfx/camera.ixx:8-22); radians.
 * Direction = (cos pitch cos yaw, cos pitch sin yaw, sin pitch), Z up (camera.ixx:26-38).
 * near_plane is only used for the reverse-Z depth of the G-buffer (math.ixx:848-860). */
typedef struct {
    float position[3]; float yaw; float pitch; float vertical_fov; float near_plane;
    uint32_t width; uint32_t height;
} mrt_camera;

typedef struct mrt_ctx mrt_ctx;
typedef struct mrt_scene mrt_scene;
typedef struct mrt_sky mrt_sky;     /* physical sky (see the sky section below) */

typedef struct { int device; uint32_t flags; } mrt_opts;

/* ---- context: replaces the Vulkan/vuk bring-up (src/sys/vulkan.ixx:106-115) ---- */
mrt_status mrt_create(mrt_ctx** out, const mrt_opts* opts);
mrt_status mrt_destroy(mrt_ctx* ctx);
const char* mrt_last_error(const mrt_ctx* ctx);   /* ctx may be NULL: last global error */
mrt_status mrt_sync(mrt_ctx* ctx);
uint32_t mrt_abi_version(void);
/* info[0] = compute units, info[1] = wavefront size, info[2] = LDS bytes per block,
 * info[3] = device index */
mrt_status mrt_device_info(mrt_ctx* ctx, uint32_t info[4]);
/* "<gcn arch>-<CUs>CU", e.g. "gfx950:sramecc+:xnack--256CU" (autotuner cache key) */
mrt_status mrt_device_name(mrt_ctx* ctx, char* buf, size_t len);

/* ---- device memory (the reference's transient vuk images, pathtracer.ixx:42-69) ---- */
mrt_status mrt_alloc(mrt_ctx* ctx, size_t bytes, void** out_dev);
mrt_status mrt_free(mrt_ctx* ctx, void* dev);
mrt_status mrt_upload(mrt_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
mrt_status mrt_download(mrt_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);
/* Stream-ordered device-to-device copy by a plain copy kernel ("copy_bytes"). */
mrt_status mrt_copy_device(mrt_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes);

/* ---- synthetic scenes: icosphere soups (no reference counterpart; replaces the
 * compile-time 5-sphere scene of src/gpu/scene.glsl:4-11) ---- */
typedef struct {
    uint32_t seed, n_spheres, level, ground;
    float box, r_min, r_max, alb_min, alb_max;
    float ground_z, ground_half, ground_albedo;
} mrt_scene_desc;
uint32_t mrt_scene_desc_tri_count(const mrt_scene_desc* desc);
/* verts: tri_count*9 floats, tri_material: tri_count, material_albedo: (n_spheres+ground)*3 */
mrt_status mrt_scene_generate(const mrt_scene_desc* desc, float* verts, uint32_t* tri_material,
                              float* material_albedo);
/* Float camera basis used by ray generation: pos, right, up, fwd (3 each), sx, sy. */
mrt_status mrt_camera_basis(const mrt_camera* cam, float out[14]);

/* ---- scene upload + BVH (no reference counterpart: the reference has no acceleration
 * structure; its scene is a brute-force loop, primaryRay.comp:26-32) ---- */
typedef struct {
    uint32_t max_leaf;  /* triangles per binary leaf, 1..8 (default 4) */
    uint32_t bins;      /* SAH bins, 4..64 (default 16) */
    uint32_t builder;   /* MRT_BVH_HOST (threaded host build) or MRT_BVH_DEVICE (on the GPU); both
                           produce the same bytes */
} mrt_bvh_opts;
enum { MRT_BVH_HOST = 0, MRT_BVH_DEVICE = 1 };
typedef struct {
    uint32_t n_tris, n_nodes, depth, leaves;
    double build_ms;
    uint64_t device_bytes;
} mrt_scene_stats;
/* Size limits of one scene or one batch. Traversal addresses a wide node by a 32-bit byte
 * offset (index * 128), so a BVH4 holds at most MRT_MAX_WIDE_NODES - 1 nodes; a leaf word holds
 * its first triangle in 28 bits. A BVH4 over n triangles has at most max(n, 1) nodes, so every
 * scene of at most MRT_MAX_WIDE_NODES - 1 triangles fits; larger ones fit if their tree does
 * (about 3.5 triangles per node with leaves of 4).
 * mrt_check_size: MRT_OK if a scene or batch of n_tris triangles whose BVH4 has n_nodes wide
 * nodes (0: not built yet, only the triangle limit applies) can be traced, else MRT_ERR_INVALID
 * with a message. mrt_scene_create and mrt_batch_create apply it and fail the same way, so an
 * over-limit scene is refused instead of being traced through wrapped node offsets. */
#define MRT_MAX_WIDE_NODES (1u << 25)
#define MRT_MAX_TRIANGLES ((1u << 28) - 1u)
mrt_status mrt_check_size(uint64_t n_tris, uint64_t n_nodes);
/* The host BVH4 build of mrt_scene_create (MRT_BVH_HOST; same bytes as MRT_BVH_DEVICE) into host
 * memory, with no device: for offline builds, serialization and inspection. nodes_host needs
 * room for max(n_tris, 1) 128-byte nodes (nodes_capacity_bytes), tris_host exactly n_tris
 * 48-byte leaf-ordered records; stats (optional) receives the node count, depth, leaf count and
 * build time (device_bytes = 0). */
mrt_status mrt_bvh_build_host(const float* verts, uint32_t n_tris, const mrt_bvh_opts* opts, void* nodes_host,
                              size_t nodes_capacity_bytes, void* tris_host, size_t tris_bytes, mrt_scene_stats* stats);
mrt_status mrt_scene_create(mrt_ctx* ctx, const float* verts, uint32_t n_tris,
                            const uint32_t* tri_material, const float* material_albedo,
                            uint32_t n_materials, const mrt_bvh_opts* opts, mrt_scene** out);
mrt_status mrt_scene_destroy(mrt_ctx* ctx, mrt_scene* scene);
mrt_status mrt_scene_stats_get(mrt_ctx* ctx, const mrt_scene* scene, mrt_scene_stats* out);
/* Copies the BVH4 out (serialization, inspection): n_nodes 128-byte wide nodes and n_tris
 * 48-byte leaf-ordered triangle records {v0, id bits}, {v1, 0}, {v2, 0}; either pointer may be
 * NULL, a non-NULL one needs exactly that many bytes. */
mrt_status mrt_scene_export(mrt_ctx* ctx, const mrt_scene* scene, void* nodes_host, size_t nodes_bytes, void* tris_host,
                            size_t tris_bytes);

/* ---- traversal: the closest-hit loop of tracePrimaryRay (primaryRay.comp:23-36) and of
 * traceSecondaryRays (secondaryRays.comp:77-83), over a BVH ---- */
mrt_status mrt_trace_closest(mrt_ctx* ctx, const mrt_scene* scene, const mrt_ray* rays_dev,
                             uint32_t n, mrt_hit* hits_dev);
/* Any hit in [tmin, tmax] (shadow rays): occluded_dev[i] = 1 or 0. */
mrt_status mrt_trace_any(mrt_ctx* ctx, const mrt_scene* scene, const mrt_ray* rays_dev,
                         uint32_t n, uint32_t* occluded_dev);

/* ---- ray reordering (no reference counterpart): a processing order that groups rays by
 * direction octant, then by the Morton code of the origin inside the scene bounds (stable
 * LSD radix sort on the device). Results of an ordered trace are still written at each
 * ray's own index, so they equal the unordered trace. ---- */
mrt_status mrt_ray_order(mrt_ctx* ctx, const mrt_scene* scene, const mrt_ray* rays_dev, uint32_t n,
                         uint32_t* order_dev);
/* The same order with no device allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. mrt_ray_order grows a context workspace on demand; this call uses instead a fixed
 * 26.2 MB device array that the library holds on each device it runs on until the process ends,
 * for at most MRT_RAY_ORDER_NOALLOC_MAX rays (more: MRT_ERR_INVALID). The order is written
 * straight into order_dev. Calls on one device share that array, so calls that overlap on one
 * device (two contexts, or work on two streams) race. */
#define MRT_RAY_ORDER_NOALLOC_MAX (1u << 21)
mrt_status mrt_ray_order_noalloc(mrt_ctx* ctx, const mrt_scene* scene, const mrt_ray* rays_dev, uint32_t n,
                                 uint32_t* order_dev);
mrt_status mrt_trace_closest_ordered(mrt_ctx* ctx, const mrt_scene* scene, const mrt_ray* rays_dev,
                                     const uint32_t* order_dev, uint32_t n, mrt_hit* hits_dev);
/* Stable radix sort of n (key, value) pairs by the low `bits` bits of the keys. */
mrt_status mrt_sort_pairs(mrt_ctx* ctx, uint32_t* keys_dev, uint32_t* values_dev, uint32_t n, uint32_t bits);

/* ---- render stages: Pathtracer::primaryRays / secondaryRays (pathtracer.ixx:29,118) ---- */

/* G-buffer of the primary pass (pathtracer.ixx:22-27, formats :42-69). hits_dev is required;
 * the other planes are optional (NULL = not written):
 *   depth_dev  W*H floats, reverse-Z z/w of the hit (primaryRay.comp:62-64), 0 on a miss
 *   normal_dev W*H*4 floats, two-sided face normal, or the ray direction on a miss
 *              (primaryRay.comp:66-71), w = 0
 *   motion_dev W*H*2 floats, (NDC_now - NDC_prev) / pitch in pixels (primaryRay.comp:73-75) */
typedef struct {
    mrt_hit* hits_dev;
    float* depth_dev;
    float* normal_dev;
    float* motion_dev;
    uint32_t width, height;
} mrt_gbuffer;

/* One closest-hit primary ray per pixel through the pixel centre (primaryRay.comp:46-51). */
mrt_status mrt_primary_rays(mrt_ctx* ctx, const mrt_scene* scene, const mrt_camera* camera,
                            const mrt_camera* prev_camera, mrt_gbuffer* gbuffer);

#define MRT_RENDER_SORT_RAYS 1u   /* sort extension rays by octant + origin Morton code */
#define MRT_RENDER_MEGAKERNEL 2u  /* after vertex 0, one launch carries every path to its end
                                     (one lane per path); same results as the wavefront. Without
                                     it the wavefront runs vertices 0 and 1 and that launch
                                     finishes the paths from vertex 2. */

/* Diffuse path tracing from the G-buffer (secondaryRays.comp:64-135), as wavefront kernels:
 * every pixel runs spp paths (samples sample_base .. sample_base+spp-1); path state carries
 * its RNG, seeded pcg(pixel ^ pcg(sample ^ pcg(frame_seed))). At every vertex that can still
 * spawn a ray, a shadow ray to the directional light adds thr*albedo*max(0,n.l)*sun_e if
 * unoccluded (sun_e = 0 turns next-event estimation off); escaped rays add thr*sky, or the
 * physical sky's radiance when sky_lut is set; roulette q = min(0.95, max(thr)) applies from
 * extension ray rr_start on (rr_start >= 2). rgb_dev (W*H*3 floats) receives the mean of
 * the spp samples. capture_bounce j > 0 stops before tracing extension ray j and exports
 * those rays (capture_rays_dev, capture_pid_dev: path id = pixel + sample*W*H, count in
 * capture_count_dev); rgb is then not written. */
typedef struct {
    uint32_t frame_seed, sample_base, spp, bounces, rr_start, flags;
    float sky[3], sun_dir[3], sun_e[3];
    uint32_t capture_bounce;
    mrt_ray* capture_rays_dev;
    uint32_t* capture_pid_dev;
    uint32_t* capture_count_dev;
    const mrt_sky* sky_lut;   /* non-NULL: escaped rays add thr * skyColor(dir) of this sky's
                                 current view instead of thr * sky (NEE still uses sun_e) */
} mrt_render_opts;
mrt_status mrt_secondary_rays(mrt_ctx* ctx, const mrt_scene* scene, const mrt_camera* camera,
                              const mrt_gbuffer* gbuffer, const mrt_render_opts* opts, float* rgb_dev);
/* primary + secondary with a context-owned G-buffer. */
mrt_status mrt_render(mrt_ctx* ctx, const mrt_scene* scene, const mrt_camera* camera,
                      const mrt_render_opts* opts, float* rgb_dev);
/* ---- batches (no reference counterpart): several scenes rendered in one frame. Their BVHs
 * are concatenated (child words rebased on the device); every bounce of every scene runs as
 * one persistent traversal launch over one shared ray queue. Triangle IDs stay local to each
 * scene and every scene's image equals its single-scene render bit for bit. ---- */
typedef struct mrt_batch mrt_batch;
typedef struct mrt_graph mrt_graph;
typedef struct { mrt_camera camera; float* rgb_dev; } mrt_batch_view;  /* rgb: W*H*3 floats */
/* scenes: n pointers (1..24); the scenes may be destroyed after the call */
mrt_status mrt_batch_create(mrt_ctx* ctx, const mrt_scene* const* scenes, uint32_t n, mrt_batch** out);
mrt_status mrt_batch_destroy(mrt_ctx* ctx, mrt_batch* batch);
/* info[0] scenes, [1] triangles, [2] BVH nodes, [3] device bytes */
mrt_status mrt_batch_info(mrt_ctx* ctx, const mrt_batch* batch, uint64_t info[4]);
/* One frame of every scene: primary rays traced inside the wavefront, then opts->bounces
 * diffuse bounces (same estimator as mrt_secondary_rays; capture is not available). */
mrt_status mrt_batch_render(mrt_ctx* ctx, mrt_batch* batch, const mrt_batch_view* views, uint32_t n_views,
                            const mrt_render_opts* opts);
/* The same frame captured once into a HIP graph. Each mrt_graph_launch replays it with the
 * frame index advanced by one (paths use frame_seed ^ frame index); mrt_graph_set_frame
 * sets the index of the next replay. The graph holds no memset nodes. */
mrt_status mrt_graph_create(mrt_ctx* ctx, mrt_batch* batch, const mrt_batch_view* views, uint32_t n_views,
                            const mrt_render_opts* opts, mrt_graph** out);
mrt_status mrt_graph_set_frame(mrt_ctx* ctx, mrt_graph* graph, uint32_t frame);
mrt_status mrt_graph_launch(mrt_ctx* ctx, mrt_graph* graph);
mrt_status mrt_graph_destroy(mrt_ctx* ctx, mrt_graph* graph);

/* Config-0 shading from a G-buffer: rgb = albedo * max(0, n.l) * light_e if the shadow ray
 * towards light_dir is unoccluded, 0 on a miss or in shadow. */
mrt_status mrt_shade_lambert(mrt_ctx* ctx, const mrt_scene* scene, const mrt_camera* camera,
                             const mrt_gbuffer* gbuffer, const float light_dir[3],
                             const float light_e[3], float* rgb_dev);

/* ---- traversal launch shapes: the autotuner's knobs. Every variant returns identical
 * results; they differ in block size, LDS stack depth and lane-refill policy. ---- */
uint32_t mrt_trace_variant_count(void);
const char* mrt_trace_variant_name(uint32_t v);
/* any = 0: closest-hit kernel, any = 1: any-hit kernel */
mrt_status mrt_trace_variant_set(mrt_ctx* ctx, int any, uint32_t v);
mrt_status mrt_trace_variant_get(mrt_ctx* ctx, int any, uint32_t* v);

/* ---- frame launch shapes: the autotuner's other knobs (no reference counterpart; the reference
 * dispatches fixed 8x8 workgroups, pathtracer.ixx:106): the shading block, which is also the
 * queue-append chunk (one atomic per block), and the megakernel's block size and LDS stack depth.
 * Every variant gives identical images. Variant 0 is the default. ---- */
uint32_t mrt_frame_variant_count(void);
const char* mrt_frame_variant_name(uint32_t v);   /* "s<shade block>_m<mega block>x<stack depth>" */
mrt_status mrt_frame_variant_set(mrt_ctx* ctx, uint32_t v);
mrt_status mrt_frame_variant_get(mrt_ctx* ctx, uint32_t* v);

/* ---- radiance cache (no reference counterpart): multiresolution hash grid (Mueller et al.
 * 2022) + six 64x64 layers (ReLU after the first five, RGB = first three outputs of the last),
 * BF16 MFMA forward/backward with FP32 accumulation, relative L2 loss, Adam. ---- */
typedef struct mrt_cache mrt_cache;
typedef struct {
    uint32_t levels, log2_table;   /* e.g. 16 levels, 2^19 entries x 2 features per level */
    float base_res, max_res;       /* coarsest / finest grid resolution over the bounds */
    float lo[3], hi[3];            /* positions are normalized to these bounds */
    float lr, adam_beta1, adam_beta2, adam_eps;
    uint32_t seed;                 /* weight and table initialization */
    float ema_decay;               /* > 0: queries use the exponential average of the trained
                                      parameters (decay a, bias-corrected); 0: the parameters.
                                      A frame's train_steps steps in mrt_render_cached apply the
                                      table's average in closed form (equal to per-step averaging
                                      in exact arithmetic); mrt_cache_train averages per step */
    float table_lr;                /* > 0: Adam step size of the hash table; 0: lr */
} mrt_cache_opts;
/* One cache query / training sample (64 B): hit position, incoming direction, normal, albedo. */
typedef struct { float pos[3]; float pad0; float dir[3]; float pad1; float nrm[3]; float pad2; float alb[3]; float pad3; }
    mrt_cache_sample;
mrt_status mrt_cache_create(mrt_ctx* ctx, const mrt_cache_opts* opts, mrt_cache** out);
mrt_status mrt_cache_destroy(mrt_ctx* ctx, mrt_cache* cache);
/* rgb4_dev: n x 4 floats (RGB + 0) predicted outgoing radiance */
mrt_status mrt_cache_query(mrt_ctx* ctx, mrt_cache* cache, const mrt_cache_sample* q_dev, uint32_t n, float* rgb4_dev);
/* One training step on n samples with targets target4_dev (n x 4): forward, loss, backward
 * (gradients accumulate), and when apply != 0 an Adam step that also clears the gradients.
 * n = 0 with apply != 0 takes the Adam step on the gradients accumulated so far. */
mrt_status mrt_cache_train(mrt_ctx* ctx, mrt_cache* cache, const mrt_cache_sample* q_dev, const float* target4_dev,
                           uint32_t n, int apply);
/* The MLP input vectors of n samples (n x 64 floats: 32 grid features, dir, normal, albedo,
 * zero padding), as the network sees them (BF16 values). */
mrt_status mrt_cache_inputs(mrt_ctx* ctx, mrt_cache* cache, const mrt_cache_sample* q_dev, uint32_t n, float* x64_dev);
/* Sum of the per-sample losses of the last training step, steps taken, 1 if any MLP weight or
 * any table entry the encoding can read (or, with ema_decay > 0, its running average, which
 * queries use) is not finite (each pointer optional; synchronizes). A dense level's slots past
 * its (res+1)^3 cells are never read, updated or checked. */
mrt_status mrt_cache_stats(mrt_ctx* ctx, mrt_cache* cache, float* last_loss_sum, uint32_t* steps, uint32_t* nonfinite);
/* which: 0 MLP weights [6][64][64] ([layer][out][in]), 1 hash table [levels][2^T][2],
 * 2 MLP weight gradient, 3 table gradient, 4 / 5 running averages of weights / table (2-5 get
 * only). count in floats must match. Adam updates the weights densely and the table sparsely:
 * a table entry with an exactly zero gradient keeps its value and moments. */
mrt_status mrt_cache_get(mrt_ctx* ctx, mrt_cache* cache, uint32_t which, float* host, size_t count);
/* Setting the weights or the table restarts its running average (ema_decay > 0) at the set
 * values, so queries see exactly them until the next step. */
mrt_status mrt_cache_set(mrt_ctx* ctx, mrt_cache* cache, uint32_t which, const float* host, size_t count);

/* ---- cache-terminated path tracing (no reference counterpart; the reference traces every
 * path in full, secondaryRays.comp:64-100). Paths end at vertex `vertex` and add
 * throughput * max(0, cache(vertex)). Training: among the frame's paths that reach that vertex on
 * a surface, each is kept with probability min(1, train_samples / their count) (a hash of the
 * path id and frame_seed against a threshold computed on the device), so about train_samples of
 * them (all when fewer reach it); those are traced in full and each gives one training sample at
 * the vertex with target (L_final - L_before) / throughput. After the frame the cache takes
 * `train_steps` Adam steps on equal shares of those samples. The image is bit-identical to
 * mrt_render when enable = 0 or when a cache weight or a table entry the encoding can read is
 * not finite. ---- */
typedef struct {
    uint32_t enable;                      /* 0: render exactly as mrt_render */
    uint32_t vertex;                      /* 0 = primary hit, 1 = hit of the first bounce ray; < bounces */
    uint32_t train_samples;               /* training samples per frame at the vertex (0: none) */
    uint32_t train_steps;                 /* optimizer steps per frame (>= 1 when training) */
    float* cache_rgb_dev;                 /* optional (spp 1): per pixel max(0, cache) where a path ended in it */
    mrt_cache_sample* train_samples_dev;  /* optional export of the frame's training batch, room for
                                             train_samples + train_samples / 8 + 256 rows ... */
    float* train_targets_dev;             /* ... n x 4 targets ... */
    uint32_t* train_pid_dev;              /* ... path ids ... */
    uint32_t* train_count_dev;            /* ... and its size n */
    uint32_t overlap_training;            /* 1: the frame's training steps are not run at its end but at the
                                             start of this context's next mrt_render_cached with the same
                                             cache, on a second stream beside that frame's traversal (it
                                             waits for them at the cache vertex, so it infers from the same
                                             parameters). Every mrt_cache_* call, and a cached frame with
                                             another cache, first runs steps still left, so the cache is
                                             never observed without them. Ignored under graph capture. */
    uint32_t limits;                      /* MRT_CACHE_LIMIT_* checks, made on the device before the cache
                                             vertex is shaded; a frame that fails one renders exactly
                                             mrt_render's image (its paths continue to full depth) */
    uint32_t query_budget;                /* MRT_CACHE_LIMIT_QUERIES: the most paths that may reach the cache
                                             vertex on a surface (0: none may end into the cache) */
    float loss_budget;                    /* MRT_CACHE_LIMIT_LOSS: the highest mean loss per sample of the
                                             cache's last training step (a loss spike disables the cache) */
} mrt_cache_render_opts;
#define MRT_CACHE_LIMIT_QUERIES 1u
#define MRT_CACHE_LIMIT_LOSS 2u
mrt_status mrt_render_cached(mrt_ctx* ctx, const mrt_scene* scene, mrt_cache* cache, const mrt_camera* camera,
                             const mrt_render_opts* opts, const mrt_cache_render_opts* copts, float* rgb_dev);

/* ---- physical sky: the reference's atmosphere (Atmosphere / Sky, src/gfx/modules/sky.ixx:18-266;
 * src/gpu/sky/gen{Transmittance,MultiScattering,View}.comp) as LUTs on the device, and its miss radiance skyColor
 * (secondaryRays.comp:36-58). Units are km. ---- */
typedef struct {
    float bottom_radius, top_radius;               /* planet and atmosphere-top radii */
    float rayleigh_density_exp_scale;
    float rayleigh_scattering[3];
    float mie_density_exp_scale;
    float mie_scattering[3], mie_extinction[3], mie_absorption[3];
    float mie_phase_g;
    float absorption_density0_layer_width, absorption_density0_constant, absorption_density0_linear;
    float absorption_density1_constant, absorption_density1_linear;
    float absorption_extinction[3];
    float ground_albedo[3];
} mrt_atmosphere_params;                           /* fields of Atmosphere::Params (sky.ixx) */
/* The reference's Earth parameters (Atmosphere::Params::earth). */
mrt_status mrt_sky_earth(mrt_atmosphere_params* out);
/* Transmittance (256x64) and multiple-scattering (32x32) LUTs (the Atmosphere constructor). */
mrt_status mrt_sky_create(mrt_ctx* ctx, const mrt_atmosphere_params* params, mrt_sky** out);
/* The 192x108 sky-view LUT seen from probe_km (Sky::createView, genView.comp) for a sun
 * direction and illuminance; sun_disc != 0 adds the sun disc to the miss radiance, as the
 * reference does (leave it out when the sun is sampled by next-event estimation). */
mrt_status mrt_sky_update_view(mrt_ctx* ctx, mrt_sky* sky, const float probe_km[3], const float sun_dir[3],
                               const float sun_illuminance[3], int sun_disc);
/* Miss radiance of n unit directions (n x 3 floats in and out, device). */
mrt_status mrt_sky_radiance(mrt_ctx* ctx, const mrt_sky* sky, const float* dirs_dev, uint32_t n, float* rgb_dev);
/* which: 0 transmittance, 1 multiple scattering, 2 sky view; 4 floats per texel, rows of x. */
mrt_status mrt_sky_get(mrt_ctx* ctx, const mrt_sky* sky, uint32_t which, float* host, size_t count);
mrt_status mrt_sky_destroy(mrt_ctx* ctx, mrt_sky* sky);

/* ---- progressive accumulation and relative MSE (the cache-vs-plain protocol) ----
 * mrt_accumulate: sum_dev += rgb_dev (n floats); mean_dev (optional) = sum_dev / frames.
 * mrt_rel_mse: mean over n values of (x - r)^2 / (r^2 + eps), summed in double. */
mrt_status mrt_accumulate(mrt_ctx* ctx, float* sum_dev, const float* rgb_dev, size_t n, uint32_t frames,
                          float* mean_dev);
mrt_status mrt_rel_mse(mrt_ctx* ctx, const float* img_dev, const float* ref_dev, size_t n, double eps, double* out);

/* ---- post-processing: the reference's bilateral denoiser (Denoiser::bilateral, src/gfx/modules/denoiser.ixx:34-97;
 * src/gpu/denoise/bilateral.comp) and tonemappers (Tonemapper, src/gfx/modules/tonemapper.ixx:57-373;
 * the compute shaders under src/gpu/tonemap). Images are row-major, device pointers. ---- */
typedef struct {
    float sigma, k_sigma, threshold;  /* BilateralParams (denoiser.ixx:20-31): 5, 2, 0.12 */
    float near_plane;                 /* the camera near plane in kilometres, as the reference's push
                                         constant (its world unit is the km; the depth term is 100x the
                                         distance difference in km): pass near * km-per-scene-unit */
    uint32_t frame;                   /* seeds the dither */
} mrt_bilateral_params;
/* rgb W*H*3, depth W*H (reverse-Z, 0 where the primary ray missed: those pixels pass through
 * unfiltered), normal W*H*4 (xyz used); out W*H*3 = smartDeNoise + 0.005 * dither. out != rgb.
 * Radius round(k_sigma * sigma) up to 64. */
mrt_status mrt_denoise_bilateral(mrt_ctx* ctx, const float* rgb, const float* depth, const float* normal, uint32_t w,
                                 uint32_t h, const mrt_bilateral_params* p, float* out);
enum {
    MRT_TONEMAP_LINEAR = 0, MRT_TONEMAP_REINHARD = 1, MRT_TONEMAP_HABLE = 2, MRT_TONEMAP_ACES = 3,
    MRT_TONEMAP_UCHIMURA = 4, MRT_TONEMAP_AMD = 5      /* renderer.ixx:163-230 order; AMD is the default */
};
typedef struct {
    uint32_t op;
    float exposure;                   /* multiplies the input first */
    float reinhard_max;               /* hdrMax of Tonemapper::reinhard (tonemapper.ixx:106) */
    float uchimura[6];                /* UchimuraParams: maxBrightness, contrast, linearStart, linearLength, blackTightness, pedestal */
    float amd[5];                     /* AMDParams: hdrMax, contrast, shoulder, midIn, midOut */
} mrt_tonemap_params;
/* The reference's defaults (tonemapper.ixx:19-55): AMD operator, exposure 1. */
mrt_status mrt_tonemap_defaults(mrt_tonemap_params* p);
/* n pixels of linear RGB (n*3 floats) -> sRGB-encoded RGB (out_rgb, n*3 floats, optional) and/or
 * RGBA8 UNORM (out_rgba8, n*4 bytes, optional: round(clamp(c,0,1)*255), alpha 255). */
mrt_status mrt_tonemap(mrt_ctx* ctx, const float* rgb, uint32_t n, const mrt_tonemap_params* p, float* out_rgb,
                       uint8_t* out_rgba8);

/* ---- profiling: HIP events around each library kernel on the context stream ---- */
mrt_status mrt_profile_enable(mrt_ctx* ctx, int on);
mrt_status mrt_profile_reset(mrt_ctx* ctx);
/* Total milliseconds and launch count of one kernel name (e.g. "trace_closest") or region. */
mrt_status mrt_profile_query(mrt_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches);
/* A named region of the context stream timed by one pair of HIP events, whatever
 * mrt_profile_enable says: begin = 1 records its start, begin = 0 its end (MRT_ERR_STATE for a
 * region already open, or not open); mrt_profile_query then reports it as one launch per pair.
 * A region brackets many calls without the per-launch events of mrt_profile_enable, which
 * separate every kernel from the next. No reference counterpart (the reference times passes
 * with its render graph's GPU queries). */
mrt_status mrt_profile_region(mrt_ctx* ctx, const char* name, int begin);

#ifdef __cplusplus
}
#endif
#endif
