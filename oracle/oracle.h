/* oracle.h — C entry points of the scalar CPU oracle (test infrastructure only).
 *
 * The oracle restates the MinoteRT hot path on the CPU: camera ray generation
 * (primaryRay.comp:46-51), closest hit (primaryRay.comp:23-36), the diffuse estimator
 * (secondaryRays.comp:64-100), the PCG RNG (random.glsl:22-31) and sphere sampling
 * (random.glsl:10-19), extended by the documented decisions for the triangle scenes
 * (watertight triangle test, binned-SAH BVH, next-event estimation, Russian roulette).
 * Only tests, the smoke check and the bench's CPU-baseline leg load this library.
 *
 * Layouts shared with the product ABI:
 *   ray  = 8 floats  {ox, oy, oz, tmin, dx, dy, dz, tmax}           (32 B)
 *   hit  = 4 words   {t (f32), tri (u32), u (f32), v (f32)}          (16 B); miss: tri = ~0u, t = +inf
 *   tris = 9 floats per triangle {v0, v1, v2}
 */
#ifndef MINOTE_ORACLE_H
#define MINOTE_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float pos[3]; float yaw, pitch, vfov; uint32_t width, height;
} orc_camera;

typedef struct {
    uint32_t seed, n_spheres, level, ground;
    float box, r_min, r_max, alb_min, alb_max;
    float ground_z, ground_half, ground_albedo;
} orc_scene_desc;

typedef struct {
    uint32_t frame_seed;   /* per-frame seed; paths use pcg(pixel ^ pcg(sample ^ pcg(frame_seed))) */
    uint32_t sample;       /* sample index inside the frame */
    uint32_t bounces;      /* extension rays after the primary ray */
    uint32_t rr_start;     /* first extension ray index (1-based) that is subject to roulette */
    float sky[3];          /* constant environment radiance seen by escaped rays */
    float sun_dir[3];      /* unit direction towards the directional light */
    float sun_e[3];        /* light scale: contribution = thr * albedo * max(0, n.l) * E * V */
    uint32_t cache_mode;   /* 0: plain; 1: terminate at vertex cache_vertex into cache_rgb */
    uint32_t cache_vertex; /* 0 = primary hit, 1 = the hit of the first extension ray, ... */
    const struct orc_sky_state* sky_state;  /* non-NULL: escaped rays take this sky's radiance */
} orc_path_params;

uint32_t orc_pcg(uint32_t v);
float orc_rand_float(uint32_t* state);
void orc_sincos_turns(float x, float* s, float* c);
void orc_cosine_dir(const float n[3], float xi1, float xi2, float out[3]);
void orc_camera_basis(const orc_camera* cam, float out[14]);
void orc_primary_rays(const orc_camera* cam, float* rays);
void orc_offset_ray(const float p[3], const float n[3], float out[3]);

uint32_t orc_scene_tri_count(const orc_scene_desc* d);
/* verts: n*9 floats, mat: n u32, albedo: (n_spheres + ground) * 3 floats */
int orc_scene_generate(const orc_scene_desc* d, float* verts, uint32_t* mat, float* albedo);

int orc_intersect(const float ray[8], const float tri[9], uint32_t id, float hit[4]);
void orc_closest_brute(const float* verts, uint32_t ntri, const float* rays, uint32_t nrays,
                       uint32_t* hits, int threads);

typedef struct orc_bvh orc_bvh;
orc_bvh* orc_bvh_build(const float* verts, uint32_t ntri);
void orc_bvh_free(orc_bvh* b);
/* info: nodes, leaves, max_depth, max_leaf_size; sah cost (Ct=Ci=1) in *sah */
void orc_bvh_info(const orc_bvh* b, uint32_t info[4], double* sah);
/* Structural check: every triangle in exactly one leaf, child boxes inside parents,
 * leaves hold 1..4 triangles. Returns 0 if valid, else a nonzero code. */
int orc_bvh_validate(const orc_bvh* b);
void orc_trace_closest(const orc_bvh* b, const float* rays, uint32_t n, uint32_t* hits,
                       uint64_t* box_tests, uint64_t* tri_tests, int threads);
void orc_trace_any(const orc_bvh* b, const float* rays, uint32_t n, uint8_t* occluded, int threads);

/* Config-0 shading: primary ray, closest hit, Lambert with one shadow ray. */
void orc_render_lambert(const orc_bvh* b, const uint32_t* mat, const float* albedo,
                        const orc_camera* cam, const float light_dir[3], const float light_e[3],
                        uint32_t* hits, float* rgb, int threads);
/* One sample of diffuse path tracing for every pixel (configs 1-4). cache_rgb (optional,
 * per pixel) supplies the radiance used when cache_mode terminates a path at vertex 1. */
void orc_render_path(const orc_bvh* b, const uint32_t* mat, const float* albedo,
                     const orc_camera* cam, const orc_path_params* p, const float* cache_rgb,
                     float* rgb, int threads);
/* The first extension ray of every pixel whose primary ray hits (config-1 benchmark set).
 * Returns the count; rays: count*8 floats, pixel: count u32 (ascending pixel order). */
uint32_t orc_bounce1_rays(const orc_bvh* b, const orc_camera* cam, const orc_path_params* p,
                          float* rays, uint32_t* pixel, int threads);
/* n rays with origins uniform in [-box, box]^3 and directions uniform on the sphere. */
void orc_random_rays(uint32_t seed, float box, uint32_t n, float* rays);

double orc_rel_mse(const float* img, const float* ref, size_t n_values, double eps);

/* Cache training data of plain paths (cache_mode ignored): for path id pid (pixel pid % npix,
 * sample p->sample + pid / npix) the cache sample at vertex cache_vertex (16 floats: position,
 * 0, incoming direction, 0, normal, 0, albedo, 0) and the target (L_final - L_before) / thr
 * (4 floats, w = 0), where L_before is the radiance gathered before the vertex and thr the
 * throughput arriving there. valid[i] = 0 when the path has no surface vertex there. */
void orc_path_suffix(const orc_bvh* b, const uint32_t* mat, const float* albedo, const orc_camera* cam,
                     const orc_path_params* p, const uint32_t* pids, uint32_t n, float* samples, float* targets,
                     uint8_t* valid, int threads);

/* ---- physical sky (oracle_sky.cpp): the reference's atmosphere LUTs and miss radiance ---- */
typedef struct {
    float bottom_radius, top_radius;               /* km */
    float rayleigh_density_exp_scale;
    float rayleigh_scattering[3];
    float mie_density_exp_scale;
    float mie_scattering[3], mie_extinction[3], mie_absorption[3];
    float mie_phase_g;
    float absorption_density0_layer_width, absorption_density0_constant, absorption_density0_linear;
    float absorption_density1_constant, absorption_density1_linear;
    float absorption_extinction[3];
    float ground_albedo[3];
} orc_atmosphere;
void orc_sky_earth(orc_atmosphere* p);
/* A sky as the path tracer sees it: LUTs (w*h*3 floats), the probe, sun and sun-disc switch. */
typedef struct orc_sky_state {
    orc_atmosphere atmo;
    const float* trans;
    uint32_t tw, th;
    const float* view;
    uint32_t vw, vh;
    float probe[3], sun_dir[3], illum[3];
    int sun_disc;
} orc_sky_state;
void orc_sky_state_radiance(const orc_sky_state* s, const float dir[3], float out[3]);
/* test hooks: the sky's exp/sin/cos/acos/x^1.5 (fn 0..4), phase functions (mie = 0: Rayleigh)
 * and the LUT parameterizations */
void orc_tx_eval(int fn, const float* x, uint32_t n, float* out);
float orc_sky_phase(int mie, float g, float cos_t);
void orc_sky_trans_uv(float h, float mu, float bottom, float top, float out[2]);
void orc_sky_trans_params(float u, float v, float bottom, float top, float out[2]);
void orc_sky_view_uv(int ground, float vz, float lv, int w, int h, float view_h, float bottom, float out[2]);
void orc_sky_view_params(float u, float v, int w, int h, float view_h, float bottom, float out[2]);
/* LUTs are w*h*3 floats, row-major (x fastest) */
void orc_sky_transmittance(const orc_atmosphere* p, uint32_t w, uint32_t h, float* out);
void orc_sky_multiscattering(const orc_atmosphere* p, const float* trans, uint32_t tw, uint32_t th, uint32_t w,
                             uint32_t h, float* out);
void orc_sky_view(const orc_atmosphere* p, const float* trans, uint32_t tw, uint32_t th, const float* ms, uint32_t mw,
                  uint32_t mh, const float probe[3], const float sun_dir[3], const float illum[3], uint32_t w, uint32_t h,
                  float* out);
void orc_sky_radiance(const orc_atmosphere* p, const float* trans, uint32_t tw, uint32_t th, const float* view,
                      uint32_t vw, uint32_t vh, const float probe[3], const float sun_dir[3], const float illum[3],
                      int sun_disc, const float* dirs, uint32_t n, float* out);

/* ---- post-processing (oracle_post.cpp): bilateral denoiser and tonemappers ---- */
typedef struct { float sigma, k_sigma, threshold, near_plane; uint32_t frame; } orc_bilateral;
/* rgb W*H*3, depth W*H (reverse-Z, 0 = no hit), normal W*H*4; out W*H*3 */
void orc_denoise_bilateral(const float* rgb, const float* depth, const float* normal, uint32_t w, uint32_t h,
                           const orc_bilateral* p, float* out);
typedef struct {
    uint32_t op;           /* 0 linear, 1 Reinhard, 2 Hable, 3 ACES, 4 Uchimura, 5 AMD (renderer.ixx:165-172) */
    float exposure;
    float reinhard_max;
    float uchimura[6];     /* max brightness, contrast, linear start, linear length, black tightness, pedestal */
    float amd[5];          /* hdr max, contrast, shoulder, mid in, mid out */
} orc_tonemap_params;
void orc_tonemap_defaults(orc_tonemap_params* p);
/* n pixels of RGB in; sRGB-encoded RGB out (optional) and RGBA8 UNORM (optional) */
void orc_tonemap(const orc_tonemap_params* p, const float* rgb, uint32_t n, float* out, uint8_t* rgba8);

/* ---- radiance cache (oracle_cache.cpp) ---- */
typedef struct {
    uint32_t levels, log2_table;
    float base_res, max_res;
    float lo[3], hi[3];
} orc_grid;
float orc_bf16_round(float f);
/* x0: n*64 MLP inputs (32 hash features, dir, normal, albedo, zero padding), BF16-rounded */
void orc_cache_encode(const orc_grid* g, const float* table, const float* pos, const float* dir, const float* nrm,
                      const float* alb, uint32_t n, float scale, float* x0);
/* W: 6*64*64 FP32 master weights ([layer][out][in]); y: n*4; acts (optional): 5*n*64 */
void orc_mlp_forward(const float* W, const float* x0, uint32_t n, float* y, float* acts);
double orc_mlp_backward(const float* W, const float* x0, uint32_t n, const float* target, float* dW, float* dx0);
void orc_cache_encode_backward(const orc_grid* g, const float* pos, const float* dx0, uint32_t n, float* grad_table);
/* Adam with bias correction; sparse != 0 skips entries whose gradient is exactly zero (parameter
 * and moments unchanged), as for hash-grid tables. */
void orc_adam(float* p, const float* g, float* m, float* v, uint32_t n, uint32_t step, float lr, float b1, float b2,
              float eps, int sparse);
int orc_hardware_threads(void);

#ifdef __cplusplus
}
#endif
#endif
