// render.hip — primary stage (camera rays, closest hit, G-buffer planes), config-0 Lambert
// shading, and the context workspace. The path-tracing stage is in paths.hip.
//
// Reference: Pathtracer::primaryRays (src/gfx/modules/pathtracer.ixx:29-116) running
// primaryRay.comp: one ray per pixel through its centre (:46-51), closest hit (:23-36),
// visibility / depth / normal / motion planes (:60-75).
#include "internal.h"

namespace {

using rt::v3;

struct Frame {
    rt::CamBasis cam;
    float proj_w, proj_h, near_plane;   // perspective (math.ixx:848-860)
};

__device__ __forceinline__ void tri_vertices(const float* __restrict__ verts, uint32_t tri, v3& v0, v3& v1, v3& v2) {
    const float* p = verts + (size_t)tri * 9;
    v0 = rt::make_v3(p[0], p[1], p[2]);
    v1 = rt::make_v3(p[3], p[4], p[5]);
    v2 = rt::make_v3(p[6], p[7], p[8]);
}

__device__ __forceinline__ void project(const Frame& f, v3 p, float& nx, float& ny, float& z) {
    const v3 q = p - f.cam.pos;
    const float vx = rt::vdot(f.cam.right, q), vy = rt::vdot(f.cam.up, q), vz = rt::vdot(f.cam.fwd, q);
    nx = (f.proj_w * vx) / vz;
    ny = (f.proj_h * vy) / vz;
    z = f.near_plane / vz;
}

__global__ __launch_bounds__(256) void k_gbuffer_planes(Frame now, Frame prev, const mrt_hit* __restrict__ hits,
                                                        const float* __restrict__ verts, float* __restrict__ depth,
                                                        float* __restrict__ normal, float* __restrict__ motion) {
    const uint32_t npix = now.cam.width * now.cam.height;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix) return;
    v3 o, d;
    rt::camera_ray(now.cam, i % now.cam.width, i / now.cam.width, o, d);
    const mrt_hit h = hits[i];
    v3 p = o, n = d;
    const bool hit = h.tri != kMissId;
    if (hit) {
        v3 v0, v1, v2;
        tri_vertices(verts, h.tri, v0, v1, v2);
        rt::surface_frame(v0, v1, v2, h.u, h.v, d, p, n);
    }
    if (depth) {
        float x, y, z;
        project(now, p, x, y, z);
        depth[i] = hit ? z : 0.0f;
    }
    if (normal) reinterpret_cast<float4*>(normal)[i] = make_float4(n.x, n.y, n.z, 0.0f);
    if (motion) {
        float x0, y0, z0, x1, y1, z1;
        project(now, p, x0, y0, z0);
        project(prev, p, x1, y1, z1);
        float2 m = hit ? make_float2((x0 - x1) / now.cam.pitch_x, (y0 - y1) / now.cam.pitch_y) : make_float2(0.0f, 0.0f);
        reinterpret_cast<float2*>(motion)[i] = m;
    }
}

// ---- config-0 Lambert + one shadow ray -----------------------------------------------------
__global__ __launch_bounds__(256) void k_lambert_prep(rt::CamBasis cam, const mrt_hit* __restrict__ hits,
                                                      const float* __restrict__ verts,
                                                      const uint32_t* __restrict__ material,
                                                      const float* __restrict__ albedo, v3 l, v3 e,
                                                      mrt_ray* __restrict__ shadow, float4* __restrict__ contrib) {
    const uint32_t npix = cam.width * cam.height;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix) return;
    const mrt_hit h = hits[i];
    float4 c = make_float4(0, 0, 0, 0);
    float4 s0 = make_float4(0, 0, 0, 0), s1 = make_float4(0, 0, 0, 0);
    if (h.tri != kMissId) {
        v3 o, d, v0, v1, v2, p, n;
        rt::camera_ray(cam, i % cam.width, i / cam.width, o, d);
        tri_vertices(verts, h.tri, v0, v1, v2);
        rt::surface_frame(v0, v1, v2, h.u, h.v, d, p, n);
        const float cosl = rt::vdot(n, l);
        if (cosl > 0.0f) {
            const float* a = albedo + (size_t)material[h.tri] * 3;
            c = make_float4((a[0] * cosl) * e.x, (a[1] * cosl) * e.y, (a[2] * cosl) * e.z, 0.0f);
            const v3 po = rt::offset_origin(p, n);
            s0 = make_float4(po.x, po.y, po.z, 0.0f);
            s1 = make_float4(l.x, l.y, l.z, __builtin_huge_valf());
        }
    }
    contrib[i] = c;
    float4* sr = reinterpret_cast<float4*>(shadow + i);
    sr[0] = s0; sr[1] = s1;
}

__global__ __launch_bounds__(256) void k_lambert_finish(const float4* __restrict__ contrib,
                                                        const uint32_t* __restrict__ occ, uint32_t npix,
                                                        float* __restrict__ rgb) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix) return;
    const float4 c = contrib[i];
    const bool lit = occ[i] == 0u;
    rgb[3 * i + 0] = lit ? c.x : 0.0f;
    rgb[3 * i + 1] = lit ? c.y : 0.0f;
    rgb[3 * i + 2] = lit ? c.z : 0.0f;
}

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + 255) / 256); }

Frame make_frame(const mrt_camera& c) {
    Frame f;
    f.cam = camera_basis(c);
    const double h = 1.0 / std::tan(0.5 * (double)c.vertical_fov);
    f.proj_h = (float)h;
    f.proj_w = (float)(h * ((double)c.height / (double)c.width));
    f.near_plane = c.near_plane;
    return f;
}

}  // namespace

// ---- workspace ---------------------------------------------------------------------------
hipError_t Workspace::ensure_paths(size_t cap) {
    if (cap <= path_cap) return hipSuccess;
    release_paths();
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc((void**)p, bytes ? bytes : 256); };
    for (int b = 0; b < 2; ++b) {
        A(&ray[b], cap * sizeof(mrt_ray));
        A(&pid[b], cap * 4);
        A(&rng[b], cap * 4);
        A(&thr[b], cap * 16);
        A(&L[b], cap * 16);
        A(&pend[b], cap * 16);
    }
    A(&hit, cap * sizeof(mrt_hit));
    A(&shadow, cap * sizeof(mrt_ray));
    A(&occ, cap * 4);
    A(&out_L, cap * 16);
    A(&root, cap * 4);
    A(&counts, 64 * 4);
    if (e != hipSuccess) { release_paths(); return e; }
    path_cap = cap;
    return hipSuccess;
}

void Workspace::release_paths() {
    for (int b = 0; b < 2; ++b) {
        hipFree(ray[b]); hipFree(pid[b]); hipFree(rng[b]); hipFree(thr[b]); hipFree(L[b]); hipFree(pend[b]);
        ray[b] = nullptr; pid[b] = rng[b] = nullptr; thr[b] = L[b] = pend[b] = nullptr;
    }
    hipFree(hit); hipFree(shadow); hipFree(occ); hipFree(out_L); hipFree(root); hipFree(counts);
    hit = nullptr; shadow = nullptr; occ = nullptr; out_L = nullptr; root = nullptr; counts = nullptr;
    path_cap = 0;
}

hipError_t Workspace::ensure_pixels(size_t cap) {
    if (cap <= pix_cap) return hipSuccess;
    release_pixels();
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc((void**)p, bytes ? bytes : 256); };
    A(&cam_rays, cap * sizeof(mrt_ray));
    A(&gb_hits, cap * sizeof(mrt_hit));
    A(&contrib, cap * 16);
    A(&pix_occ, cap * 4);
    if (e != hipSuccess) { release_pixels(); return e; }
    pix_cap = cap;
    return hipSuccess;
}

hipError_t Workspace::ensure_cache(size_t queries, size_t train_paths, size_t train_rows) {
    // train_paths: entries of the training radiance, indexed by path id
    if (queries <= cq_cap && train_paths <= ct_cap && train_rows <= ct_rows) return hipSuccess;
    release_cache();
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc((void**)p, bytes ? bytes : 256); };
    A(&cq, queries * sizeof(mrt_cache_sample));
    A(&cq_pid, queries * 4);
    A(&cq_L, queries * 16);
    A(&cq_thr, queries * 16);
    A(&ct_L, train_paths * 16);
    A(&ct_samples, train_rows * sizeof(mrt_cache_sample));
    A(&ct_targets, train_rows * 16);
    A(&c_counts, (kStepCounts + 32) * 4);  // 32: the largest train_steps
    if (e != hipSuccess) { release_cache(); return e; }
    cq_cap = queries;
    ct_cap = train_paths;
    ct_rows = train_rows;
    return hipSuccess;
}

void Workspace::release_cache() {
    hipFree(cq); hipFree(cq_pid); hipFree(cq_L); hipFree(cq_thr); hipFree(ct_L); hipFree(ct_samples);
    hipFree(ct_targets); hipFree(c_counts);
    cq = nullptr; cq_pid = nullptr; cq_L = cq_thr = nullptr; ct_L = nullptr; ct_samples = nullptr;
    ct_targets = nullptr; c_counts = nullptr;
    cq_cap = ct_cap = ct_rows = 0;
}

void Workspace::release_pixels() {
    hipFree(cam_rays); hipFree(gb_hits); hipFree(contrib); hipFree(pix_occ);
    cam_rays = nullptr; gb_hits = nullptr; contrib = nullptr; pix_occ = nullptr;
    pix_cap = 0;
}

// ---- stage launchers -----------------------------------------------------------------------
hipError_t run_primary(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera& cam, const mrt_camera& prev,
                       const mrt_gbuffer& gb) {
    const size_t npix = (size_t)cam.width * cam.height;
    hipError_t e = ctx->ws.ensure_pixels(npix);
    if (e != hipSuccess) return e;
    Frame now = make_frame(cam), before = make_frame(prev);
    // The previous position is projected with the current projection and the previous view
    // (projection * prevView, primaryRay.comp:73; only prevCamera.view() is passed,
    // pathtracer.ixx:102): the previous camera's field of view and viewport do not enter.
    before.proj_w = now.proj_w;
    before.proj_h = now.proj_h;
    // the traversal generates each pixel's camera ray itself (no ray array written and read back)
    if ((e = launch_trace_camera(ctx, s, now.cam, gb.hits_dev)) != hipSuccess) return e;
    if (gb.depth_dev || gb.normal_dev || gb.motion_dev) {
        ProfScope p(ctx, "gbuffer_planes");
        hipLaunchKernelGGL(k_gbuffer_planes, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, now, before,
                           gb.hits_dev, s->d_verts, gb.depth_dev, gb.normal_dev, gb.motion_dev);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t run_lambert(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera& cam, const mrt_gbuffer& gb,
                       const float* l, const float* le, float* rgb) {
    const size_t npix = (size_t)cam.width * cam.height;
    Workspace& w = ctx->ws;
    hipError_t e = w.ensure_pixels(npix);
    if (e != hipSuccess) return e;
    const rt::CamBasis cb = camera_basis(cam);
    {
        ProfScope p(ctx, "lambert_prep");
        hipLaunchKernelGGL(k_lambert_prep, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, cb, gb.hits_dev, s->d_verts,
                           s->d_material, s->d_albedo, rt::make_v3(l[0], l[1], l[2]), rt::make_v3(le[0], le[1], le[2]),
                           w.cam_rays, w.contrib);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = launch_trace(ctx, s, w.cam_rays, (uint32_t)npix, nullptr, w.pix_occ, 1)) != hipSuccess) return e;
    {
        ProfScope p(ctx, "lambert_finish");
        hipLaunchKernelGGL(k_lambert_finish, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, w.contrib, w.pix_occ,
                           (uint32_t)npix, rgb);
    }
    return hipGetLastError();
}
