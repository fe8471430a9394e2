// render.hip — wavefront render stages: primary rays + G-buffer, diffuse path tracing with
// next-event estimation, roulette and stream compaction, and config-0 Lambert shading.
//
// Reference: Pathtracer::primaryRays / secondaryRays (src/gfx/modules/pathtracer.ixx:29-195)
// running primaryRay.comp (one ray per pixel, G-buffer :60-75) and secondaryRays.comp (a
// per-pixel loop over samples and bounces, :64-135). Here the per-pixel loop becomes one
// kernel launch per stage and bounce over SoA path queues:
//   init  -> [shade_k -> trace_any(shadow) -> trace_closest(extension)]_k -> resolve
// A path's random numbers come from its own counter-seeded PCG state, so compaction order
// (wave ballots + one atomic per wave) never changes any path's result.
#include "internal.h"

namespace {

using rt::v3;

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Compacted append: every active lane of the wave must call it. Returns the slot for lanes
// with keep = true.
__device__ __forceinline__ uint32_t wave_append(bool keep, uint32_t* counter) {
    const uint64_t m = __ballot(keep);
    const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt());
    uint32_t base = 0;
    const uint32_t total = (uint32_t)__popcll(m);
    // lowest active lane performs the atomic
    const uint64_t active = __ballot(true);
    const uint32_t leader = (uint32_t)(__ffsll((long long)active) - 1);
    if (__lane_id() == leader && total) base = atomicAdd(counter, total);
    base = __shfl(base, (int)leader, 64);
    return base + rank;
}

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

// ---- primary stage ----------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_camera_rays(rt::CamBasis cam, mrt_ray* __restrict__ rays) {
    const uint32_t npix = cam.width * cam.height;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix) return;
    v3 o, d;
    rt::camera_ray(cam, i % cam.width, i / cam.width, o, d);
    float4* r = reinterpret_cast<float4*>(rays + i);
    r[0] = make_float4(o.x, o.y, o.z, 0.0f);
    r[1] = make_float4(d.x, d.y, d.z, __builtin_huge_valf());
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

// ---- path tracing ------------------------------------------------------------------------
struct PathBuffers {
    mrt_ray* ray;
    uint32_t* pid;
    uint32_t* rng;
    float4* thr;    // xyz throughput, w = 1 while the path may spawn rays, 0 once only NEE is pending
    float4* L;      // xyz accumulated radiance
    float4* pend;   // xyz NEE contribution waiting for its shadow ray
};

struct ShadeParams {
    v3 sky, sun_dir, sun_e;
    uint32_t bounces, rr_start, bounce, capture;
};

__global__ __launch_bounds__(256) void k_paths_init(rt::CamBasis cam, uint32_t spp, uint32_t sample_base,
                                                    uint32_t frame_seed, const mrt_hit* __restrict__ gb_hits,
                                                    PathBuffers cur, mrt_hit* __restrict__ hits,
                                                    uint32_t* __restrict__ counts, uint32_t ncounts) {
    const uint32_t npix = cam.width * cam.height;
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s == 0) {
        counts[0] = npix * spp;
        for (uint32_t k = 1; k < ncounts; ++k) counts[k] = 0;
    }
    if (s >= npix * spp) return;
    const uint32_t pixel = s % npix, j = s / npix;
    v3 o, d;
    rt::camera_ray(cam, pixel % cam.width, pixel / cam.width, o, d);
    float4* r = reinterpret_cast<float4*>(cur.ray + s);
    r[0] = make_float4(o.x, o.y, o.z, 0.0f);
    r[1] = make_float4(d.x, d.y, d.z, __builtin_huge_valf());
    hits[s] = gb_hits[pixel];
    cur.pid[s] = s;
    cur.rng[s] = rt::path_seed(pixel, sample_base + j, frame_seed);
    cur.thr[s] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    cur.L[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    cur.pend[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// One shading step for every path of bounce k (the body of traceSecondaryRays,
// secondaryRays.comp:67-98, for one iteration).
__global__ __launch_bounds__(256) void k_shade(ShadeParams P, PathBuffers cur, PathBuffers nxt,
                                               const mrt_hit* __restrict__ hits, const uint32_t* __restrict__ occ,
                                               mrt_ray* __restrict__ shadow, const float* __restrict__ verts,
                                               const uint32_t* __restrict__ material,
                                               const float* __restrict__ albedo, float4* __restrict__ out_L,
                                               uint32_t* __restrict__ counts) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    const uint32_t n = counts[P.bounce];
    const bool active = s < n;
    bool keep = false;
    // values carried to the compacted write
    uint32_t pid = 0, rng = 0;
    float4 thr4 = make_float4(0, 0, 0, 0), L4 = make_float4(0, 0, 0, 0), pend4 = make_float4(0, 0, 0, 0);
    float4 ray0 = make_float4(0, 0, 0, 0), ray1 = make_float4(0, 0, 0, 0);
    float4 sh0 = make_float4(0, 0, 0, 0), sh1 = make_float4(0, 0, 0, 0);
    if (active) {
        pid = cur.pid[s];
        rng = cur.rng[s];
        thr4 = cur.thr[s];
        L4 = cur.L[s];
        if (P.bounce > 0 && occ[s] == 0u) {
            const float4 pe = cur.pend[s];
            L4.x = L4.x + pe.x; L4.y = L4.y + pe.y; L4.z = L4.z + pe.z;
        }
        const mrt_hit h = hits[s];
        const bool alive = thr4.w != 0.0f;
        bool finished = true;
        if (alive && h.tri == kMissId) {
            L4.x = L4.x + thr4.x * P.sky.x; L4.y = L4.y + thr4.y * P.sky.y; L4.z = L4.z + thr4.z * P.sky.z;
        } else if (alive && P.bounce < P.bounces) {
            const float4 r1 = reinterpret_cast<const float4*>(cur.ray + s)[1];
            const v3 d = rt::make_v3(r1.x, r1.y, r1.z);
            v3 v0, v1, v2, p, nrm;
            tri_vertices(verts, h.tri, v0, v1, v2);
            rt::surface_frame(v0, v1, v2, h.u, h.v, d, p, nrm);
            const float* a3 = albedo + (size_t)material[h.tri] * 3;
            const v3 a = rt::make_v3(a3[0], a3[1], a3[2]);
            const v3 po = rt::offset_origin(p, nrm);
            const float cosl = rt::vdot(nrm, P.sun_dir);
            bool nee = false;
            if (cosl > 0.0f) {
                nee = true;
                pend4 = make_float4(((thr4.x * a.x) * cosl) * P.sun_e.x, ((thr4.y * a.y) * cosl) * P.sun_e.y,
                                    ((thr4.z * a.z) * cosl) * P.sun_e.z, 0.0f);
                sh0 = make_float4(po.x, po.y, po.z, 0.0f);
                sh1 = make_float4(P.sun_dir.x, P.sun_dir.y, P.sun_dir.z, __builtin_huge_valf());
            }
            thr4.x = thr4.x * a.x; thr4.y = thr4.y * a.y; thr4.z = thr4.z * a.z;
            bool spawn = true;
            if (P.bounce + 1 >= P.rr_start) {
                const float q = fminf(0.95f, fmaxf(thr4.x, fmaxf(thr4.y, thr4.z)));
                const float xi = rt::next_float(rng);
                if (!(xi < q)) spawn = false;
                else { thr4.x = thr4.x / q; thr4.y = thr4.y / q; thr4.z = thr4.z / q; }
            }
            if (spawn) {
                const float xi1 = rt::next_float(rng);
                const float xi2 = rt::next_float(rng);
                const v3 nd = rt::cosine_direction(nrm, xi1, xi2);
                ray0 = make_float4(po.x, po.y, po.z, 0.0f);
                ray1 = make_float4(nd.x, nd.y, nd.z, __builtin_huge_valf());
            }
            thr4.w = spawn ? 1.0f : 0.0f;
            finished = !spawn && !nee;
        }
        keep = !finished;
        if (finished && !P.capture) out_L[pid] = L4;
    }
    const uint32_t slot = wave_append(keep, &counts[P.bounce + 1]);
    if (keep) {
        nxt.pid[slot] = pid;
        nxt.rng[slot] = rng;
        nxt.thr[slot] = thr4;
        nxt.L[slot] = L4;
        nxt.pend[slot] = pend4;
        float4* r = reinterpret_cast<float4*>(nxt.ray + slot);
        r[0] = ray0; r[1] = ray1;
        float4* sr = reinterpret_cast<float4*>(shadow + slot);
        sr[0] = sh0; sr[1] = sh1;
    }
}

// Exports the extension rays of the live paths (capture mode).
__global__ __launch_bounds__(256) void k_capture(PathBuffers nxt, const uint32_t* __restrict__ count,
                                                 mrt_ray* __restrict__ out_rays, uint32_t* __restrict__ out_pid,
                                                 uint32_t* __restrict__ out_count) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    const bool live = s < *count && nxt.thr[s].w != 0.0f;
    const uint32_t slot = wave_append(live, out_count);
    if (live) {
        const float4* r = reinterpret_cast<const float4*>(nxt.ray + s);
        float4* o = reinterpret_cast<float4*>(out_rays + slot);
        o[0] = r[0]; o[1] = r[1];
        out_pid[slot] = nxt.pid[s];
    }
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// rgb[pixel] = (sum over samples j in order of L[pixel + j*npix]) / spp  (secondaryRays.comp:131-133)
__global__ __launch_bounds__(256) void k_resolve(const float4* __restrict__ out_L, uint32_t npix, uint32_t spp,
                                                 float* __restrict__ rgb) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix) return;
    float4 acc = out_L[i];
    for (uint32_t j = 1; j < spp; ++j) {
        const float4 v = out_L[i + j * npix];
        acc.x = acc.x + v.x; acc.y = acc.y + v.y; acc.z = acc.z + v.z;
    }
    const float fs = (float)spp;
    rgb[3 * i + 0] = acc.x / fs;
    rgb[3 * i + 1] = acc.y / fs;
    rgb[3 * i + 2] = acc.z / fs;
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
    hipFree(hit); hipFree(shadow); hipFree(occ); hipFree(out_L); hipFree(counts);
    hit = nullptr; shadow = nullptr; occ = nullptr; out_L = nullptr; counts = nullptr;
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
    {
        ProfScope p(ctx, "camera_rays");
        hipLaunchKernelGGL(k_camera_rays, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, now.cam, ctx->ws.cam_rays);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_trace(ctx, s, ctx->ws.cam_rays, (uint32_t)npix, gb.hits_dev, nullptr, 0)) != hipSuccess) return e;
    if (gb.depth_dev || gb.normal_dev || gb.motion_dev) {
        ProfScope p(ctx, "gbuffer_planes");
        hipLaunchKernelGGL(k_gbuffer_planes, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, now, before,
                           gb.hits_dev, s->d_verts, gb.depth_dev, gb.normal_dev, gb.motion_dev);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t run_secondary(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera& cam, const mrt_gbuffer& gb,
                         const mrt_render_opts& o, float* rgb) {
    const uint32_t npix = cam.width * cam.height;
    const size_t cap = (size_t)npix * o.spp;
    Workspace& w = ctx->ws;
    hipError_t e = w.ensure_paths(cap);
    if (e != hipSuccess) return e;
    const rt::CamBasis cb = camera_basis(cam);
    int c = 0;
    PathBuffers buf[2];
    for (int b = 0; b < 2; ++b) buf[b] = PathBuffers{w.ray[b], w.pid[b], w.rng[b], w.thr[b], w.L[b], w.pend[b]};
    {
        ProfScope p(ctx, "paths_init");
        hipLaunchKernelGGL(k_paths_init, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, cb, o.spp, o.sample_base,
                           o.frame_seed, gb.hits_dev, buf[0], w.hit, w.counts, 64u);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (o.capture_bounce) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, ctx->stream, o.capture_count_dev, 1u);
    }
    ShadeParams P;
    P.sky = rt::make_v3(o.sky[0], o.sky[1], o.sky[2]);
    P.sun_dir = rt::make_v3(o.sun_dir[0], o.sun_dir[1], o.sun_dir[2]);
    P.sun_e = rt::make_v3(o.sun_e[0], o.sun_e[1], o.sun_e[2]);
    P.bounces = o.bounces;
    P.rr_start = o.rr_start;
    for (uint32_t k = 0; k <= o.bounces; ++k) {
        P.bounce = k;
        const bool capture_now = o.capture_bounce && k + 1 == o.capture_bounce;
        P.capture = capture_now ? 1u : 0u;
        {
            ProfScope p(ctx, "shade");
            hipLaunchKernelGGL(k_shade, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, P, buf[c], buf[c ^ 1], w.hit,
                               w.occ, w.shadow, s->d_verts, s->d_material, s->d_albedo, w.out_L, w.counts);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        if (capture_now) {
            hipLaunchKernelGGL(k_capture, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, buf[c ^ 1], w.counts + k + 1,
                               o.capture_rays_dev, o.capture_pid_dev, o.capture_count_dev);
            return hipGetLastError();
        }
        if (k == o.bounces) break;
        c ^= 1;
        if ((e = launch_trace(ctx, s, w.shadow, (uint32_t)cap, nullptr, w.occ, 1, nullptr, w.counts + k + 1)) != hipSuccess)
            return e;
        if ((e = launch_trace(ctx, s, buf[c].ray, (uint32_t)cap, w.hit, nullptr, 0, nullptr, w.counts + k + 1)) != hipSuccess)
            return e;
    }
    {
        ProfScope p(ctx, "resolve");
        hipLaunchKernelGGL(k_resolve, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, w.out_L, npix, o.spp, rgb);
    }
    return hipGetLastError();
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
