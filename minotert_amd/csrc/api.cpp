// api.cpp — the C ABI of include/minotert.h: context, device memory, scenes, traversal
// entry points and kernel profiling. No exceptions cross this boundary; every failure is a
// status code plus a message retrievable with mrt_last_error.
#include "internal.h"
#include "sky.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

hipError_t launch_copy_bytes(mrt_ctx* ctx, void* dst, const void* src, size_t bytes);
hipError_t launch_fill_u32(mrt_ctx* ctx, uint32_t* dst, size_t n, uint32_t value);

namespace {

thread_local std::string g_last_error;

mrt_status fail(mrt_ctx* ctx, mrt_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    if (ctx) ctx->last_error = buf;
    return st;
}

#define HIP_TRY(ctx, expr)                                                                     \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail((ctx), _e == hipErrorOutOfMemory ? MRT_ERR_OOM : MRT_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(_e), __FILE__, __LINE__);                    \
    } while (0)

template <typename T>
hipError_t dev_alloc(T** p, size_t count, size_t& total) {
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    total += bytes;
    return hipMalloc((void**)p, bytes);
}

void free_scene(mrt_scene* s) {
    if (!s) return;
    hipFree(s->d_nodes);
    hipFree(s->d_tris);
    hipFree(s->d_verts);
    hipFree(s->d_material);
    hipFree(s->d_albedo);
    hipFree(s->d_overflow);
    delete s;
}

mrt_status collect_profile(mrt_ctx* ctx) {
    if (ctx->pending.empty()) return MRT_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& ev : ctx->pending) {
        float ms = 0.0f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ev.a, ev.b));
        ProfAccum& acc = ctx->prof[ev.name];
        acc.ms += ms;
        acc.launches += 1;
        ctx->event_pool.push_back(ev.a);
        ctx->event_pool.push_back(ev.b);
    }
    ctx->pending.clear();
    return MRT_OK;
}

}  // namespace

mrt_status api_fail(mrt_ctx* ctx, mrt_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    if (ctx) ctx->last_error = buf;
    return st;
}

ProfScope::ProfScope(mrt_ctx* c, const char* n, hipStream_t s) : ctx(c), name(n), stream(s) {
    if (!ctx || !ctx->profiling) return;
    if (!stream) stream = ctx->stream;
    for (hipEvent_t* e : {&a, &b}) {
        if (!ctx->event_pool.empty()) { *e = ctx->event_pool.back(); ctx->event_pool.pop_back(); }
        else if (hipEventCreate(e) != hipSuccess) { *e = nullptr; }
    }
    if (a) hipEventRecord(a, stream);
}

ProfScope::~ProfScope() {
    if (!ctx || !ctx->profiling || !a || !b) return;
    hipEventRecord(b, stream);
    ctx->pending.push_back(ProfEvent{name, a, b});
}

extern "C" {

uint32_t mrt_abi_version(void) { return MRT_ABI_VERSION; }

const char* mrt_last_error(const mrt_ctx* ctx) {
    if (ctx && !ctx->last_error.empty()) return ctx->last_error.c_str();
    return g_last_error.c_str();
}

mrt_status mrt_create(mrt_ctx** out, const mrt_opts* opts) {
    if (!out) return fail(nullptr, MRT_ERR_INVALID, "mrt_create: out is null");
    *out = nullptr;
    int device = opts ? opts->device : 0;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(nullptr, MRT_ERR_HIP, "mrt_create: no HIP device (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count) return fail(nullptr, MRT_ERR_INVALID, "mrt_create: device %d of %d", device, count);
    auto* ctx = new (std::nothrow) mrt_ctx();
    if (!ctx) return fail(nullptr, MRT_ERR_OOM, "mrt_create: host allocation failed");
    ctx->device = device;
    auto bail = [&](hipError_t err, const char* what) {
        fail(nullptr, MRT_ERR_HIP, "mrt_create: %s: %s", what, hipGetErrorString(err));
        if (ctx->stream) hipStreamDestroy(ctx->stream);
        if (ctx->side) hipStreamDestroy(ctx->side);
        if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
        if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
        if (ctx->ev_train) hipEventDestroy(ctx->ev_train);
        if (ctx->ev_pending) hipEventDestroy(ctx->ev_pending);
        if (ctx->d_counters) hipFree(ctx->d_counters);
        delete ctx;
        return MRT_ERR_HIP;
    };
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return bail(e, "hipGetDeviceProperties");
    ctx->num_cus = prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking)) != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreateWithFlags(&ctx->ev_train, hipEventDisableTiming)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreateWithFlags(&ctx->ev_pending, hipEventDisableTiming)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipMalloc((void**)&ctx->d_counters, 2048 * sizeof(uint32_t))) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMemset(ctx->d_counters, 0, 2048 * sizeof(uint32_t))) != hipSuccess) return bail(e, "hipMemset");
    if ((e = noalloc_scratch_init(device, &ctx->noalloc_scratch)) != hipSuccess) return bail(e, "hipGetSymbolAddress");
    *out = ctx;
    return MRT_OK;
}

mrt_status mrt_destroy(mrt_ctx* ctx) {
    if (!ctx) return MRT_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipStreamSynchronize(ctx->side);
    for (auto& ev : ctx->pending) { hipEventDestroy(ev.a); hipEventDestroy(ev.b); }
    for (auto ev : ctx->event_pool) hipEventDestroy(ev);
    for (auto& r : ctx->region_open) hipEventDestroy(r.second);
    ctx->ws.release_paths();
    ctx->ws.release_pixels();
    ctx->ws.release_cache();
    hipFree(ctx->d_partials);
    ctx->d_partials = nullptr;
    ctx->sort.release();
    hipFree(ctx->d_counters);
    hipStreamDestroy(ctx->stream);
    hipStreamDestroy(ctx->side);
    hipEventDestroy(ctx->ev_fork);
    hipEventDestroy(ctx->ev_join);
    hipEventDestroy(ctx->ev_train);
    hipEventDestroy(ctx->ev_pending);
    delete ctx;
    return MRT_OK;
}

mrt_status mrt_sync(mrt_ctx* ctx) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_sync: null context");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MRT_OK;
}

mrt_status mrt_device_info(mrt_ctx* ctx, uint32_t info[4]) {
    if (!ctx || !info) return fail(ctx, MRT_ERR_INVALID, "mrt_device_info: null argument");
    hipDeviceProp_t prop;
    HIP_TRY(ctx, hipGetDeviceProperties(&prop, ctx->device));
    info[0] = (uint32_t)prop.multiProcessorCount;
    info[1] = (uint32_t)prop.warpSize;
    info[2] = (uint32_t)prop.sharedMemPerBlock;
    info[3] = (uint32_t)ctx->device;
    return MRT_OK;
}

mrt_status mrt_device_name(mrt_ctx* ctx, char* buf, size_t len) {
    if (!ctx || !buf || len == 0) return fail(ctx, MRT_ERR_INVALID, "mrt_device_name: null argument");
    hipDeviceProp_t prop;
    HIP_TRY(ctx, hipGetDeviceProperties(&prop, ctx->device));
    std::snprintf(buf, len, "%s-%dCU", prop.gcnArchName, prop.multiProcessorCount);
    return MRT_OK;
}

mrt_status mrt_alloc(mrt_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(ctx, MRT_ERR_INVALID, "mrt_alloc: null argument");
    *out = nullptr;
    HIP_TRY(ctx, hipMalloc(out, bytes ? bytes : 256));
    return MRT_OK;
}

mrt_status mrt_free(mrt_ctx* ctx, void* p) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_free: null context");
    if (!p) return MRT_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipFree(p));
    return MRT_OK;
}

mrt_status mrt_upload(mrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(ctx, MRT_ERR_INVALID, "mrt_upload: null argument");
    if (!bytes) return MRT_OK;
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MRT_OK;
}

mrt_status mrt_download(mrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(ctx, MRT_ERR_INVALID, "mrt_download: null argument");
    if (!bytes) return MRT_OK;
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MRT_OK;
}

mrt_status mrt_copy_device(mrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(ctx, MRT_ERR_INVALID, "mrt_copy_device: null argument");
    HIP_TRY(ctx, launch_copy_bytes(ctx, dst, src, bytes));
    return MRT_OK;
}

mrt_status mrt_check_size(uint64_t n_tris, uint64_t n_nodes) {
    if (n_tris > MRT_MAX_TRIANGLES)
        return fail(nullptr, MRT_ERR_INVALID, "%llu triangles exceed the limit of %u (28-bit leaf triangle index)",
                    (unsigned long long)n_tris, MRT_MAX_TRIANGLES);
    if (n_nodes >= MRT_MAX_WIDE_NODES)
        return fail(nullptr, MRT_ERR_INVALID, "%llu BVH4 nodes exceed the limit of %u (32-bit node byte offsets)",
                    (unsigned long long)n_nodes, MRT_MAX_WIDE_NODES - 1u);
    return MRT_OK;
}

mrt_status mrt_bvh_build_host(const float* verts, uint32_t n_tris, const mrt_bvh_opts* opts, void* nodes_host,
                              size_t nodes_capacity_bytes, void* tris_host, size_t tris_bytes, mrt_scene_stats* stats) {
    if (n_tris && (!verts || !tris_host)) return fail(nullptr, MRT_ERR_INVALID, "mrt_bvh_build_host: null triangle buffer");
    if (!nodes_host) return fail(nullptr, MRT_ERR_INVALID, "mrt_bvh_build_host: null node buffer");
    const size_t need_nodes = (size_t)std::max<uint32_t>(n_tris, 1u) * sizeof(WNode);
    if (nodes_capacity_bytes < need_nodes)
        return fail(nullptr, MRT_ERR_INVALID, "mrt_bvh_build_host: nodes need room for %zu bytes", need_nodes);
    if (tris_bytes != (size_t)n_tris * 3 * sizeof(float4))
        return fail(nullptr, MRT_ERR_INVALID, "mrt_bvh_build_host: triangles need %zu bytes", (size_t)n_tris * 3 * sizeof(float4));
    const uint32_t max_leaf = opts && opts->max_leaf ? opts->max_leaf : 4;
    const uint32_t bins = opts && opts->bins ? opts->bins : 16;
    if (max_leaf < 1 || max_leaf > 8) return fail(nullptr, MRT_ERR_INVALID, "mrt_bvh_build_host: max_leaf %u not in 1..8", max_leaf);
    if (bins < 4 || bins > 64) return fail(nullptr, MRT_ERR_INVALID, "mrt_bvh_build_host: bins %u not in 4..64", bins);
    if (mrt_check_size(n_tris, 0) != MRT_OK) return MRT_ERR_INVALID;
    for (size_t i = 0; i < (size_t)n_tris * 9; ++i)
        if (!std::isfinite(verts[i]))
            return fail(nullptr, MRT_ERR_INVALID, "mrt_bvh_build_host: triangle %zu has a non-finite vertex coordinate", i / 9);
    HostBVH h;
    build_bvh4(verts, n_tris, max_leaf, bins, h);
    if (h.nodes.size() * sizeof(WNode) > nodes_capacity_bytes)  // cannot happen: at most max(n, 1) nodes
        return fail(nullptr, MRT_ERR_STATE, "mrt_bvh_build_host: %zu nodes exceed the stated bound", h.nodes.size());
    std::memcpy(nodes_host, h.nodes.data(), h.nodes.size() * sizeof(WNode));
    if (n_tris) std::memcpy(tris_host, h.tri4.data(), h.tri4.size() * sizeof(float));
    if (stats) {
        stats->n_tris = n_tris;
        stats->n_nodes = (uint32_t)h.nodes.size();
        stats->depth = h.depth;
        stats->leaves = h.leaves;
        stats->build_ms = h.build_ms;
        stats->device_bytes = 0;
    }
    return mrt_check_size(n_tris, h.nodes.size());
}

mrt_status mrt_scene_create(mrt_ctx* ctx, const float* verts, uint32_t n_tris, const uint32_t* tri_material,
                            const float* material_albedo, uint32_t n_materials, const mrt_bvh_opts* opts,
                            mrt_scene** out) {
    if (!ctx || !out) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: null argument");
    *out = nullptr;
    if (n_tris && (!verts || !tri_material)) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: null triangle data");
    if (n_materials && !material_albedo) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: null albedo");
    if (mrt_check_size(n_tris, 0) != MRT_OK) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: %s", g_last_error.c_str());
    for (uint32_t t = 0; t < n_tris; ++t)
        if (tri_material[t] >= n_materials)
            return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: triangle %u has material %u of %u", t, tri_material[t], n_materials);
    uint32_t max_leaf = opts && opts->max_leaf ? opts->max_leaf : 4;
    uint32_t bins = opts && opts->bins ? opts->bins : 16;
    const uint32_t builder = opts ? opts->builder : MRT_BVH_HOST;
    if (max_leaf < 1 || max_leaf > 8) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: max_leaf %u not in 1..8", max_leaf);
    if (bins < 4 || bins > 64) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: bins %u not in 4..64", bins);
    if (builder != MRT_BVH_HOST && builder != MRT_BVH_DEVICE)
        return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: unknown builder %u", builder);
    for (size_t i = 0; i < (size_t)n_tris * 9; ++i)
        if (!std::isfinite(verts[i]))
            return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: triangle %zu has a non-finite vertex coordinate", i / 9);
    auto* s = new (std::nothrow) mrt_scene();
    if (!s) return fail(ctx, MRT_ERR_OOM, "mrt_scene_create: host allocation failed");
    s->n_tris = n_tris;
    s->n_materials = n_materials;
    for (int k = 0; k < 3; ++k) { s->bounds[k] = 3.4e38f; s->bounds[3 + k] = -3.4e38f; }
    for (size_t i = 0; i < (size_t)n_tris * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            s->bounds[k] = std::min(s->bounds[k], verts[3 * i + k]);
            s->bounds[3 + k] = std::max(s->bounds[3 + k], verts[3 * i + k]);
        }
    size_t total = 0;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    chk(dev_alloc(&s->d_verts, (size_t)n_tris * 9, total));
    chk(dev_alloc(&s->d_material, (size_t)n_tris, total));
    chk(dev_alloc(&s->d_albedo, (size_t)n_materials * 3, total));
    if (e == hipSuccess) {
        if (n_tris) {
            chk(hipMemcpy(s->d_verts, verts, (size_t)n_tris * 9 * sizeof(float), hipMemcpyHostToDevice));
            chk(hipMemcpy(s->d_material, tri_material, (size_t)n_tris * sizeof(uint32_t), hipMemcpyHostToDevice));
        }
        if (n_materials)
            chk(hipMemcpy(s->d_albedo, material_albedo, (size_t)n_materials * 3 * sizeof(float), hipMemcpyHostToDevice));
    }
    if (e == hipSuccess && builder == MRT_BVH_DEVICE && n_tris > 0) {
        DeviceBVH d;
        chk(build_bvh4_device(ctx, s->d_verts, n_tris, max_leaf, bins, d));
        if (e == hipSuccess) {
            s->d_nodes = d.d_nodes;
            s->d_tris = d.d_tris;
            s->n_nodes = d.n_nodes;
            s->depth = d.depth;
            s->leaves = d.leaves;
            s->build_ms = d.build_ms;
            total += (size_t)d.n_nodes * sizeof(WNode) + (size_t)n_tris * 3 * sizeof(float4);
        }
    } else if (e == hipSuccess) {
        HostBVH h;
        build_bvh4(verts, n_tris, max_leaf, bins, h);
        s->n_nodes = (uint32_t)h.nodes.size();
        s->depth = h.depth;
        s->leaves = h.leaves;
        s->build_ms = h.build_ms;
        chk(dev_alloc(&s->d_nodes, h.nodes.size(), total));
        chk(dev_alloc(&s->d_tris, (size_t)n_tris * 3, total));
        if (e == hipSuccess) {
            chk(hipMemcpy(s->d_nodes, h.nodes.data(), h.nodes.size() * sizeof(WNode), hipMemcpyHostToDevice));
            if (n_tris) chk(hipMemcpy(s->d_tris, h.tri4.data(), h.tri4.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    if (e == hipSuccess && mrt_check_size(n_tris, s->n_nodes) != MRT_OK) {
        const std::string why = g_last_error;
        free_scene(s);
        return fail(ctx, MRT_ERR_INVALID, "mrt_scene_create: %s", why.c_str());
    }
    // Traversal stack bound: at most 3 pending siblings per internal level on the path.
    const size_t max_stack = 3 * (size_t)s->depth + 1;
    const size_t lds = (size_t)trace_min_lds_depth();
    s->overflow_entries = max_stack > lds ? max_stack - lds : 1;
    s->overflow_threads = (size_t)ctx->num_cus * 2048;
    if (e == hipSuccess) chk(dev_alloc(&s->d_overflow, 2 * s->overflow_entries * s->overflow_threads * 2, total));
    if (e != hipSuccess) {
        free_scene(s);
        return fail(ctx, e == hipErrorOutOfMemory ? MRT_ERR_OOM : MRT_ERR_HIP, "mrt_scene_create: %s", hipGetErrorString(e));
    }
    s->device_bytes = total;
    *out = s;
    return MRT_OK;
}

mrt_status mrt_scene_export(mrt_ctx* ctx, const mrt_scene* scene, void* nodes_host, size_t nodes_bytes, void* tris_host,
                            size_t tris_bytes) {
    if (!ctx || !scene) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_export: null argument");
    if (nodes_host && nodes_bytes != (size_t)scene->n_nodes * sizeof(WNode))
        return fail(ctx, MRT_ERR_INVALID, "mrt_scene_export: nodes need %zu bytes", (size_t)scene->n_nodes * sizeof(WNode));
    if (tris_host && tris_bytes != (size_t)scene->n_tris * 3 * sizeof(float4))
        return fail(ctx, MRT_ERR_INVALID, "mrt_scene_export: triangles need %zu bytes", (size_t)scene->n_tris * 3 * sizeof(float4));
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && nodes_host && nodes_bytes) e = hipMemcpy(nodes_host, scene->d_nodes, nodes_bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess && tris_host && tris_bytes) e = hipMemcpy(tris_host, scene->d_tris, tris_bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(ctx, MRT_ERR_HIP, "mrt_scene_export: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_scene_destroy(mrt_ctx* ctx, mrt_scene* s) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_scene_destroy: null context");
    if (!s) return MRT_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    free_scene(s);
    return MRT_OK;
}

mrt_status mrt_scene_stats_get(mrt_ctx* ctx, const mrt_scene* s, mrt_scene_stats* out) {
    if (!ctx || !s || !out) return fail(ctx, MRT_ERR_INVALID, "mrt_scene_stats_get: null argument");
    out->n_tris = s->n_tris;
    out->n_nodes = s->n_nodes;
    out->depth = s->depth;
    out->leaves = s->leaves;
    out->build_ms = s->build_ms;
    out->device_bytes = s->device_bytes;
    return MRT_OK;
}

mrt_status mrt_trace_closest(mrt_ctx* ctx, const mrt_scene* s, const mrt_ray* rays, uint32_t n, mrt_hit* hits) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_trace_closest: null context or scene");
    if (n && (!rays || !hits)) return fail(ctx, MRT_ERR_INVALID, "mrt_trace_closest: null ray or hit buffer");
    HIP_TRY(ctx, launch_trace(ctx, s, rays, n, hits, nullptr, 0));
    return MRT_OK;
}

mrt_status mrt_trace_any(mrt_ctx* ctx, const mrt_scene* s, const mrt_ray* rays, uint32_t n, uint32_t* occluded) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_trace_any: null context or scene");
    if (n && (!rays || !occluded)) return fail(ctx, MRT_ERR_INVALID, "mrt_trace_any: null ray or output buffer");
    HIP_TRY(ctx, launch_trace(ctx, s, rays, n, nullptr, occluded, 1));
    return MRT_OK;
}

mrt_status mrt_ray_order(mrt_ctx* ctx, const mrt_scene* s, const mrt_ray* rays, uint32_t n, uint32_t* order) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_ray_order: null context or scene");
    if (n && (!rays || !order)) return fail(ctx, MRT_ERR_INVALID, "mrt_ray_order: null ray or order buffer");
    if (n == 0) return MRT_OK;
    uint32_t* res = nullptr;
    HIP_TRY(ctx, launch_ray_sort(ctx, s->bounds, rays, n, nullptr, &res));
    HIP_TRY(ctx, hipMemcpyAsync(order, res, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    return MRT_OK;
}

mrt_status mrt_ray_order_noalloc(mrt_ctx* ctx, const mrt_scene* s, const mrt_ray* rays, uint32_t n, uint32_t* order) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_ray_order_noalloc: null context or scene");
    if (n && (!rays || !order)) return fail(ctx, MRT_ERR_INVALID, "mrt_ray_order_noalloc: null ray or order buffer");
    if (n > noalloc_max_rays())
        return fail(ctx, MRT_ERR_INVALID, "mrt_ray_order_noalloc: %u rays exceed the limit of %u of the allocation-free "
                    "order (use mrt_ray_order)", n, noalloc_max_rays());
    if (n == 0) return MRT_OK;
    HIP_TRY(ctx, launch_ray_sort_noalloc(ctx, s->bounds, rays, n, order));
    return MRT_OK;
}

mrt_status mrt_trace_closest_ordered(mrt_ctx* ctx, const mrt_scene* s, const mrt_ray* rays, const uint32_t* order,
                                     uint32_t n, mrt_hit* hits) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_trace_closest_ordered: null context or scene");
    if (n && (!rays || !hits || !order))
        return fail(ctx, MRT_ERR_INVALID, "mrt_trace_closest_ordered: null ray, order or hit buffer");
    HIP_TRY(ctx, launch_trace(ctx, s, rays, n, hits, nullptr, 0, order, nullptr));
    return MRT_OK;
}

mrt_status mrt_sort_pairs(mrt_ctx* ctx, uint32_t* keys, uint32_t* values, uint32_t n, uint32_t bits) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_sort_pairs: null context");
    if (n && (!keys || !values)) return fail(ctx, MRT_ERR_INVALID, "mrt_sort_pairs: null buffer");
    if (bits == 0 || bits > 32) return fail(ctx, MRT_ERR_INVALID, "mrt_sort_pairs: bits %u not in 1..32", bits);
    if (n == 0) return MRT_OK;
    HIP_TRY(ctx, launch_sort_pairs(ctx, keys, values, n, bits));
    return MRT_OK;
}

static mrt_status check_camera(mrt_ctx* ctx, const mrt_camera* cam, const char* fn) {
    if (!cam) return fail(ctx, MRT_ERR_INVALID, "%s: null camera", fn);
    if (cam->width == 0 || cam->height == 0 || (uint64_t)cam->width * cam->height > (1ull << 28))
        return fail(ctx, MRT_ERR_INVALID, "%s: viewport %ux%u out of range", fn, cam->width, cam->height);
    if (!(cam->vertical_fov > 0.0f && cam->vertical_fov < 3.14159f))
        return fail(ctx, MRT_ERR_INVALID, "%s: vertical_fov %g out of (0, pi)", fn, (double)cam->vertical_fov);
    return MRT_OK;
}

static mrt_status check_gbuffer(mrt_ctx* ctx, const mrt_camera* cam, const mrt_gbuffer* gb, const char* fn) {
    if (!gb || !gb->hits_dev) return fail(ctx, MRT_ERR_INVALID, "%s: G-buffer without hits", fn);
    if (gb->width != cam->width || gb->height != cam->height)
        return fail(ctx, MRT_ERR_INVALID, "%s: G-buffer %ux%u does not match camera %ux%u", fn, gb->width, gb->height,
                    cam->width, cam->height);
    return MRT_OK;
}

mrt_status mrt_primary_rays(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera* cam, const mrt_camera* prev,
                            mrt_gbuffer* gb) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_primary_rays: null context or scene");
    mrt_status st;
    if ((st = check_camera(ctx, cam, "mrt_primary_rays")) != MRT_OK) return st;
    if ((st = check_gbuffer(ctx, cam, gb, "mrt_primary_rays")) != MRT_OK) return st;
    HIP_TRY(ctx, run_primary(ctx, s, *cam, prev ? *prev : *cam, *gb));
    return MRT_OK;
}

static mrt_status check_render_opts(mrt_ctx* ctx, const mrt_camera* cam, const mrt_render_opts* o, const char* fn) {
    if (!o) return fail(ctx, MRT_ERR_INVALID, "%s: null options", fn);
    if (o->spp == 0 || o->spp > 4096) return fail(ctx, MRT_ERR_INVALID, "%s: spp %u not in 1..4096", fn, o->spp);
    if ((uint64_t)cam->width * cam->height * o->spp > (1ull << 28))
        return fail(ctx, MRT_ERR_INVALID, "%s: %llu paths exceed 2^28", fn,
                    (unsigned long long)cam->width * cam->height * o->spp);
    if (o->bounces > 60) return fail(ctx, MRT_ERR_INVALID, "%s: bounces %u > 60", fn, o->bounces);
    if (o->rr_start < 2) return fail(ctx, MRT_ERR_INVALID, "%s: rr_start must be >= 2", fn);
    if (o->flags & ~(MRT_RENDER_SORT_RAYS | MRT_RENDER_MEGAKERNEL))
        return fail(ctx, MRT_ERR_INVALID, "%s: unknown flags %#x", fn, o->flags);
    SkyDev sd;
    if (o->sky_lut && !sky_device(o->sky_lut, &sd))
        return fail(ctx, MRT_ERR_STATE, "%s: sky_lut has no view yet (mrt_sky_update_view)", fn);
    if (o->capture_bounce) {
        if (o->capture_bounce > o->bounces) return fail(ctx, MRT_ERR_INVALID, "%s: capture_bounce beyond bounces", fn);
        if (!o->capture_rays_dev || !o->capture_pid_dev || !o->capture_count_dev)
            return fail(ctx, MRT_ERR_INVALID, "%s: capture buffers missing", fn);
    }
    return MRT_OK;
}

mrt_status mrt_secondary_rays(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera* cam, const mrt_gbuffer* gb,
                              const mrt_render_opts* o, float* rgb) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_secondary_rays: null context or scene");
    mrt_status st;
    if ((st = check_camera(ctx, cam, "mrt_secondary_rays")) != MRT_OK) return st;
    if ((st = check_gbuffer(ctx, cam, gb, "mrt_secondary_rays")) != MRT_OK) return st;
    if ((st = check_render_opts(ctx, cam, o, "mrt_secondary_rays")) != MRT_OK) return st;
    if (!rgb && !o->capture_bounce) return fail(ctx, MRT_ERR_INVALID, "mrt_secondary_rays: null rgb");
    HIP_TRY(ctx, run_secondary(ctx, s, *cam, *gb, *o, rgb));
    return MRT_OK;
}

mrt_status mrt_render(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera* cam, const mrt_render_opts* o, float* rgb) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_render: null context or scene");
    mrt_status st;
    if ((st = check_camera(ctx, cam, "mrt_render")) != MRT_OK) return st;
    if ((st = check_render_opts(ctx, cam, o, "mrt_render")) != MRT_OK) return st;
    if (!rgb && !o->capture_bounce) return fail(ctx, MRT_ERR_INVALID, "mrt_render: null rgb");
    HIP_TRY(ctx, ctx->ws.ensure_pixels((size_t)cam->width * cam->height));
    mrt_gbuffer gb{ctx->ws.gb_hits, nullptr, nullptr, nullptr, cam->width, cam->height};
    HIP_TRY(ctx, run_primary(ctx, s, *cam, *cam, gb));
    HIP_TRY(ctx, run_secondary(ctx, s, *cam, gb, *o, rgb));
    return MRT_OK;
}

mrt_status mrt_render_cached(mrt_ctx* ctx, const mrt_scene* s, mrt_cache* cache, const mrt_camera* cam,
                             const mrt_render_opts* o, const mrt_cache_render_opts* co, float* rgb) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_render_cached: null context or scene");
    if (!co || !co->enable || !cache) return mrt_render(ctx, s, cam, o, rgb);
    mrt_status st;
    if ((st = check_camera(ctx, cam, "mrt_render_cached")) != MRT_OK) return st;
    if ((st = check_render_opts(ctx, cam, o, "mrt_render_cached")) != MRT_OK) return st;
    if (!rgb) return fail(ctx, MRT_ERR_INVALID, "mrt_render_cached: null rgb");
    if (o->capture_bounce) return fail(ctx, MRT_ERR_UNSUPPORTED, "mrt_render_cached: no ray capture with a cache");
    if (co->vertex >= o->bounces)
        return fail(ctx, MRT_ERR_INVALID, "mrt_render_cached: vertex %u must be < bounces %u", co->vertex, o->bounces);
    if (co->train_samples && (co->train_steps == 0 || co->train_steps > 32))
        return fail(ctx, MRT_ERR_INVALID, "mrt_render_cached: train_steps %u not in 1..32", co->train_steps);
    if (co->cache_rgb_dev && o->spp != 1)
        return fail(ctx, MRT_ERR_INVALID, "mrt_render_cached: cache_rgb export needs spp 1");
    HIP_TRY(ctx, ctx->ws.ensure_pixels((size_t)cam->width * cam->height));
    mrt_gbuffer gb{ctx->ws.gb_hits, nullptr, nullptr, nullptr, cam->width, cam->height};
    HIP_TRY(ctx, run_primary(ctx, s, *cam, *cam, gb));
    HIP_TRY(ctx, run_cached(ctx, s, cache, *cam, gb, *o, *co, rgb));
    return MRT_OK;
}

mrt_status mrt_accumulate(mrt_ctx* ctx, float* sum, const float* rgb, size_t n, uint32_t frames, float* mean) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_accumulate: null context");
    if (n && (!sum || !rgb)) return fail(ctx, MRT_ERR_INVALID, "mrt_accumulate: null buffer");
    if (mean && frames == 0) return fail(ctx, MRT_ERR_INVALID, "mrt_accumulate: frames must be >= 1");
    HIP_TRY(ctx, launch_accumulate(ctx, sum, rgb, n, frames, mean));
    return MRT_OK;
}

mrt_status mrt_rel_mse(mrt_ctx* ctx, const float* img, const float* ref, size_t n, double eps, double* out) {
    if (!ctx || !out) return fail(ctx, MRT_ERR_INVALID, "mrt_rel_mse: null argument");
    if (n && (!img || !ref)) return fail(ctx, MRT_ERR_INVALID, "mrt_rel_mse: null buffer");
    if (!(eps > 0.0)) return fail(ctx, MRT_ERR_INVALID, "mrt_rel_mse: eps must be > 0");
    HIP_TRY(ctx, launch_rel_mse(ctx, img, ref, n, eps, out));
    return MRT_OK;
}

mrt_status mrt_shade_lambert(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera* cam, const mrt_gbuffer* gb,
                             const float light_dir[3], const float light_e[3], float* rgb) {
    if (!ctx || !s) return fail(ctx, MRT_ERR_INVALID, "mrt_shade_lambert: null context or scene");
    mrt_status st;
    if ((st = check_camera(ctx, cam, "mrt_shade_lambert")) != MRT_OK) return st;
    if ((st = check_gbuffer(ctx, cam, gb, "mrt_shade_lambert")) != MRT_OK) return st;
    if (!light_dir || !light_e || !rgb) return fail(ctx, MRT_ERR_INVALID, "mrt_shade_lambert: null argument");
    HIP_TRY(ctx, run_lambert(ctx, s, *cam, *gb, light_dir, light_e, rgb));
    return MRT_OK;
}

uint32_t mrt_trace_variant_count(void) { return (uint32_t)trace_variant_count(); }

const char* mrt_trace_variant_name(uint32_t v) { return trace_variant_name((int)v); }

mrt_status mrt_trace_variant_set(mrt_ctx* ctx, int any, uint32_t v) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_trace_variant_set: null context");
    if ((int)v >= trace_variant_count()) return fail(ctx, MRT_ERR_INVALID, "mrt_trace_variant_set: variant %u of %d", v,
                                                     trace_variant_count());
    (any ? ctx->trace_variant_any : ctx->trace_variant_closest) = (int)v;
    return MRT_OK;
}

mrt_status mrt_trace_variant_get(mrt_ctx* ctx, int any, uint32_t* v) {
    if (!ctx || !v) return fail(ctx, MRT_ERR_INVALID, "mrt_trace_variant_get: null argument");
    *v = (uint32_t)(any ? ctx->trace_variant_any : ctx->trace_variant_closest);
    return MRT_OK;
}

mrt_status mrt_profile_enable(mrt_ctx* ctx, int on) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_profile_enable: null context");
    ctx->profiling = on != 0;
    return MRT_OK;
}

mrt_status mrt_profile_reset(mrt_ctx* ctx) {
    if (!ctx) return fail(nullptr, MRT_ERR_INVALID, "mrt_profile_reset: null context");
    mrt_status st = collect_profile(ctx);
    ctx->prof.clear();
    for (auto& r : ctx->region_open) ctx->event_pool.push_back(r.second);
    ctx->region_open.clear();
    return st;
}

mrt_status mrt_profile_region(mrt_ctx* ctx, const char* name, int begin) {
    if (!ctx || !name) return fail(ctx, MRT_ERR_INVALID, "mrt_profile_region: null argument");
    auto open = ctx->region_open.find(name);
    if (begin ? open != ctx->region_open.end() : open == ctx->region_open.end())
        return fail(ctx, MRT_ERR_STATE, "mrt_profile_region: region \"%s\" is %s", name,
                    begin ? "already open" : "not open");
    hipEvent_t ev = nullptr;
    if (!ctx->event_pool.empty()) {
        ev = ctx->event_pool.back();
        ctx->event_pool.pop_back();
    } else {
        HIP_TRY(ctx, hipEventCreate(&ev));
    }
    const hipError_t e = hipEventRecord(ev, ctx->stream);
    if (e != hipSuccess) {
        ctx->event_pool.push_back(ev);
        HIP_TRY(ctx, e);
    }
    if (begin) {
        ctx->region_open[name] = ev;
    } else {
        ctx->pending.push_back(ProfEvent{name, open->second, ev});
        ctx->region_open.erase(open);
    }
    return MRT_OK;
}

mrt_status mrt_profile_query(mrt_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches) {
    if (!ctx || !kernel) return fail(ctx, MRT_ERR_INVALID, "mrt_profile_query: null argument");
    mrt_status st = collect_profile(ctx);
    if (st != MRT_OK) return st;
    auto it = ctx->prof.find(kernel);
    if (total_ms) *total_ms = it == ctx->prof.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == ctx->prof.end() ? 0 : it->second.launches;
    return MRT_OK;
}

}  // extern "C"
